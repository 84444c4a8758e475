// Persistent stepping kernels, f32 engine precision, advect nonlinearity, n = 2^4 .. 2^13.
#include "rks_kernels.h"

template __global__ void rks_kernel<float, 4, NLK_ADVECT>(rks_launch);
template __global__ void rks_kernel<float, 5, NLK_ADVECT>(rks_launch);
template __global__ void rks_kernel<float, 6, NLK_ADVECT>(rks_launch);
template __global__ void rks_kernel<float, 7, NLK_ADVECT>(rks_launch);
template __global__ void rks_kernel<float, 8, NLK_ADVECT>(rks_launch);
template __global__ void rks_kernel<float, 9, NLK_ADVECT>(rks_launch);
template __global__ void rks_kernel<float, 10, NLK_ADVECT>(rks_launch);
template __global__ void rks_kernel<float, 11, NLK_ADVECT>(rks_launch);
template __global__ void rks_kernel<float, 12, NLK_ADVECT>(rks_launch);
template __global__ void rks_kernel<float, 13, NLK_ADVECT>(rks_launch);

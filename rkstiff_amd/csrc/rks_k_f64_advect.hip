// Persistent stepping kernels, f64 engine precision, advect nonlinearity, n = 2^4 .. 2^13.
#include "rks_kernels.h"

template __global__ void rks_kernel<double, 4, NLK_ADVECT>(rks_launch);
template __global__ void rks_kernel<double, 5, NLK_ADVECT>(rks_launch);
template __global__ void rks_kernel<double, 6, NLK_ADVECT>(rks_launch);
template __global__ void rks_kernel<double, 7, NLK_ADVECT>(rks_launch);
template __global__ void rks_kernel<double, 8, NLK_ADVECT>(rks_launch);
template __global__ void rks_kernel<double, 9, NLK_ADVECT>(rks_launch);
template __global__ void rks_kernel<double, 10, NLK_ADVECT>(rks_launch);
template __global__ void rks_kernel<double, 11, NLK_ADVECT>(rks_launch);
template __global__ void rks_kernel<double, 12, NLK_ADVECT>(rks_launch);
template __global__ void rks_kernel<double, 13, NLK_ADVECT>(rks_launch);

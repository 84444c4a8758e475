// Persistent stepping kernels: state/FFT float, NL and coefficient storage float, advect nonlinearity, n = 2^4 .. 2^13.
#include "rks_kernels.h"

template __global__ void rks_kernel<float, float, 4, NLK_ADVECT>(rks_launch);
template __global__ void rks_kernel<float, float, 5, NLK_ADVECT>(rks_launch);
template __global__ void rks_kernel<float, float, 6, NLK_ADVECT>(rks_launch);
template __global__ void rks_kernel<float, float, 7, NLK_ADVECT>(rks_launch);
template __global__ void rks_kernel<float, float, 8, NLK_ADVECT>(rks_launch);
template __global__ void rks_kernel<float, float, 9, NLK_ADVECT>(rks_launch);
template __global__ void rks_kernel<float, float, 10, NLK_ADVECT>(rks_launch);
template __global__ void rks_kernel<float, float, 11, NLK_ADVECT>(rks_launch);
template __global__ void rks_kernel<float, float, 12, NLK_ADVECT>(rks_launch);
template __global__ void rks_kernel<float, float, 13, NLK_ADVECT>(rks_launch);

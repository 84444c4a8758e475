// Persistent stepping kernels: state/FFT double, NL and coefficient storage double, advect nonlinearity, n = 2^4 .. 2^13.
#include "rks_kernels.h"

template __global__ void rks_kernel<double, double, 4, NLK_ADVECT>(rks_launch);
template __global__ void rks_kernel<double, double, 5, NLK_ADVECT>(rks_launch);
template __global__ void rks_kernel<double, double, 6, NLK_ADVECT>(rks_launch);
template __global__ void rks_kernel<double, double, 7, NLK_ADVECT>(rks_launch);
template __global__ void rks_kernel<double, double, 8, NLK_ADVECT>(rks_launch);
template __global__ void rks_kernel<double, double, 9, NLK_ADVECT>(rks_launch);
template __global__ void rks_kernel<double, double, 10, NLK_ADVECT>(rks_launch);
template __global__ void rks_kernel<double, double, 11, NLK_ADVECT>(rks_launch);
template __global__ void rks_kernel<double, double, 12, NLK_ADVECT>(rks_launch);
template __global__ void rks_kernel<double, double, 13, NLK_ADVECT>(rks_launch);

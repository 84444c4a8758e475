// Persistent stepping kernels: state/FFT double, NL and coefficient storage float, cubic nonlinearity, n = 2^4 .. 2^13.
#include "rks_kernels.h"

template __global__ void rks_kernel<double, float, 4, NLK_CUBIC>(rks_launch);
template __global__ void rks_kernel<double, float, 5, NLK_CUBIC>(rks_launch);
template __global__ void rks_kernel<double, float, 6, NLK_CUBIC>(rks_launch);
template __global__ void rks_kernel<double, float, 7, NLK_CUBIC>(rks_launch);
template __global__ void rks_kernel<double, float, 8, NLK_CUBIC>(rks_launch);
template __global__ void rks_kernel<double, float, 9, NLK_CUBIC>(rks_launch);
template __global__ void rks_kernel<double, float, 10, NLK_CUBIC>(rks_launch);
template __global__ void rks_kernel<double, float, 11, NLK_CUBIC>(rks_launch);
template __global__ void rks_kernel<double, float, 12, NLK_CUBIC>(rks_launch);
template __global__ void rks_kernel<double, float, 13, NLK_CUBIC>(rks_launch);

// Persistent stepping kernels, f64 engine precision, cubic nonlinearity, n = 2^4 .. 2^13.
#include "rks_kernels.h"

template __global__ void rks_kernel<double, 4, NLK_CUBIC>(rks_launch);
template __global__ void rks_kernel<double, 5, NLK_CUBIC>(rks_launch);
template __global__ void rks_kernel<double, 6, NLK_CUBIC>(rks_launch);
template __global__ void rks_kernel<double, 7, NLK_CUBIC>(rks_launch);
template __global__ void rks_kernel<double, 8, NLK_CUBIC>(rks_launch);
template __global__ void rks_kernel<double, 9, NLK_CUBIC>(rks_launch);
template __global__ void rks_kernel<double, 10, NLK_CUBIC>(rks_launch);
template __global__ void rks_kernel<double, 11, NLK_CUBIC>(rks_launch);
template __global__ void rks_kernel<double, 12, NLK_CUBIC>(rks_launch);
template __global__ void rks_kernel<double, 13, NLK_CUBIC>(rks_launch);

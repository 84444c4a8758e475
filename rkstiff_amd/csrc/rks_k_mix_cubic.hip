// Persistent stepping kernels: state/FFT double, NL and coefficient storage float, cubic nonlinearity, n = 2^4 .. 2^13
// (and the proposer variants, n = 2^7 .. 2^13).
#include "rks_kernels.h"

RKS_KERNEL_SIZES(RKS_INSTANTIATE, double, float, NLK_CUBIC)

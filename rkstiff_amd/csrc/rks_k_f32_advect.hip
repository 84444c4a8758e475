// Persistent stepping kernels: state/FFT float, NL and coefficient storage float, advect nonlinearity, n = 2^4 .. 2^13
// (and the proposer variants, n = 2^7 .. 2^13).
#include "rks_kernels.h"

RKS_KERNEL_SIZES(RKS_INSTANTIATE, float, float, NLK_ADVECT)

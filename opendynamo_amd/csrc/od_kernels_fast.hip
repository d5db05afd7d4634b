// Fast-math build of the step kernels (the [solver] fast_math key, f32
// compute): the kernel source compiled with OD_FAST = 1 and contraction on
// (-ffp-contract=fast), into namespace od::fastmath.
#define OD_FAST 1
#include "od_kernels.hip"

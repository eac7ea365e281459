// bitset-store kernels (filled in later)
#include <hip/hip_runtime.h>

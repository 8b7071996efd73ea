// FP32 instantiation: the served search.
#define SGPU_T float
#include "kernels.inc"

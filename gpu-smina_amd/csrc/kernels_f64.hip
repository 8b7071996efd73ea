// FP64 instantiation: the oracle-parity search.  Built with
// -ffp-contract=off so that no multiply-add is fused (the reference builds its
// FP64 kernels with --fmad=false, src/gpu/CMakeLists.txt:31-34).
#define SGPU_T double
#include "kernels.inc"

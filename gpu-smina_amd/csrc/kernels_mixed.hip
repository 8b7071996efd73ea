// Mixed instantiation: FP32 evaluation (maps, pair tables, forces) with the
// search state in FP64 (conf accumulation, BFGS Hessian update, line search,
// Metropolis energies).
#define SGPU_T float
#define SGPU_S double
#include "kernels.inc"

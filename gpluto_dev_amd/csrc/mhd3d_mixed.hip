// Mixed-precision instantiation: FP32 primitives, PLM, Riemann solver, UCT
// data and EMFs; FP64 conserved variables, RK accumulation, face field, CT
// update and dt.  Built with FMA contraction and the fast FP32 divide/sqrt.
#include "mhd3d_kernels.hpp"

namespace gpl {
template void mhd3d_launch<float>(int, const Mhd3dBufs<float> &, const Box3 &, const StepCtl *,
                                  RKW, int, int, int, int, int, double, double, double,
                                  hipStream_t);
}

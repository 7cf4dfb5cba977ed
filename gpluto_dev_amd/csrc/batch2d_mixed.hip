// FP32 instantiation of the batched 2D MHD kernels: FP32 primitives, PLM,
// Riemann solver, UCT data, EMFs and within-stage sums; FP64 conserved
// state, RK combination, face field, CT update and dt.  Built with FMA
// contraction and the fast FP32 divide/sqrt.
#include "batch2d_kernels.hpp"

namespace gpl {
template void batch2d_launch<float>(int, const Batch2dBufs<float> &, RKW, int, int, int, int,
                                    int, int, double, double, double, hipStream_t);
}

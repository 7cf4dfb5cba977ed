// FP64 drop-in instantiation of the batched 2D MHD kernels (built with
// -ffp-contract=off: every product and sum rounded as in the reference).
#include "batch2d_kernels.hpp"

namespace gpl {
template void batch2d_launch<double>(int, const Batch2dBufs<double> &, RKW, int, int, int, int,
                                     int, int, double, double, double, hipStream_t);
}

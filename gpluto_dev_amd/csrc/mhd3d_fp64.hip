// FP64 drop-in instantiation of the 3D MHD stage kernels.  Built with
// -ffp-contract=off so every product and sum is rounded as in the reference.
#include "mhd3d_kernels.hpp"

namespace gpl {
template void mhd3d_launch<double>(int, const Mhd3dBufs<double> &, const Box3 &, const StepCtl *,
                                   RKW, int, int, int, int, int, double, double, double,
                                   hipStream_t);
}

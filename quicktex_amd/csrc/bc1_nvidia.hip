// bc1_nvidia.hip -- the BC1/BC3 encode kernels of the Nvidia interpolator.
#include "bc1_kernels.h"

namespace qtx {
template hipError_t launch_bc1_interp<QTX_INTERP_NVIDIA>(const Bc1Launch &);
template hipError_t launch_bc1_batch_interp<QTX_INTERP_NVIDIA>(bool, const BatchJob *, int, uint32_t, uint32_t *, unsigned,
                                                 const DeviceTables *, hipStream_t);
}  // namespace qtx

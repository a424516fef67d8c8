// Volume right-hand side kernels, FP64 storage and arithmetic (parity path),
// 256 grid lines per workgroup.  The kernel itself is in volume_impl.hpp; this
// unit instantiates it for orders 1-4 and every fused epilogue.
#include "volume_impl.hpp"

namespace drf {

hipError_t launch_volume_fp64_w256(int order, const VolGeom<double> &geo, const Ops &ops,
                                   const Phys &ph, View<const double> U, const Epi<double> &epi,
                                   hipStream_t s)
{
    return vol::dispatch<double, double, 256>(order, geo, ops, ph, U, epi, s);
}

} // namespace drf

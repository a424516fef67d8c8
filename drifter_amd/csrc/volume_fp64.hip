// Volume right-hand side kernels, FP64 storage and arithmetic (parity path),
// 128 grid lines per workgroup.  The kernel itself is in volume_impl.hpp; this
// unit instantiates it for orders 1-8 and every fused epilogue.
#include "volume_impl.hpp"

namespace drf {

hipError_t launch_volume_fp64_w128(int order, const VolGeom<double> &geo, const Ops &ops,
                                   const Phys &ph, View<const double> U, const Epi<double> &epi,
                                   hipStream_t s)
{
    return vol::dispatch<double, double, 128>(order, geo, ops, ph, U, epi, s);
}

} // namespace drf

// Volume right-hand side kernels, FP32 storage and arithmetic (mixed-precision
// path; the free-surface column integrals stay FP64),
// 512 grid lines per workgroup.  The kernel itself is in volume_impl.hpp; this
// unit instantiates it for orders 1-4 and every fused epilogue.
#include "volume_impl.hpp"

namespace drf {

hipError_t launch_volume_fp32_w512(int order, const VolGeom<float> &geo, const Ops &ops,
                                   const Phys &ph, View<const float> U, const Epi<float> &epi,
                                   hipStream_t s)
{
    return vol::dispatch<float, float, 512>(order, geo, ops, ph, U, epi, s);
}

} // namespace drf

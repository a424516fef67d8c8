// The CFL wave-speed maximum with no device allocation: the partial maxima
// of the reduction live in a file-scope device array instead of the
// context's lazily allocated reduction buffer, and the maximum is written to
// a device double the caller owns.  For callers that cannot allocate (code
// under HIP graph capture, serving loops that forbid hipMalloc).
//
// `scratch` is one array per GPU, held until the process ends; every call on
// a GPU shares it, so calls that overlap on one GPU race.  Each call writes
// every partial it reads (one per workgroup of the wave-speed kernel), so
// nothing has to start at zero.
#include <mutex>
#include <vector>

#include "drf_internal.hpp"

namespace drf {

// at most this many partial maxima (the wave-speed launch never uses more)
constexpr int64_t kScratchPartials = 2048;

} // namespace drf

static __device__ __align__(256) double scratch[drf::kScratchPartials];

namespace drf {
namespace {

std::mutex g_scratch_mu;
std::vector<double *> g_scratch_ptr; // per device: the device address of `scratch`

} // namespace

// Looks up the device address of `scratch` on `device` (which must be the
// current device) once, at context creation.
drf_status register_scratch(int device)
{
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if ((int)g_scratch_ptr.size() <= device)
        g_scratch_ptr.resize(device + 1, nullptr);
    if (g_scratch_ptr[device])
        return DRF_OK;
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
        return fail(DRF_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") +
                                     hipGetErrorString(e));
    g_scratch_ptr[device] = (double *)p;
    return DRF_OK;
}

double *scratch_ptr(int device)
{
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    return device < (int)g_scratch_ptr.size() ? g_scratch_ptr[device] : nullptr;
}

} // namespace drf

using namespace drf;

extern "C" drf_status drf_max_wave_speed_noalloc_dev(drf_ctx *c, const void *U, int layout,
                                                     void *c_max, void *stream)
{
    if (!c || c->E <= 0)
        return fail(DRF_ERR_STATE, "drf_max_wave_speed_noalloc_dev: context not initialized");
    if (!U || !c_max)
        return fail(DRF_ERR_ARG, "drf_max_wave_speed_noalloc_dev: null argument");
    if (layout != DRF_LAYOUT_PACKED && layout != DRF_LAYOUT_SOA && layout != DRF_LAYOUT_LINES &&
        layout != DRF_LAYOUT_BLOCKS)
        return fail(DRF_ERR_ARG, "drf_max_wave_speed_noalloc_dev: unknown layout");
    if (c->precision != DRF_PREC_FP64)
        return fail(DRF_ERR_UNSUPPORTED,
                    "drf_max_wave_speed_noalloc_dev: wave speed on the FP64 state only");
    // the same partial count as drf_max_wave_speed_dev, so the two agree bit for bit
    const int64_t npart = std::min<int64_t>(2048, (c->E + 3) / 4);
    if (npart > kScratchPartials)
        return fail(DRF_ERR_ARG, "drf_max_wave_speed_noalloc_dev: " + std::to_string(npart) +
                                     " partial maxima exceed the scratch array's " +
                                     std::to_string(kScratchPartials));
    double *part = scratch_ptr(c->device);
    if (!part)
        return fail(DRF_ERR_STATE,
                    "drf_max_wave_speed_noalloc_dev: scratch not registered on this device");
    hipError_t e = hipSetDevice(c->device);
    View<const double> Uv{(const double *)U, 0, 0, 0, 0};
    layout_strides(layout, c->E, c->n, 5, Uv.fs, Uv.es, Uv.ls, Uv.is, Uv.bs, Uv.bb);
    View<const double> hv{c->d_h, 0, 0, 0, 0}; // bathymetry: LINES
    layout_strides(DRF_LAYOUT_LINES, c->E, c->n, 5, hv.fs, hv.es, hv.ls, hv.is, hv.bs, hv.bb);
    const hipStream_t s = (hipStream_t)stream; // as given; NULL is HIP's null stream
    if (e == hipSuccess)
        e = launch_wave_speed(c->order, c->E, Uv, hv, part, npart, s);
    if (e == hipSuccess)
        e = launch_max_reduce(part, npart, (double *)c_max, s);
    if (e != hipSuccess)
        return fail(DRF_ERR_HIP, std::string("drf_max_wave_speed_noalloc_dev: ") +
                                     hipGetErrorString(e));
    return DRF_OK;
}

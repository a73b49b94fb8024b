// The entry point that allocates nothing: get_difference on arrays the caller
// already holds in device memory, for callers that cannot allocate (code
// under HIP graph capture, serving loops that keep every buffer resident,
// harnesses that forbid hipMalloc).
//
// Its only device memory beyond the caller's inputs and output is `scratch`,
// the per-block partial sums of the reduction, a file-scope device array of
// REDUCE_BLOCKS doubles (8 KiB) that every GPU the library runs on holds
// until the process ends.  The reduction always runs REDUCE_BLOCKS blocks, so
// the scratch it needs does not grow with the arrays.  Each block writes its
// own partial before the finishing kernel reads it: nothing has to start at
// zero, and the array keeps the last call's partials.  Calls that overlap on
// one GPU (other streams, other threads) race on the array.
#include <hip/hip_runtime.h>

#include <mutex>

#include "host.hpp"
#include "tpls_kernels.hpp"

namespace {

using namespace tpls;

constexpr int NOALLOC_PARTS = REDUCE_BLOCKS;
constexpr int NOALLOC_MAX_DEVICES = 64;

static __device__ __align__(256) double scratch[NOALLOC_PARTS];

std::mutex g_scratch_mu;
double* g_scratch[NOALLOC_MAX_DEVICES] = {};

// The device address of `scratch` on the current device, looked up once per
// device with that device current.  The host-side name `scratch` is not a
// device address and is never passed to a kernel or a copy.
int scratch_of_current_device(double** out)
{
    int dev = -1;
    if (hipGetDevice(&dev) != hipSuccess) return set_error(TPLS_ERR_NOGPU, "no current HIP device");
    if (dev < 0 || dev >= NOALLOC_MAX_DEVICES)
        return set_error(TPLS_ERR_ARG, "device index " + std::to_string(dev) + " beyond the " +
                                           std::to_string(NOALLOC_MAX_DEVICES) + " devices the scratch table holds");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch[dev]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return set_error(TPLS_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        g_scratch[dev] = static_cast<double*>(p);
    }
    *out = g_scratch[dev];
    return TPLS_OK;
}

bool on_device(const void* p)
{
    hipPointerAttribute_t at;
    if (hipPointerGetAttributes(&at, p) != hipSuccess) {
        (void)hipGetLastError();  // an unknown host pointer leaves an error behind
        return false;
    }
    return at.type == hipMemoryTypeDevice || at.type == hipMemoryTypeManaged;
}

}  // namespace

extern "C" int tpls_get_difference_noalloc(const double* ua, const double* ub, int ex, int ey, int ez, int sx, int sy,
                                           int sz, double* diff, void* stream)
{
    static_assert(REDUCE_BLOCKS <= NOALLOC_PARTS, "the reduction's partials must fit the scratch");
    if (!ua || !ub || !diff) return set_error(TPLS_ERR_ARG, "null array");
    if (sx != 1 || sy != 1 || sz != 0) return set_error(TPLS_ERR_DECOMP, "single-rank decomposition required");
    if (ex < 1 || ey < 2 || ez < 0) return set_error(TPLS_ERR_ARG, "empty array bounds");
    if (!on_device(ua) || !on_device(ub) || !on_device(diff))
        return set_error(TPLS_ERR_ARG, "ua, ub and diff must be device memory of the current device");
    double* part = nullptr;
    const int rc = scratch_of_current_device(&part);
    if (rc != TPLS_OK) return rc;
    // the arrays as the caller holds them: (sx-1:ex+1, sy-1:ey+1, sz:ez), unpadded
    const long nx = ex + 2, ny = ey + 2;
    const Grid g{ex + 1, ey + 1, ez + 2, ey};
    FieldView<double> a{const_cast<double*>(ua), nx, nx * ny}, b{const_cast<double*>(ub), nx, nx * ny};
    const hipError_t e = launch_difference<double>(static_cast<hipStream_t>(stream), g, ez + 1, a, b,
                                                   Reduce{part, diff});
    if (e != hipSuccess) return set_error(TPLS_ERR_HIP, std::string("get_difference launch: ") + hipGetErrorString(e));
    return TPLS_OK;
}

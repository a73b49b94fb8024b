// The halo exchanges of mpi.f90 for one rank with y-periodic neighbours
// (exchange2d, exchange2d_aug1, exchange2d_aug2), and the device side of the
// host-array mirror (mirror.hpp): its allocation and copy backend, and the
// exchange itself run on the mirror's device copy when the device holds the
// newest data.
#include <hip/hip_runtime.h>

#include <cstring>
#include <string>

#include "host.hpp"
#include "mirror.hpp"
#include "tpls_kernels.hpp"

namespace tpls {

namespace {

constexpr int ROWS_BX = 256;

__global__ void __launch_bounds__(ROWS_BX) k_copy_rows(double* a, long nx, long nxy, int i0, int ni, RowPairs rp)
{
    const int i = blockIdx.x * ROWS_BX + threadIdx.x;
    if (i >= ni) return;
    const int q = blockIdx.z;
    double* p = a + long(blockIdx.y) * nxy + i0 + i;
    p[long(rp.dst[q]) * nx] = p[long(rp.src[q]) * nx];
}

}  // namespace

hipError_t launch_copy_rows(hipStream_t s, double* a, long nx, long ny, long nz, int i0, int ni, const RowPairs& rp)
{
    if (!a || rp.n < 1 || rp.n > MAX_ROW_PAIRS || i0 < 0 || ni < 1 || i0 + ni > nx || nz < 1 || nz > 65535)
        return hipErrorInvalidValue;
    for (int q = 0; q < rp.n; ++q)
        if (rp.dst[q] < 0 || rp.dst[q] >= ny || rp.src[q] < 0 || rp.src[q] >= ny) return hipErrorInvalidValue;
    const dim3 grid((ni + ROWS_BX - 1) / ROWS_BX, unsigned(nz), unsigned(rp.n));
    hipLaunchKernelGGL(k_copy_rows, grid, dim3(ROWS_BX), 0, s, a, nx, nx * ny, i0, ni, rp);
    return hipGetLastError();
}

namespace {

// ---- the mirror's device backend: plain synchronous runtime copies ----
void* dev_alloc(size_t n)
{
    void* p = nullptr;
    return hipMalloc(&p, n) == hipSuccess ? p : nullptr;
}
void dev_release(void* p) { (void)hipFree(p); }
int dev_upload(void* d, const void* h, size_t n) { return hipMemcpy(d, h, n, hipMemcpyHostToDevice) == hipSuccess ? 0 : -1; }
int dev_download(void* h, const void* d, size_t n)
{
    return hipMemcpy(h, d, n, hipMemcpyDeviceToHost) == hipSuccess ? 0 : -1;
}
int dev_fail(int code, const char* msg) { return set_error(code, msg); }

struct InstallBackend {
    InstallBackend() { mirror::set_backend({dev_alloc, dev_release, dev_upload, dev_download, dev_fail}); }
} g_install_backend;

}  // namespace
}  // namespace tpls

using namespace tpls;

extern "C" int tpls_exchange2d(double* uu, int ex, int ey, int ez, int sx, int sy, int sz, int kind)
{
    if (!uu) return set_error(TPLS_ERR_ARG, "null array");
    if (sx != 1 || sy != 1) return set_error(TPLS_ERR_DECOMP, "single-rank decomposition required: sx = sy = 1");
    if (ex < 1 || ey < 2 || ez < sz || kind < 0 || kind > 2)
        return set_error(TPLS_ERR_ARG, "exchange2d: bad bounds or kind");
    // (sx-1:ex+1, sy-1:ey+1, sz:ez) for exchange2d, (sx-2:ex+2, sy-2:ey+2, sz:ez)
    // for the augmented arrays; rows and columns 0-based from the lower bound
    const int halo = kind == 0 ? 1 : 2;
    const long nx = ex - sx + 1 + 2 * halo, ny = ey - sy + 1 + 2 * halo, nz = ez - sz + 1;
    auto row = [&](int j) { return j - (sy - halo); };
    RowPairs rp{};
    rp.n = 2;
    if (kind == 2) {  // second halo rows: sy-2 <- ey-1, ey+2 <- sy+1 (mpi.f90:294-311)
        rp.dst[0] = row(sy - 2), rp.src[0] = row(ey - 1);
        rp.dst[1] = row(ey + 2), rp.src[1] = row(sy + 1);
    } else {          // first halo rows: sy-1 <- ey, ey+1 <- sy (mpi.f90:124-137, 225-238)
        rp.dst[0] = row(sy - 1), rp.src[0] = row(ey);
        rp.dst[1] = row(ey + 1), rp.src[1] = row(sy);
    }
    // x from sx-1 to ex+1 in every kind (the vector types' block length ex-sx+3)
    const int i0 = halo - 1, ni = ex - sx + 3;
    const size_t bytes = size_t(nx * ny * nz) * sizeof(double);
    if (mirror::any() && mirror::device_of(uu, bytes) && !mirror::host_current(uu, bytes)) {
        // the device holds data the host has not seen: exchange there
        double* d = static_cast<double*>(mirror::device_of(uu, bytes));
        if (int rc = mirror::to_device(uu, bytes)) return rc;
        hipError_t e = launch_copy_rows(nullptr, d, nx, ny, nz, i0, ni, rp);
        if (e == hipSuccess) e = hipStreamSynchronize(nullptr);
        if (e != hipSuccess) return set_error(TPLS_ERR_HIP, std::string("exchange2d on the device: ") + hipGetErrorString(e));
        return mirror::written(uu, bytes);
    }
    for (long k = 0; k < nz; ++k)
        for (int q = 0; q < rp.n; ++q)
            std::memcpy(uu + (k * ny + rp.dst[q]) * nx + i0, uu + (k * ny + rp.src[q]) * nx + i0,
                        size_t(ni) * sizeof(double));
    return TPLS_OK;
}

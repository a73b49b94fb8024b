// Host side of the MI355X-native TPLS stencil core: the device-resident case
// (the timestep of main_ns_hybrid.F90:583-918, 2480-2509 as a sequence of
// kernel launches on one HIP stream) and the C ABI of include/tpls_core.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <memory>
#include <atomic>
#include <cerrno>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/tpls_core.h"
#include "../../include/tpls_io.h"
#include "group.hpp"
#include "host.hpp"
#include "mirror.hpp"
#include "tpls_kernels.hpp"

using namespace tpls;

namespace {

thread_local std::string g_err;

int fail(int code, const std::string& msg)
{
    g_err = msg;
    return code;
}

#define HIPCHK(x)                                                                                  \
    do {                                                                                           \
        hipError_t e_ = (x);                                                                       \
        if (e_ != hipSuccess) return fail(TPLS_ERR_HIP, std::string(#x) + ": " + hipGetErrorString(e_)); \
    } while (0)

// Fortran bounds of a field (main_ns_hybrid.F90:416-549).
struct Shape {
    int lx, ux, ly, uy, lz, uz;
    long nx() const { return ux - lx + 1; }
    long ny() const { return uy - ly + 1; }
    long nz() const { return uz - lz + 1; }
    long count() const { return nx() * ny() * nz(); }
};

enum Grid3 { G_UV, G_W, G_P, G_AUG, G_INLET };

int grid_of(int f)
{
    switch (f) {
    case TPLS_F_W2: case TPLS_F_W3: case TPLS_F_CONV0_W: case TPLS_F_CONV1_W: case TPLS_F_CONV2_W:
    case TPLS_F_CSF_W0: case TPLS_F_CSF_W1: case TPLS_F_CSF_W2: case TPLS_F_DIFF_W: case TPLS_F_RHS_W:
        return G_W;
    case TPLS_F_PHI2: case TPLS_F_PHI3: case TPLS_F_PRES: case TPLS_F_PRES_OLD: case TPLS_F_CONV0_PHI:
    case TPLS_F_CONV1_PHI: case TPLS_F_CONV2_PHI: case TPLS_F_FX_CSF: case TPLS_F_FY_CSF: case TPLS_F_FZ_CSF:
    case TPLS_F_VISC2: case TPLS_F_RHS_P: case TPLS_F_CONVPREV_PHI:
        return G_P;
    case TPLS_F_VISC2_AUG: return G_AUG;
    case TPLS_F_U_INLET: return G_INLET;
    default: return G_UV;
    }
}

Shape shape_of(int grid, int maxl, int maxm, int maxn)
{
    switch (grid) {
    case G_UV: return {0, maxl, 0, maxm, 0, maxn - 2};
    case G_W: return {0, maxl, 0, maxm, 0, maxn - 1};
    case G_P: return {0, maxl, 0, maxm, 0, maxn};
    case G_AUG: return {-1, maxl + 1, -1, maxm + 1, 0, maxn};
    default: return {0, maxn - 2, 0, 0, 0, 0};
    }
}

// Device allocation of a Fortran-shaped field with a padded x pitch.
struct DevArray {
    void* ptr = nullptr;
    Shape sh{};
    long pitch = 0;  // elements per (j,k) row
    int esz = 8;
    long sy() const { return pitch; }
    long sz() const { return pitch * sh.ny(); }
    size_t bytes() const { return size_t(pitch) * sh.ny() * sh.nz() * esz; }
    template <typename T>
    FieldView<T> view() const
    {
        FieldView<T> v;
        v.sy = sy();
        v.sz = sz();
        v.p = static_cast<T*>(ptr) - sh.lx - long(sh.ly) * v.sy - long(sh.lz) * v.sz;
        return v;
    }
};

long pitch_for(long nx, int esz) { return (nx * esz + 127) / 128 * 128 / esz; }

}  // namespace

// ---------------------------------------------------------------------------
// The case
// ---------------------------------------------------------------------------
struct tpls_case {
    int maxl, maxm, maxn, prec;
    // a step of this case failed part-way (its buffer roles may be half
    // rotated): the case refuses further steps
    int broken = 0;
    std::string broken_msg;
    int fail_next = 0;  // test hook: the next step_begin fails before any launch
    uint64_t uid = 0;   // unique for the life of the process (cache keys)
    // the step runs grouped with other cases' (tpls_batch_step): 1, the
    // pressure solve's iterations (the case's own launches leave out the
    // set-up's V-cycle); 2, also the Helmholtz loops and the solve's set-up
    // (the case's own launches stop after the momentum right-hand sides)
    int group_solve = 0;
    // the first step's launches and the first CG chunk run eagerly, so that
    // every launcher's one-time device queries (resident_blocks) happen
    // before any stream capture, never inside one
    int pre_warm = 0, cg_warm = 0;
    tpls_params prm;
    Grid g;
    Phys ph;
    hipStream_t s = nullptr;
    int esz;
    std::vector<DevArray*> owned;

    // state; x[c][0..2] are the three buffers of component c (u,v,w):
    // x2 (the state) and the two Helmholtz iterates
    DevArray* x[3][3];
    int i2[3];
    int i3[3];          // index of the current x3 (after the last step)
    DevArray* conv[3][3];  // [comp][0,1,2]
    DevArray* csf[3][3];
    DevArray* cphi[3];
    DevArray* rhs[3];
    DevArray *phi2, *phi3, *pra, *prb, *D, *sg, *ist;
    DevArray *fx, *fy, *fz, *mu;
    DevArray* pres;   // storage precision: outlet BC and projection
    DevArray* pold;   // double: CG unknown, pres_old
    DevArray *cg_r, *cg_p0, *cg_p1;
    // multigrid preconditioner (ksp_pc >= 1): z and t on the pressure grid,
    // levels 1.. in their own arrays
    DevArray* cg_z = nullptr;
    MgHierarchy mg{};
    double* mg_spec = nullptr;  // the coarsest level's eigenbasis tables (mg.spec)
    MgLevel mg_coef[MG_MAXLEV];
    Tiling tiling;              // launch shapes of the Helmholtz sweeps
    bool use_mg = false;
    DevArray* inlet;
    // diffuse interface method (prm.idm_dim): ConvPrevPhi, the sweeps' inlet
    // profile (maxn+1 values of the storage type, outside `owned` so that
    // loads keep it) and the scalar factors of dim
    DevArray* cprev = nullptr;
    void* dim_prof = nullptr;
    DimCoef dimc{};
    double* red_part = nullptr;
    double* red_out = nullptr;
    CgState* cg = nullptr;
    CgState* h_cg = nullptr;   // pinned host mirror
    double* h_out = nullptr;   // pinned host mirror of red_out
    tpls_diag diag{};
    long steps_done = 0;
    int last_its = 0;
    // pressure solve in flight (the step is split into begin / poll / end so
    // that several cases can be in flight on their own streams)
    int fin[3] = {0, 0, 0};     // Helmholtz result buffer per component
    long ps_launched = 0;       // CG iterations enqueued
    long ps_chunk = 0;          // size of the chunk in flight
    int ps_cur = 0;             // p double-buffer index
    hipEvent_t ps_ev = nullptr; // completes when the chunk's state read-back lands
    bool diag_pending = false;  // residual diagnostics copied back, not yet parsed
    // HIP graphs (prm.graphs): the launches of a step up to the first CG
    // chunk, keyed by the buffer roles of the step, and one CG iteration per
    // p-buffer parity
    struct PreGraph {
        std::vector<const void*> key;
        hipGraphExec_t exec;
    };
    std::vector<PreGraph> pre_graphs;
    hipGraphExec_t it_graph[2] = {nullptr, nullptr};
    // kernel timing: the selected TPLS_TIME_* class
    int timed = TPLS_TIME_OFF;
    std::vector<hipEvent_t> ev_pool;
    size_t ev_used = 0;
    double t_ms = 0.0;
    long t_n = 0;
    // pinned staging of host <-> device field copies, two halves so that the
    // host copy into one overlaps the DMA of the other (first use allocates)
    char* stage = nullptr;
    hipEvent_t stage_ev[2] = {nullptr, nullptr};
};

namespace {

int alloc_field(tpls_case* c, DevArray** out, const Shape& sh, int esz, bool zero = true)
{
    DevArray* a = new DevArray;
    a->sh = sh;
    a->esz = esz;
    a->pitch = pitch_for(sh.nx(), esz);
    hipError_t e = hipMalloc(&a->ptr, a->bytes());
    if (e != hipSuccess) {
        delete a;
        return fail(TPLS_ERR_NOMEM, std::string("hipMalloc: ") + hipGetErrorString(e));
    }
    c->owned.push_back(a);
    if (zero) HIPCHK(hipMemsetAsync(a->ptr, 0, a->bytes(), c->s));
    *out = a;
    return TPLS_OK;
}

Phys phys_of(const tpls_params& p)
{
    return Phys{p.dx, p.dy, p.dz, p.dt, p.Re, p.mu_plus, p.mu_minus, p.height, p.smooth_width, p.scap};
}

// host <-> device copies of a field, converting to/from the storage type.
// Pitched copies straight from pageable memory run at a fraction of a GB/s;
// going through a pinned buffer in chunks of whole rows keeps the DMA engine
// at PCIe rate while the host copies (and converts) the next chunk.
constexpr size_t STAGE_HALF = size_t(16) << 20;

int stage_init(tpls_case* c)
{
    if (c->stage) return TPLS_OK;
    for (hipEvent_t& e : c->stage_ev)
        if (!e) HIPCHK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    HIPCHK(hipHostMalloc(reinterpret_cast<void**>(&c->stage), 2 * STAGE_HALF, hipHostMallocDefault));
    return TPLS_OK;
}

// Rows of a field from the host: fill(buf, first, n) writes elements
// [first, first + n) of the packed host array, in storage type, to buf.
template <class Fill>
int put_rows(tpls_case* c, DevArray* a, Fill&& fill)
{
    int rc = stage_init(c);
    if (rc) return rc;
    const Shape& sh = a->sh;
    const size_t row = size_t(sh.nx()), esz = size_t(a->esz), dpitch = a->pitch * esz;
    const long rows = sh.ny() * sh.nz();
    const long per = long(std::max<size_t>(1, STAGE_HALF / (row * esz)));
    for (long r0 = 0, k = 0; r0 < rows; r0 += per, ++k) {
        const long nr = std::min(per, rows - r0);
        char* buf = c->stage + (k & 1) * STAGE_HALF;
        if (k >= 2) HIPCHK(hipEventSynchronize(c->stage_ev[k & 1]));  // its previous DMA is done
        if ((rc = fill(buf, size_t(r0) * row, size_t(nr) * row))) return rc;
        HIPCHK(hipMemcpy2DAsync(static_cast<char*>(a->ptr) + size_t(r0) * dpitch, dpitch, buf, row * esz, row * esz,
                                size_t(nr), hipMemcpyHostToDevice, c->s));
        HIPCHK(hipEventRecord(c->stage_ev[k & 1], c->s));
    }
    HIPCHK(hipStreamSynchronize(c->s));
    return TPLS_OK;
}

// Rows of a field to the host: drain(buf, first, n) takes elements
// [first, first + n) of the packed array, in storage type, from buf.
template <class Drain>
int get_rows(tpls_case* c, const DevArray* a, Drain&& drain)
{
    int rc = stage_init(c);
    if (rc) return rc;
    const Shape& sh = a->sh;
    const size_t row = size_t(sh.nx()), esz = size_t(a->esz), spitch = a->pitch * esz;
    const long rows = sh.ny() * sh.nz();
    const long per = long(std::max<size_t>(1, STAGE_HALF / (row * esz)));
    const long n = (rows + per - 1) / per;
    auto issue = [&](long k) {
        const long r0 = k * per, nr = std::min(per, rows - r0);
        if (hipMemcpy2DAsync(c->stage + (k & 1) * STAGE_HALF, row * esz,
                             static_cast<const char*>(a->ptr) + size_t(r0) * spitch, spitch, row * esz, size_t(nr),
                             hipMemcpyDeviceToHost, c->s) != hipSuccess ||
            hipEventRecord(c->stage_ev[k & 1], c->s) != hipSuccess)
            return fail(TPLS_ERR_HIP, "device-to-host copy of a field");
        return int(TPLS_OK);
    };
    for (long k = 0; k < std::min(n, 2L); ++k)
        if ((rc = issue(k))) return rc;
    for (long k = 0; k < n; ++k) {
        HIPCHK(hipEventSynchronize(c->stage_ev[k & 1]));
        const long r0 = k * per, nr = std::min(per, rows - r0);
        if ((rc = drain(c->stage + (k & 1) * STAGE_HALF, size_t(r0) * row, size_t(nr) * row))) {
            hipStreamSynchronize(c->s);
            return rc;
        }
        if (k + 2 < n && (rc = issue(k + 2))) return rc;
    }
    return TPLS_OK;
}

int put_array(tpls_case* c, DevArray* a, const double* host)
{
    return put_rows(c, a, [&](char* buf, size_t q0, size_t n) {
        if (a->esz == 8) {
            std::memcpy(buf, host + q0, n * 8);
        } else {
            float* f = reinterpret_cast<float*>(buf);
            for (size_t q = 0; q < n; ++q) f[q] = float(host[q0 + q]);
        }
        return int(TPLS_OK);
    });
}

int get_array(tpls_case* c, const DevArray* a, double* host)
{
    return get_rows(c, a, [&](const char* buf, size_t q0, size_t n) {
        if (a->esz == 8) {
            std::memcpy(host + q0, buf, n * 8);
        } else {
            const float* f = reinterpret_cast<const float*>(buf);
            for (size_t q = 0; q < n; ++q) host[q0 + q] = double(f[q]);
        }
        return int(TPLS_OK);
    });
}

// periodic y wrap of halo rows (what exchange2d* leaves in them)
void wrap_rows(double* h, const Shape& sh, int maxm)
{
    const long nx = sh.nx(), ny = sh.ny();
    const int P = maxm - 1;
    for (long k = 0; k < sh.nz(); ++k)
        for (int j = sh.ly; j <= sh.uy; ++j) {
            if (j >= 1 && j <= P) continue;
            int src = j < 1 ? j + P : j - P;
            std::memcpy(h + (k * ny + (j - sh.ly)) * nx, h + (k * ny + (src - sh.ly)) * nx, nx * sizeof(double));
        }
}

template <typename T>
int step_impl(tpls_case* c);
int harvest_events(tpls_case* c);
int finish_diag(tpls_case* c);

DevArray* field_ptr(tpls_case* c, int f)
{
    switch (f) {
    case TPLS_F_U2: return c->x[0][c->i2[0]];
    case TPLS_F_V2: return c->x[1][c->i2[1]];
    case TPLS_F_W2: return c->x[2][c->i2[2]];
    case TPLS_F_U3: return c->x[0][c->i3[0]];
    case TPLS_F_V3: return c->x[1][c->i3[1]];
    case TPLS_F_W3: return c->x[2][c->i3[2]];
    case TPLS_F_PHI2: return c->phi2;
    case TPLS_F_PHI3: return c->phi3;
    case TPLS_F_PRES: return c->pres;
    case TPLS_F_PRES_OLD: return c->pold;
    case TPLS_F_CONV0_PHI: return c->cphi[0];
    case TPLS_F_CONV1_PHI: return c->cphi[1];
    case TPLS_F_CONV2_PHI: return c->cphi[2];
    case TPLS_F_FX_CSF: return c->fx;
    case TPLS_F_FY_CSF: return c->fy;
    case TPLS_F_FZ_CSF: return c->fz;
    case TPLS_F_VISC2: return c->mu;
    case TPLS_F_RHS_U: return c->rhs[0];
    case TPLS_F_RHS_V: return c->rhs[1];
    case TPLS_F_RHS_W: return c->rhs[2];
    case TPLS_F_U_INLET: return c->inlet;
    case TPLS_F_CONVPREV_PHI: return c->cprev;
    default: break;
    }
    if (f >= TPLS_F_CONV0_U && f <= TPLS_F_CONV2_W) {
        int q = f - TPLS_F_CONV0_U;
        return c->conv[q / 3][q % 3];
    }
    if (f >= TPLS_F_CSF_U0 && f <= TPLS_F_CSF_W2) {
        int q = f - TPLS_F_CSF_U0;
        return c->csf[q / 3][q % 3];
    }
    return nullptr;  // VISC2_AUG, RHS_P, DIFF_*: not kept by the case
}

// The arrays that carry state from one step to the next, by role: the
// velocities u2, v2, w2, the two older convection and CSF histories of each
// component, the two older phi-advection histories, phi2, the pressure (the
// Helmholtz boundary conditions read it), the CG unknown pres_old (FP64) and
// u_inlet.  Every other array is written within a step before it is read.
std::vector<std::pair<std::string, DevArray*>> state_arrays(tpls_case* c)
{
    static const char* comp = "uvw";
    std::vector<std::pair<std::string, DevArray*>> v;
    for (int q = 0; q < 3; ++q) v.emplace_back(std::string(1, comp[q]) + "2", c->x[q][c->i2[q]]);
    for (int q = 0; q < 3; ++q)
        for (int h = 0; h < 2; ++h) {
            v.emplace_back("conv" + std::to_string(h) + "_" + comp[q], c->conv[q][h]);
            v.emplace_back(std::string("csf_") + comp[q] + std::to_string(h), c->csf[q][h]);
        }
    v.emplace_back("conv0_phi", c->cphi[0]);
    v.emplace_back("conv1_phi", c->cphi[1]);
    v.emplace_back("convprev_phi", c->cprev);
    v.emplace_back("phi2", c->phi2);
    v.emplace_back("pres", c->pres);
    v.emplace_back("pres_old", c->pold);
    v.emplace_back("u_inlet", c->inlet);
    return v;
}

// Order-sensitive 64-bit checksum of a byte range, eight bytes at a step.
uint64_t checksum(uint64_t h, const char* p, size_t n)
{
    size_t q = 0;
    for (; q + 8 <= n; q += 8) {
        uint64_t w;
        std::memcpy(&w, p + q, 8);
        h = (h ^ w) * 0x100000001b3ULL;
        h ^= h >> 29;
    }
    for (; q < n; ++q) h = (h ^ uint8_t(p[q])) * 0x100000001b3ULL;
    return h;
}

constexpr char CKPT_MAGIC[8] = {'T', 'P', 'L', 'S', 'C', 'K', 'P', '1'};

struct CkptHeader {
    char magic[8];
    int32_t maxl, maxm, maxn, precision;
    int64_t steps_done;
    int32_t narrays, last_its;
};

struct CkptArray {
    int32_t esz, pad;
    int64_t count;
    uint64_t sum;
};

}  // namespace

// The eigenbasis tables of a hierarchy's coarsest level in device memory
// (mg_spectral_tables), or *out = nullptr when the level is too large for
// them.  The caller frees *out.
static int upload_spectral(const MgLevel& top, hipStream_t s, double** out)
{
    *out = nullptr;
    const int len = mg_spectral_tables(top.nx, top.ny, top.nz, nullptr);
    if (!len) return TPLS_OK;
    std::vector<double> tab(static_cast<size_t>(len));
    mg_spectral_tables(top.nx, top.ny, top.nz, tab.data());
    double* d = nullptr;
    if (hipMalloc(&d, sizeof(double) * size_t(len)) != hipSuccess) return fail(TPLS_ERR_NOMEM, "hipMalloc");
    if (hipMemcpyAsync(d, tab.data(), sizeof(double) * size_t(len), hipMemcpyHostToDevice, s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess) {
        hipFree(d);
        return fail(TPLS_ERR_HIP, "copy of the coarsest level's tables");
    }
    *out = d;
    return TPLS_OK;
}

int tpls::set_error(int code, const std::string& msg) { return fail(code, msg); }

int tpls::case_dims(const tpls_case* c, int* maxl, int* maxm, int* maxn)
{
    if (!c || !maxl || !maxm || !maxn) return fail(TPLS_ERR_ARG, "null argument");
    *maxl = c->maxl;
    *maxm = c->maxm;
    *maxn = c->maxn;
    return TPLS_OK;
}

// ---------------------------------------------------------------------------
extern "C" {

const char* tpls_last_error(void) { return g_err.c_str(); }

size_t tpls_case_bytes(int maxl, int maxm, int maxn, int precision)
{
    const int esz = precision == TPLS_FP32 ? 4 : 8;
    const size_t rows = size_t(maxm) + 1;
    const size_t pt = pitch_for(maxl + 1, esz), pd = pitch_for(maxl + 1, 8);
    const size_t uv = pt * rows * (maxn - 1), w = pt * rows * maxn, p = pt * rows * (maxn + 1);
    size_t n = 0;
    n += (6 * uv + 3 * w) * esz;             // u,v,w x 3 buffers
    n += (6 * uv + 3 * w) * esz * 2;         // conv + csf histories
    n += (2 * uv + w) * esz;                 // rhs
    n += p * esz * (3 + 4 + 3 + 4 + 1 + 1);  // cphi, phi2/3/ra/rb, D/sg/ist, fx/fy/fz, mu, ConvPrevPhi
    n += p * esz;                            // pres
    n += pd * rows * (maxn + 1) * 8 * 4;     // pold + CG vectors
    n += pd * rows * (maxn + 1) * 8 * 2;     // multigrid z and coarse levels (< 1/7 x 2), if enabled
    return n + (1 << 20);
}

int tpls_case_create(const tpls_params* prm, int maxl, int maxm, int maxn, tpls_case** out)
{
    if (!prm || !out) return fail(TPLS_ERR_ARG, "null argument");
    if (maxl < 6 || maxm < 5 || maxn < 6) return fail(TPLS_ERR_ARG, "grid too small");
    if ((maxm - 1) % 2) return fail(TPLS_ERR_ARG, "maxm-1 must be even (red-black colouring across the periodic seam)");
    if (prm->mom_u < 2 || prm->mom_v < 2 || prm->mom_w < 2) return fail(TPLS_ERR_ARG, "mom_* must be >= 2");
    if (prm->idm_dim && !(prm->epn > 0.0 && prm->Pe > 0.0 && prm->dim_its >= 1))
        return fail(TPLS_ERR_ARG, "idm dim needs epn > 0, Pe > 0 and dim_its (max_iteration_dim) >= 1");
    if (prm->precision != TPLS_FP64 && prm->precision != TPLS_FP32) return fail(TPLS_ERR_ARG, "bad precision");
    if (prm->ksp_pc < 0 || prm->ksp_pc > 8) return fail(TPLS_ERR_ARG, "ksp_pc must be 0 (none) or 1..8 (multigrid sweeps)");
    auto rows_ok = [](int r) { return r == 0 || r == 1 || r == 2 || r == 4 || r == 8; };
    if (!rows_ok(prm->tile_rows) || !rows_ok(prm->tile_rows_odd))
        return fail(TPLS_ERR_ARG, "tile rows must be 0 (default), 1, 2, 4 or 8");
    if (prm->prefetch_odd != 0 && prm->prefetch_odd != 4 && prm->prefetch_odd != 8)
        return fail(TPLS_ERR_ARG, "prefetch_odd must be 0 (default), 4 or 8");
    auto rb_ok = [](int r) { return r == 0 || r == 2 || r == 4 || r == 8; };
    if (prm->sor_fused < -1 || prm->sor_fused > 7) return fail(TPLS_ERR_ARG, "sor_fused must be -1 .. 7");
    if (prm->mg_fp32 < -1 || prm->mg_fp32 > 1) return fail(TPLS_ERR_ARG, "mg_fp32 must be -1, 0 or 1");
    if (!rb_ok(prm->tile_rows_rb) || !rb_ok(prm->tile_rows_rb_odd) ||
        (prm->tile_rows_jst != -1 && prm->tile_rows_jst != 16 && !rb_ok(prm->tile_rows_jst)))
        return fail(TPLS_ERR_ARG, "tile_rows_rb and tile_rows_rb_odd must be 0 (default), 2, 4 or 8 and "
                                  "tile_rows_jst -1 (unstaged), 0 (default), 2, 4, 8 or 16");
    if (prm->march_levels < 0) return fail(TPLS_ERR_ARG, "march_levels must be 0 (sized to the device) or >= 1");
    if (prm->mg_pair < -1 || prm->mg_pair > 1) return fail(TPLS_ERR_ARG, "mg_pair must be -1, 0 or 1");
    if (prm->ksp_pc > 0) {
        MgLevel L[MG_MAXLEV];
        const int nl = mg_plan(maxl - 1, maxm - 1, maxn - 1, 1, 1, 1, 6, L);
        const MgLevel& t = L[nl - 1];
        if (long(t.nx) * t.ny * t.nz > MG_COARSE_MAX)
            return fail(TPLS_ERR_ARG, "multigrid: the coarsest level has " + std::to_string(long(t.nx) * t.ny * t.nz) +
                                          " cells (at most " + std::to_string(MG_COARSE_MAX) +
                                          "); halving stops when nx or nz is odd or ny is not a multiple of 4");
    }
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return fail(TPLS_ERR_NOGPU, "no HIP device");
    tpls_case* c = new tpls_case;
    static std::atomic<uint64_t> next_uid{1};
    c->uid = next_uid++;
    c->maxl = maxl;
    c->maxm = maxm;
    c->maxn = maxn;
    c->prm = *prm;
    c->prec = prm->precision;
    c->esz = c->prec == TPLS_FP32 ? 4 : 8;
    c->g = Grid{maxl, maxm, maxn, maxm - 1};
    c->ph = phys_of(*prm);
    if (prm->tile_rows) c->tiling.rows = prm->tile_rows;
    if (prm->tile_rows_odd) c->tiling.rows_odd = prm->tile_rows_odd;
    if (prm->prefetch_odd) c->tiling.depth_odd = prm->prefetch_odd;
    c->tiling.fused = prm->sor_fused < 0 ? 0 : (prm->sor_fused ? prm->sor_fused : 1);
    if (prm->tile_rows_rb) c->tiling.rb_rows = prm->tile_rows_rb;
    if (prm->tile_rows_rb_odd) c->tiling.rb_rows_odd = prm->tile_rows_rb_odd;
    c->tiling.jst = prm->tile_rows_jst < 0 ? 0 : (prm->tile_rows_jst ? prm->tile_rows_jst : 4);
    c->tiling.ks = prm->march_levels;
    int rc = TPLS_OK;
    auto bail = [&](int r) {
        tpls_case_destroy(c);
        return r;
    };
    if (hipStreamCreateWithFlags(&c->s, hipStreamNonBlocking) != hipSuccess) return bail(fail(TPLS_ERR_HIP, "stream"));
    const Shape suv = shape_of(G_UV, maxl, maxm, maxn), sw = shape_of(G_W, maxl, maxm, maxn);
    const Shape sp = shape_of(G_P, maxl, maxm, maxn), si = shape_of(G_INLET, maxl, maxm, maxn);
    const int e = c->esz;
    for (int q = 0; q < 3; ++q) {
        const Shape& s = q == 2 ? sw : suv;
        for (int b = 0; b < 3; ++b) {
            if ((rc = alloc_field(c, &c->x[q][b], s, e))) return bail(rc);
            if ((rc = alloc_field(c, &c->conv[q][b], s, e))) return bail(rc);
            if ((rc = alloc_field(c, &c->csf[q][b], s, e))) return bail(rc);
        }
        if ((rc = alloc_field(c, &c->rhs[q], s, e))) return bail(rc);
        c->i2[q] = 0;
        c->i3[q] = 1;
    }
    for (int b = 0; b < 3; ++b)
        if ((rc = alloc_field(c, &c->cphi[b], sp, e))) return bail(rc);
    DevArray** pg[] = {&c->phi2, &c->phi3, &c->pra, &c->prb, &c->D, &c->sg, &c->fx, &c->fy, &c->fz, &c->mu, &c->pres};
    for (DevArray** a : pg)
        if ((rc = alloc_field(c, a, sp, e))) return bail(rc);
    if ((rc = alloc_field(c, &c->ist, sp, 4))) return bail(rc);
    DevArray** pd[] = {&c->pold, &c->cg_r, &c->cg_p0, &c->cg_p1};
    for (DevArray** a : pd)
        if ((rc = alloc_field(c, a, sp, 8))) return bail(rc);
    if ((rc = alloc_field(c, &c->inlet, si, e))) return bail(rc);
    if ((rc = alloc_field(c, &c->cprev, sp, e))) return bail(rc);
    if (prm->idm_dim) {
        dim_coefficients(prm->dx, prm->dy, prm->dz, prm->dt, prm->epn, prm->Pe, &c->dimc);
        std::vector<double> prof(maxn + 1);
        dim_inlet(maxn, prm->dz, prm->height, prm->epn, prof.data());
        std::vector<float> pf(prof.begin(), prof.end());
        const void* src = e == 4 ? static_cast<const void*>(pf.data()) : static_cast<const void*>(prof.data());
        if (hipMalloc(&c->dim_prof, size_t(maxn + 1) * e) != hipSuccess) return bail(fail(TPLS_ERR_NOMEM, "profile"));
        if (hipMemcpy(c->dim_prof, src, size_t(maxn + 1) * e, hipMemcpyHostToDevice) != hipSuccess)
            return bail(fail(TPLS_ERR_HIP, "profile upload"));
    }
    if (prm->ksp_pc > 0) {
        c->use_mg = true;
        c->mg.nu = prm->ksp_pc;
        c->mg.nlev = mg_plan(maxl - 1, maxm - 1, maxn - 1, 1, 1, 1, 6, c->mg.L);
        // the FP32 cycle needs a coarse level (a one-level hierarchy solves
        // on the FP64 residual itself)
        c->mg.fp32 = c->mg.nlev > 1 && (prm->mg_fp32 > 0 || (prm->mg_fp32 == 0 && prm->precision == TPLS_FP32));
        c->mg.L[0].b = c->cg_r->view<double>();
        if ((rc = upload_spectral(c->mg.L[c->mg.nlev - 1], c->s, &c->mg_spec))) return bail(rc);
        c->mg.spec = c->mg_spec;
        if (c->mg.fp32) {
            for (int l = 0; l < c->mg.nlev; ++l) {
                MgLevel32& o = c->mg.F[l];
                const MgLevel& d = c->mg.L[l];
                const Shape sl = l ? Shape{0, d.nx + 1, 0, d.ny + 1, 0, d.nz + 1} : sp;
                DevArray* ax;
                if ((rc = alloc_field(c, &ax, sl, 4))) return bail(rc);
                o.x = ax->view<float>();
                // level 0: the CG residual converted once per cycle
                DevArray* ab;
                if ((rc = alloc_field(c, &ab, sl, 4))) return bail(rc);
                o.b = ab->view<float>();
            }
        } else {
            if ((rc = alloc_field(c, &c->cg_z, sp, 8))) return bail(rc);
            c->mg.L[0].x = c->cg_z->view<double>();
            for (int l = 1; l < c->mg.nlev; ++l) {
                MgLevel& o = c->mg.L[l];
                const Shape sl{0, o.nx + 1, 0, o.ny + 1, 0, o.nz + 1};
                DevArray *ax, *ab;
                if ((rc = alloc_field(c, &ax, sl, 8)) || (rc = alloc_field(c, &ab, sl, 8))) return bail(rc);
                o.x = ax->view<double>();
                o.b = ab->view<double>();
            }
        }
        // level 0's second vector: its smoothing pairs as one pass each
        const MgLevel& l0 = c->mg.L[0];
        const bool pair = prm->mg_pair > 0 ? l0.ny % 8 == 0 : (prm->mg_pair == 0 && mg_pair_fits(l0.nx, l0.ny, l0.nz));
        if (c->mg.nlev > 1 && pair) {
            DevArray* a2;
            if ((rc = alloc_field(c, &a2, sp, c->mg.fp32 ? 4 : 8))) return bail(rc);
            if (c->mg.fp32) c->mg.x2f = a2->view<float>();
            else c->mg.x2 = a2->view<double>();
            // and level 1's, where it has a coarser level and fits the pass
            const MgLevel& l1 = c->mg.L[1];
            if (c->mg.nlev > 2 && mg_pair_fits(l1.nx, l1.ny, l1.nz)) {
                DevArray* a21;
                const Shape s1{0, l1.nx + 1, 0, l1.ny + 1, 0, l1.nz + 1};
                if ((rc = alloc_field(c, &a21, s1, c->mg.fp32 ? 4 : 8))) return bail(rc);
                if (c->mg.fp32) c->mg.x21f = a21->view<float>();
                else c->mg.x21 = a21->view<double>();
            }
        }
    }
    const long nb = std::max<long>(cg_num_blocks(c->g), (long)REDUCE_BLOCKS);
    if (hipMalloc(&c->red_part, sizeof(double) * 3 * nb) != hipSuccess) return bail(fail(TPLS_ERR_NOMEM, "partials"));
    if (hipMalloc(&c->red_out, sizeof(double) * 8) != hipSuccess) return bail(fail(TPLS_ERR_NOMEM, "outs"));
    if (hipMalloc(&c->cg, sizeof(CgState)) != hipSuccess) return bail(fail(TPLS_ERR_NOMEM, "cg state"));
    if (hipHostMalloc(&c->h_cg, sizeof(CgState)) != hipSuccess) return bail(fail(TPLS_ERR_NOMEM, "pinned"));
    if (hipHostMalloc(&c->h_out, sizeof(double) * 8) != hipSuccess) return bail(fail(TPLS_ERR_NOMEM, "pinned"));
    if (hipEventCreateWithFlags(&c->ps_ev, hipEventDisableTiming) != hipSuccess)
        return bail(fail(TPLS_ERR_HIP, "event"));
    if (hipMemsetAsync(c->red_out, 0, sizeof(double) * 8, c->s) != hipSuccess) return bail(fail(TPLS_ERR_HIP, "memset"));
    if (hipStreamSynchronize(c->s) != hipSuccess) return bail(fail(TPLS_ERR_HIP, "sync"));
    *out = c;
    return TPLS_OK;
}

int tpls_case_destroy(tpls_case* c)
{
    if (!c) return TPLS_OK;
    if (c->s) hipStreamSynchronize(c->s);
    for (DevArray* a : c->owned) {
        hipFree(a->ptr);
        delete a;
    }
    hipFree(c->red_part);
    hipFree(c->red_out);
    hipFree(c->cg);
    hipFree(c->mg_spec);
    hipFree(c->dim_prof);
    if (c->h_cg) hipHostFree(c->h_cg);
    if (c->h_out) hipHostFree(c->h_out);
    for (hipEvent_t e : c->ev_pool) hipEventDestroy(e);
    if (c->ps_ev) hipEventDestroy(c->ps_ev);
    for (hipEvent_t e : c->stage_ev)
        if (e) hipEventDestroy(e);
    if (c->stage) hipHostFree(c->stage);
    for (auto& pg : c->pre_graphs) hipGraphExecDestroy(pg.exec);
    for (hipGraphExec_t e : c->it_graph)
        if (e) hipGraphExecDestroy(e);
    if (c->s) hipStreamDestroy(c->s);
    delete c;
    return TPLS_OK;
}

int tpls_case_load(tpls_case* c, const double* u, const double* v, const double* w, const double* pres,
                   const double* phi)
{
    if (!c || !u || !v || !w || !pres || !phi) return fail(TPLS_ERR_ARG, "null argument");
    for (DevArray* a : c->owned) HIPCHK(hipMemsetAsync(a->ptr, 0, a->bytes(), c->s));
    for (int q = 0; q < 3; ++q) {
        c->i2[q] = 0;
        c->i3[q] = 1;
    }
    int rc;
    if ((rc = put_array(c, c->x[0][0], u))) return rc;
    if ((rc = put_array(c, c->x[1][0], v))) return rc;
    if ((rc = put_array(c, c->x[2][0], w))) return rc;
    if ((rc = put_array(c, c->pres, pres))) return rc;
    if ((rc = put_array(c, c->phi2, phi))) return rc;
    // u_inlet = u(0,0,:), main_ns_hybrid.F90:412
    const long nx = c->maxl + 1, ny = c->maxm + 1;
    std::vector<double> inl(c->maxn - 1);
    for (int k = 0; k <= c->maxn - 2; ++k) inl[k] = u[k * nx * ny];
    if ((rc = put_array(c, c->inlet, inl.data()))) return rc;
    c->steps_done = 0;
    c->last_its = 0;
    std::memset(&c->diag, 0, sizeof c->diag);
    return TPLS_OK;
}

int tpls_case_get_field(tpls_case* c, int f, double* host)
{
    if (!c || !host || f < 0 || f >= TPLS_NFIELDS) return fail(TPLS_ERR_ARG, "bad argument");
    DevArray* a = field_ptr(c, f);
    const Shape sh = shape_of(grid_of(f), c->maxl, c->maxm, c->maxn);
    if (f == TPLS_F_VISC2_AUG) {
        // the case keeps one viscosity array on the p grid; the augmented
        // copy of main_ns_hybrid.F90:650-718 holds the same values
        const Shape sp = shape_of(G_P, c->maxl, c->maxm, c->maxn);
        std::vector<double> t(sp.count());
        int rc = get_array(c, c->mu, t.data());
        if (rc) return rc;
        std::fill(host, host + sh.count(), 0.0);
        const int P = c->maxm - 1;
        for (long k = 0; k < sh.nz(); ++k)
            for (int j = -1; j <= c->maxm + 1; ++j) {
                int sj = j < 1 ? j + P : (j > P ? j - P : j);
                for (int i = 0; i <= c->maxl; ++i)
                    host[(i + 1) + (j + 1) * sh.nx() + k * sh.nx() * sh.ny()] = t[i + sj * sp.nx() + k * sp.nx() * sp.ny()];
            }
        return TPLS_OK;
    }
    if (f == TPLS_F_PRES_OLD && c->steps_done > 0) a = c->pres;  // pres_old = pres after every step
    if (!a) return fail(TPLS_ERR_ARG, "field not kept by the device case");
    int rc = get_array(c, a, host);
    if (rc) return rc;
    if (grid_of(f) != G_INLET) wrap_rows(host, sh, c->maxm);
    return TPLS_OK;
}

int tpls_case_set_field(tpls_case* c, int f, const double* host)
{
    if (!c || !host || f < 0 || f >= TPLS_NFIELDS) return fail(TPLS_ERR_ARG, "bad argument");
    DevArray* a = field_ptr(c, f);
    if (!a) return fail(TPLS_ERR_ARG, "field not kept by the device case");
    if (f == TPLS_F_PRES_OLD) {
        // the CG unknown is kept in double precision
        return put_array(c, c->pold, host);
    }
    return put_array(c, a, host);
}

int tpls_case_save_field(tpls_case* c, int f, const char* filename, const char* description)
{
    if (!c || !filename || !description || f < 0 || f >= TPLS_NFIELDS) return fail(TPLS_ERR_ARG, "bad argument");
    const Shape sh = shape_of(grid_of(f), c->maxl, c->maxm, c->maxn);
    std::vector<double> host(size_t(sh.count()));
    int rc = tpls_case_get_field(c, f, host.data());
    if (rc) return rc;
    return tpls_save_data(filename, description, host.data(), sh.lx, sh.ux, sh.ly, sh.uy, sh.lz, sh.uz);
}

int tpls_case_checkpoint(tpls_case* c, const char* filename)
{
    if (!c || !filename) return fail(TPLS_ERR_ARG, "null argument");
    HIPCHK(hipStreamSynchronize(c->s));
    const std::string tmp = std::string(filename) + ".tmp";
    FILE* fp = std::fopen(tmp.c_str(), "wb");
    if (!fp) return fail(TPLS_ERR_IO, tmp + ": " + std::strerror(errno));
    auto arrays = state_arrays(c);
    CkptHeader h{};
    std::memcpy(h.magic, CKPT_MAGIC, 8);
    h.maxl = c->maxl;
    h.maxm = c->maxm;
    h.maxn = c->maxn;
    h.precision = c->prm.precision;
    h.steps_done = c->steps_done;
    h.narrays = int32_t(arrays.size());
    h.last_its = c->last_its;
    int rc = std::fwrite(&h, sizeof h, 1, fp) == 1 ? TPLS_OK : fail(TPLS_ERR_IO, tmp + ": write failed");
    std::vector<char> buf;
    for (auto& na : arrays) {
        if (rc) break;
        const DevArray* a = na.second;
        const size_t esz = size_t(a->esz), count = size_t(a->sh.count());
        buf.resize(count * esz);
        rc = get_rows(c, a, [&](const char* b, size_t q0, size_t n) {
            std::memcpy(buf.data() + q0 * esz, b, n * esz);
            return int(TPLS_OK);
        });
        if (rc) break;
        CkptArray ah{int32_t(esz), 0, int64_t(count), checksum(0xcbf29ce484222325ULL, buf.data(), buf.size())};
        if (std::fwrite(&ah, sizeof ah, 1, fp) != 1 || std::fwrite(buf.data(), 1, buf.size(), fp) != buf.size())
            rc = fail(TPLS_ERR_IO, tmp + ": write failed");
    }
    if (std::fclose(fp) != 0 && !rc) rc = fail(TPLS_ERR_IO, tmp + ": close failed");
    if (!rc && std::rename(tmp.c_str(), filename) != 0)
        rc = fail(TPLS_ERR_IO, std::string(filename) + ": " + std::strerror(errno));
    if (rc) std::remove(tmp.c_str());
    return rc;
}

int tpls_case_restore(tpls_case* c, const char* filename)
{
    if (!c || !filename) return fail(TPLS_ERR_ARG, "null argument");
    FILE* fp = std::fopen(filename, "rb");
    if (!fp) return fail(TPLS_ERR_IO, std::string(filename) + ": " + std::strerror(errno));
    struct Closer {
        FILE* f;
        ~Closer() { std::fclose(f); }
    } closer{fp};
    HIPCHK(hipStreamSynchronize(c->s));
    auto arrays = state_arrays(c);
    CkptHeader h{};
    if (std::fread(&h, sizeof h, 1, fp) != 1 || std::memcmp(h.magic, CKPT_MAGIC, 8) != 0)
        return fail(TPLS_ERR_FORMAT, std::string(filename) + ": not a case checkpoint");
    if (h.maxl != c->maxl || h.maxm != c->maxm || h.maxn != c->maxn || h.precision != c->prm.precision)
        return fail(TPLS_ERR_ARG, std::string(filename) + ": checkpoint of a " + std::to_string(h.maxl) + "x" +
                                      std::to_string(h.maxm) + "x" + std::to_string(h.maxn) +
                                      (h.precision == TPLS_FP32 ? " FP32" : " FP64") + " case");
    if (h.narrays != int32_t(arrays.size()))
        return fail(TPLS_ERR_FORMAT, std::string(filename) + ": unexpected number of arrays");
    std::vector<char> buf;
    for (auto& na : arrays) {
        DevArray* a = na.second;
        const size_t esz = size_t(a->esz), count = size_t(a->sh.count());
        CkptArray ah{};
        if (std::fread(&ah, sizeof ah, 1, fp) != 1 || size_t(ah.esz) != esz || size_t(ah.count) != count)
            return fail(TPLS_ERR_FORMAT, std::string(filename) + ": bad record for " + na.first);
        buf.resize(count * esz);
        if (std::fread(buf.data(), 1, buf.size(), fp) != buf.size())
            return fail(TPLS_ERR_FORMAT, std::string(filename) + ": truncated at " + na.first);
        if (checksum(0xcbf29ce484222325ULL, buf.data(), buf.size()) != ah.sum)
            return fail(TPLS_ERR_FORMAT, std::string(filename) + ": checksum mismatch in " + na.first +
                                             " (the case state is now undefined)");
        int rc = put_rows(c, a, [&](char* b, size_t q0, size_t n) {
            std::memcpy(b, buf.data() + q0 * esz, n * esz);
            return int(TPLS_OK);
        });
        if (rc) return rc;
    }
    c->steps_done = h.steps_done;
    c->last_its = h.last_its;
    c->diag_pending = false;
    std::memset(&c->diag, 0, sizeof c->diag);
    return TPLS_OK;
}

int tpls_case_steps(tpls_case* c, long* steps)
{
    if (!c || !steps) return fail(TPLS_ERR_ARG, "null argument");
    *steps = c->steps_done;
    return TPLS_OK;
}

int tpls_case_diag(tpls_case* c, tpls_diag* out)
{
    if (!c || !out) return fail(TPLS_ERR_ARG, "null argument");
    int rc = finish_diag(c);
    if (rc) return rc;
    *out = c->diag;
    return TPLS_OK;
}

int tpls_case_sync(tpls_case* c)
{
    if (!c) return fail(TPLS_ERR_ARG, "null case");
    HIPCHK(hipStreamSynchronize(c->s));
    return TPLS_OK;
}

int tpls_case_time_kernels(tpls_case* c, int kernel_class)
{
    if (!c) return fail(TPLS_ERR_ARG, "null case");
    if (kernel_class < TPLS_TIME_OFF || kernel_class > TPLS_TIME_ADVECT) return fail(TPLS_ERR_ARG, "bad kernel class");
    HIPCHK(hipStreamSynchronize(c->s));
    c->timed = kernel_class;
    c->t_ms = 0.0;
    c->t_n = 0;
    c->ev_used = 0;
    return TPLS_OK;
}

int tpls_case_kernel_time(tpls_case* c, double* ms, long* launches)
{
    if (!c || !ms || !launches) return fail(TPLS_ERR_ARG, "null argument");
    HIPCHK(hipStreamSynchronize(c->s));
    int rc = harvest_events(c);
    if (rc) return rc;
    *ms = c->t_ms;
    *launches = c->t_n;
    return TPLS_OK;
}

int tpls_case_set_tiling(tpls_case* c, int rows, int rows_odd, int prefetch_odd, int rows_rb, int rows_rb_odd,
                         int rows_jst, int sor_fused)
{
    if (sor_fused < -1 || sor_fused > 7) return fail(TPLS_ERR_ARG, "sor_fused must be -1 .. 7");
    if (!c) return fail(TPLS_ERR_ARG, "null case");
    auto rows_ok = [](int r) { return r == 1 || r == 2 || r == 4 || r == 8; };
    auto rb_ok = [](int r) { return r == 2 || r == 4 || r == 8; };
    if (!rows_ok(rows) || !rows_ok(rows_odd) || (prefetch_odd != 4 && prefetch_odd != 8) || !rb_ok(rows_rb) ||
        !rb_ok(rows_rb_odd) || (rows_jst != -1 && rows_jst != 16 && !rb_ok(rows_jst)))
        return fail(TPLS_ERR_ARG, "tile rows must be 1, 2, 4 or 8, prefetch_odd 4 or 8, the fused pass's rows 2, 4 "
                                  "or 8 and the staged Jacobi's 2, 4, 8 or 16 (-1: unstaged)");
    HIPCHK(hipStreamSynchronize(c->s));
    c->tiling.rows = rows;
    c->tiling.rows_odd = rows_odd;
    c->tiling.depth_odd = prefetch_odd;
    c->tiling.rb_rows = rows_rb;
    c->tiling.rb_rows_odd = rows_rb_odd;
    c->tiling.jst = rows_jst < 0 ? 0 : rows_jst;
    c->tiling.fused = sor_fused < 0 ? 0 : (sor_fused ? sor_fused : 1);
    for (auto& pg : c->pre_graphs) hipGraphExecDestroy(pg.exec);  // captured with the old shapes
    c->pre_graphs.clear();
    return TPLS_OK;
}

int tpls_case_volume(tpls_case* c, double* volume)
{
    if (!c || !volume) return fail(TPLS_ERR_ARG, "null argument");
    const Reduce red = Reduce{c->red_part, c->red_out + 6};
    if (c->prec == TPLS_FP32)
        HIPCHK(launch_volume<float>(c->s, c->g, c->ph, c->phi2->view<float>(), red, c->prm.idm_dim ? 0.5 : 0.0));
    else
        HIPCHK(launch_volume<double>(c->s, c->g, c->ph, c->phi2->view<double>(), red, c->prm.idm_dim ? 0.5 : 0.0));
    HIPCHK(hipMemcpyAsync(c->h_out + 6, c->red_out + 6, sizeof(double), hipMemcpyDeviceToHost, c->s));
    HIPCHK(hipStreamSynchronize(c->s));
    *volume = c->h_out[6];
    return TPLS_OK;
}

int tpls_step(tpls_case* c, int nsteps)
{
    if (!c || nsteps < 0) return fail(TPLS_ERR_ARG, "bad argument");
    if (c->broken) return fail(TPLS_ERR_ARG, "the case failed in an earlier step: " + c->broken_msg);
    for (int n = 0; n < nsteps; ++n) {
        int rc = c->prec == TPLS_FP32 ? step_impl<float>(c) : step_impl<double>(c);
        if (rc) {
            c->broken = 1;
            c->broken_msg = g_err;
            return rc;
        }
    }
    return TPLS_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// One timestep
// ---------------------------------------------------------------------------
namespace {

Reduce red_of(tpls_case* c, int slot) { return Reduce{c->red_part, c->red_out + slot}; }

#define LCHK(x)                                                                              \
    do {                                                                                     \
        hipError_t e_ = (x);                                                                 \
        if (e_ != hipSuccess) return fail(TPLS_ERR_HIP, std::string(#x) + ": " + hipGetErrorString(e_)); \
    } while (0)

#define RC(x)                 \
    do {                      \
        int r_ = (x);         \
        if (r_) return r_;    \
    } while (0)

hipEvent_t next_event(tpls_case* c)
{
    if (c->ev_used == c->ev_pool.size()) {
        hipEvent_t e;
        hipEventCreate(&e);
        c->ev_pool.push_back(e);
    }
    return c->ev_pool[c->ev_used++];
}

int harvest_events(tpls_case* c)
{
    for (size_t q = 0; q + 1 < c->ev_used; q += 2) {
        float ms = 0.f;
        HIPCHK(hipEventElapsedTime(&ms, c->ev_pool[q], c->ev_pool[q + 1]));
        c->t_ms += ms;
        c->t_n += 1;
    }
    c->ev_used = 0;
    return TPLS_OK;
}

// The pressure solve: KSPSolve replacement (main_ns_hybrid.F90:896-913).
// Iterations are enqueued in chunks ahead of the convergence test (every
// kernel of a converged solve returns at once); after each chunk the device
// state is copied back to pinned memory and ps_ev recorded.  ps_start
// enqueues the set-up and the first chunk, ps_poll (once ps_ev has completed)
// either finishes the solve or enqueues the next chunk.  The unknown is pold,
// the residual cg_r.
// Record the launches of fn on the case's stream into an executable graph.
template <typename F>
int capture(tpls_case* c, hipGraphExec_t* out, F&& fn)
{
    HIPCHK(hipStreamBeginCapture(c->s, hipStreamCaptureModeThreadLocal));
    int rc = fn();
    hipGraph_t gr = nullptr;
    hipError_t e = hipStreamEndCapture(c->s, &gr);
    if (!rc && e != hipSuccess) rc = fail(TPLS_ERR_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
    if (!rc) {
        e = hipGraphInstantiate(out, gr, nullptr, nullptr, 0);
        if (e != hipSuccess) rc = fail(TPLS_ERR_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
    }
    if (gr) hipGraphDestroy(gr);
    return rc;
}

int ps_enqueue_chunk(tpls_case* c)
{
    const Grid& g = c->g;
    const bool mg = c->use_mg;
    FieldView<double> x = c->pold->view<double>(), r = c->cg_r->view<double>();
    const bool z32 = mg && c->mg.fp32;
    FieldView<double> z = mg && !z32 ? c->cg_z->view<double>() : r;
    DevArray* pb[2] = {c->cg_p0, c->cg_p1};
    long chunk = std::min<long>(c->ps_chunk, std::max<long>(1, c->prm.ksp_maxits - c->ps_launched));
    c->ps_chunk = chunk;
    auto iteration = [&](int cur) -> int {
        FieldView<double> po = pb[cur]->view<double>(), pn = pb[cur ^ 1]->view<double>();
        if (z32)
            LCHK(launch_cg_p(c->s, g, c->ph, c->mg.F[0].x, po, pn, 1, red_of(c, 7), c->cg));
        else
            LCHK(launch_cg_p(c->s, g, c->ph, z, po, pn, mg ? 1 : 0, red_of(c, 7), c->cg));
        const bool tx = c->timed == TPLS_TIME_CG_X;
        if (tx) LCHK(hipEventRecord(next_event(c), c->s));
        LCHK(launch_cg_x(c->s, g, c->ph, pn, x, r, red_of(c, 7), c->cg));
        if (tx) LCHK(hipEventRecord(next_event(c), c->s));
        if (mg) LCHK(launch_mg_vcycle(c->s, g, c->mg, red_of(c, 7), c->cg));
        return TPLS_OK;
    };
    const bool graphs = c->prm.graphs && !c->timed && c->cg_warm;
    c->cg_warm = 1;
    for (long q = 0; q < chunk; ++q) {
        if (graphs) {
            hipGraphExec_t& ex = c->it_graph[c->ps_cur];
            if (!ex) {
                const int cur = c->ps_cur;
                RC(capture(c, &ex, [&]() { return iteration(cur); }));
            }
            LCHK(hipGraphLaunch(ex, c->s));
        } else {
            RC(iteration(c->ps_cur));
        }
        c->ps_cur ^= 1;
    }
    c->ps_launched += chunk;
    LCHK(hipMemcpyAsync(c->h_cg, c->cg, sizeof(CgState), hipMemcpyDeviceToHost, c->s));
    LCHK(hipEventRecord(c->ps_ev, c->s));
    return TPLS_OK;
}

// Coefficients of the reduced operator (pressure_solver.F90:394-421) on every
// multigrid level, a quarter per coarser level.
void mg_coefficients(tpls_case* c)
{
    const Phys& ph = c->ph;
    const Grid& g = c->g;
    mg_plan(g.maxl - 1, g.P, g.maxn - 1, 1.0 / (ph.dy * ph.dy), 1.0 / (ph.dx * ph.dx), 1.0 / (ph.dz * ph.dz),
            2.0 * (1.0 / (ph.dx * ph.dx) + 1.0 / (ph.dy * ph.dy) + 1.0 / (ph.dz * ph.dz)), c->mg_coef);
    for (int l = 0; l < c->mg.nlev; ++l) {
        c->mg.L[l].cx = c->mg_coef[l].cx;
        c->mg.L[l].cy = c->mg_coef[l].cy;
        c->mg.L[l].cz = c->mg_coef[l].cz;
        c->mg.L[l].diag = c->mg_coef[l].diag;
    }
}

// Launches of the solve's set-up: device state, right-hand side, residual
// and the first preconditioned residual.  No host state changes (capturable).
template <typename T>
int ps_setup(tpls_case* c, FieldView<T> u3, FieldView<T> v3, FieldView<T> w3, FieldView<double> rhs_p, int use_rhs,
             const T* inlet)
{
    const Grid& g = c->g;
    FieldView<double> r = c->cg_r->view<double>();
    LCHK(launch_cg_init(c->s, c->cg, c->prm.ksp_rtol, c->prm.ksp_abstol, c->prm.ksp_dtol, c->prm.ksp_maxits));
    LCHK(launch_cg_setup_b<T>(c->s, g, c->ph, u3, v3, w3, inlet, rhs_p, use_rhs, r, red_of(c, 7), c->cg));
    LCHK(launch_cg_setup_r<T>(c->s, g, c->ph, c->pold->view<double>(), r, red_of(c, 7), c->cg));
    if (c->use_mg && !c->group_solve) LCHK(launch_mg_vcycle(c->s, g, c->mg, red_of(c, 7), c->cg));
    return TPLS_OK;
}

void ps_reset(tpls_case* c)
{
    c->ps_cur = 0;
    c->ps_launched = 0;
    // a multigrid iteration costs tens of launches: enqueue close to the last
    // solve's count
    c->ps_chunk = c->use_mg ? std::max(4, c->last_its) : std::max(8, c->last_its - 2);
}

template <typename T>
int ps_start(tpls_case* c, FieldView<T> u3, FieldView<T> v3, FieldView<T> w3, FieldView<double> rhs_p, int use_rhs,
             const T* inlet)
{
    if (c->use_mg) mg_coefficients(c);
    RC(ps_setup<T>(c, u3, v3, w3, rhs_p, use_rhs, inlet));
    ps_reset(c);
    return ps_enqueue_chunk(c);
}

// After ps_ev has completed: finish the solve (*finished = true) or enqueue
// the next chunk.
int ps_poll(tpls_case* c, bool* finished)
{
    *finished = false;
    if (c->timed == TPLS_TIME_CG_X) {
        // drop pairs recorded after convergence (those launches returned at once)
        long live = std::max<long>(0, c->h_cg->its - (c->ps_launched - c->ps_chunk));
        if (live < c->ps_chunk) c->ev_used -= size_t(2 * (c->ps_chunk - live));
    }
    if (c->timed) {
        int rc = harvest_events(c);
        if (rc) return rc;
    }
    if (c->h_cg->done || c->ps_launched >= c->prm.ksp_maxits) {
        c->diag.ksp_its = c->h_cg->its;
        c->diag.ksp_reason = c->h_cg->reason;
        c->diag.ksp_rnorm = c->h_cg->dp;
        c->last_its = c->h_cg->its;
        *finished = true;
        return TPLS_OK;
    }
    c->ps_chunk = c->use_mg ? 2 : 16;
    return ps_enqueue_chunk(c);
}

template <typename T>
int pressure_solve(tpls_case* c, FieldView<T> u3, FieldView<T> v3, FieldView<T> w3, FieldView<double> rhs_p,
                   int use_rhs, const T* inlet)
{
    int rc = ps_start<T>(c, u3, v3, w3, rhs_p, use_rhs, inlet);
    for (bool fin = false; !rc && !fin;) {
        LCHK(hipEventSynchronize(c->ps_ev));
        rc = ps_poll(c, &fin);
    }
    return rc;
}

// Buffer roles of one step, a pure function of the case state: the
// reinitialisation ping-pong leaves its final iterate as phi3 and the old
// phi3 in the freed slot (two_phase_levelset.f90:100-116); the Helmholtz
// iterations alternate between the two buffers that are not u2 and end in
// fin (main_ns_hybrid.F90:758-887).
struct StepPlan {
    DevArray *phi3, *pra, *prb;
    int fin[3];
};

StepPlan plan_step(const tpls_case* c)
{
    StepPlan p{c->phi3, c->pra, c->prb, {0, 0, 0}};
    const int nit = c->prm.idm_dim ? 0 : c->prm.levelset_its;  // dim writes phi3 in place
    if (nit > 0) {
        DevArray* last = ((nit - 1) & 1) ? c->prb : c->pra;
        p.phi3 = last;
        if (last == c->pra)
            p.pra = c->phi3;
        else
            p.prb = c->phi3;
    }
    const int mom[3] = {c->prm.mom_u, c->prm.mom_v, c->prm.mom_w};
    for (int q = 0; q < 3; ++q) {
        int others[2], n = 0;
        for (int b = 0; b < 3; ++b)
            if (b != c->i2[q]) others[n++] = b;
        p.fin[q] = ((mom[q] / 2) & 1) ? others[1] : others[0];
    }
    return p;
}

// Every launch of a step up to the pressure solve's set-up (main:585-893),
// in the buffer roles of pl; changes no host state, so it can be captured.
template <typename T>
int enqueue_pre(tpls_case* c, const StepPlan& pl)
{
    const Grid& g = c->g;
    const Phys& ph = c->ph;
    hipStream_t s = c->s;
    const bool dg = c->prm.diagnostics != 0;
    auto V = [](DevArray* a) { return a->view<T>(); };

    // ---- do_advect_phi (main:585) ----
    const bool ta = c->timed == TPLS_TIME_ADVECT;
    if (ta) LCHK(hipEventRecord(next_event(c), s));
    LCHK(launch_advect_phi<T>(s, g, ph, V(c->phi2), V(c->x[0][c->i2[0]]), V(c->x[1][c->i2[1]]), V(c->x[2][c->i2[2]]),
                              V(c->cphi[2])));
    if (ta) LCHK(hipEventRecord(next_event(c), s));
    if (c->prm.idm_dim) {
        // ---- dim (main:606-611): phiNew iterates in pra (and prb, the
        // fused sweeps' ping-pong partner), d2phi in D, source0 in sg;
        // phiNew starts from phi3, i.e. phi2 after a step and zero before
        // the first ----
        LCHK(launch_dim<T>(s, g, ph, c->dimc, c->prm.dim_its, c->steps_done == 0 ? 1 : 0, V(c->phi2), V(c->cphi[2]),
                           V(c->cprev), V(c->phi3), V(c->pra), V(c->prb), V(c->D), V(c->sg),
                           static_cast<const T*>(c->dim_prof), dg ? c->red_out : nullptr, c->prm.dim_split ? 0 : 1,
                           c->tiling.ks));
    } else {
        // ---- do_tpls (main:596) ----
        LCHK(launch_tpls_update<T>(s, g, ph, V(c->phi2), V(c->cphi[0]), V(c->cphi[1]), V(c->cphi[2]), V(c->phi3)));
        LCHK(launch_phi_bc<T>(s, g, ph.dz, ph.height, V(c->phi3), 1));
        LCHK(launch_levelset_status_code<T>(s, g, ph.dx, V(c->phi3), V(c->D), c->ist->view<int>()));
        {
            DevArray* src = c->phi3;
            DevArray* bufs[2] = {c->pra, c->prb};
            const int nit = c->prm.levelset_its;
            for (int it = 1; it <= nit; ++it) {
                DevArray* dst = bufs[(it - 1) & 1];
                LCHK(launch_levelset_iter_code<T>(s, g, ph.dx, V(c->D), c->ist->view<int>(), V(src), V(dst)));
                // levelset.f90:168-216 (h = dx is passed as dz there)
                LCHK(launch_phi_bc<T>(s, g, ph.dx, ph.height, V(dst), 1));
                if (dg && it == nit) LCHK(launch_levelset_err<T>(s, g, V(src), V(dst), red_of(c, 0)));
                src = dst;
            }
        }
    }
    // ---- CSF from the reinitialised phi3 (main:618-627) ----
    LCHK(launch_csf<T>(s, g, ph, V(pl.phi3), V(c->fx), V(c->fy), V(c->fz)));
    LCHK(launch_csf_bc<T>(s, g, V(c->fx), V(c->fy), V(c->fz)));
    // ---- viscosity incl. inlet/outlet planes (main:646-710) ----
    LCHK(launch_viscosity<T>(s, g, ph, V(c->phi2), V(c->mu), 0, g.maxl, c->prm.idm_dim ? 1 : 0));
    // ---- conv, diffusion, CSF faces, AB3/CN right-hand sides (main:629-641, 732-757) ----
    DevArray* fcs[3] = {c->fx, c->fy, c->fz};
    for (int q = 0; q < 3; ++q)
        LCHK(launch_momentum<T>(s, g, ph, q, 4, V(c->x[0][c->i2[0]]), V(c->x[1][c->i2[1]]), V(c->x[2][c->i2[2]]),
                                V(c->mu), V(c->conv[q][2]), V(c->rhs[q]), V(fcs[q]), V(c->conv[q][0]),
                                V(c->conv[q][1]), V(c->csf[q][0]), V(c->csf[q][1]), V(c->csf[q][2]), V(c->rhs[q])));
    // a step grouped with the Helmholtz loops runs the rest with other cases'
    if (c->group_solve == 2) return TPLS_OK;
    // ---- Helmholtz outer loops (main:758-887) ----
    // The three components' loops are independent.  In FP32 they run in
    // lock-step, iteration by iteration, so that an odd iteration's fused
    // passes of u, v and w go out as one launch (launch_sor_pair3): 17.3
    // against 17.9 ms per step at 512 x 128 x 256.  In FP64 each component's
    // loop runs whole, as its iterations read what the previous one has just
    // written while the device's last-level cache still holds part of it:
    // the lock-step order there measured 25.7 against 25.6 ms.
    const int mom[3] = {c->prm.mom_u, c->prm.mom_v, c->prm.mom_w};
    int srcb[3], dstb[3];
    for (int q = 0; q < 3; ++q) {
        int others[2], n = 0;
        for (int b = 0; b < 3; ++b)
            if (b != c->i2[q]) others[n++] = b;
        srcb[q] = others[0];
        dstb[q] = others[1];
    }
    const bool lock = sizeof(T) == 4 && (c->tiling.fused & 1) && !c->timed && sor_pair3_fits(g, c->tiling);
    const int nmin = std::min({mom[0], mom[1], mom[2]}) / 2;
    // one outer iteration of component q; three: its fused pass and the
    // boundary conditions after it are left to the three-component launch
    auto iteration = [&](int q, int it, bool three) -> int {
        // bit 0: odd iteration_time fused, bit 1: even
        const bool fused = (c->tiling.fused >> (it % 2 ? 0 : 1)) & 1;
        const bool tj = c->timed == TPLS_TIME_JACOBI;
        const bool ts = c->timed == (it % 2 ? TPLS_TIME_SOR_ODD : TPLS_TIME_SOR_EVEN);
        // fused SOR pass: the Jacobi step writes the third buffer (u2's,
        // dead once the momentum right-hand sides exist) and the pass reads
        // it and writes x3; split: Jacobi into x3, then both sweeps in place
        FieldView<T> xs = V(c->x[q][srcb[q]]), xd = V(c->x[q][dstb[q]]);
        FieldView<T> xj = fused ? V(c->x[q][c->i2[q]]) : xd;
        if (!fused && it % 2 == 0 && (c->tiling.fused & FUSED_JSOR) && c->tiling.jst && jsor0_fits(g) && !tj && !ts) {
            // the Jacobi step, its boundary conditions and the first
            // colour sweep in one pass, then the second sweep as below
            LCHK(launch_jsor0<T>(s, g, ph, q, xs, xd, V(c->rhs[q]), V(c->mu), V(c->pres), c->tiling));
            LCHK(launch_sor<T>(s, g, ph, q, xd, V(c->rhs[q]), V(c->mu), 1, it, c->tiling));
            LCHK(launch_helmholtz_bc<T>(s, g, ph, q, xd, V(c->pres)));
            std::swap(srcb[q], dstb[q]);
            return TPLS_OK;
        }
        if (tj) LCHK(hipEventRecord(next_event(c), s));
        LCHK(launch_jacobi<T>(s, g, ph, q, xs, V(c->rhs[q]), V(c->mu), xj, it == 1, c->tiling));
        if (tj) LCHK(hipEventRecord(next_event(c), s));
        LCHK(launch_helmholtz_bc<T>(s, g, ph, q, xj, V(c->pres)));
        if (three) return TPLS_OK;
        if (fused) {
            if (ts) LCHK(hipEventRecord(next_event(c), s));
            LCHK(launch_sor_pair<T>(s, g, ph, q, xj, xd, V(c->rhs[q]), V(c->mu), it, c->tiling));
            if (ts) LCHK(hipEventRecord(next_event(c), s));
        } else {
            for (int flg = 0; flg < 2; ++flg) {
                if (ts) LCHK(hipEventRecord(next_event(c), s));
                LCHK(launch_sor<T>(s, g, ph, q, xd, V(c->rhs[q]), V(c->mu), flg, it, c->tiling));
                if (ts) LCHK(hipEventRecord(next_event(c), s));
            }
        }
        LCHK(launch_helmholtz_bc<T>(s, g, ph, q, xd, V(c->pres)));
        std::swap(srcb[q], dstb[q]);
        return TPLS_OK;
    };
    if (lock) {
        const int nits = std::max({mom[0], mom[1], mom[2]}) / 2;
        for (int it = 1; it <= nits; ++it) {
            const bool three = it % 2 && it <= nmin;
            // an odd iteration's Jacobi steps go out right after the same
            // component's even iteration, whose output they read while the
            // cache still holds it (the first odd iteration's on their own)
            const bool next3 = it % 2 == 0 && it + 1 <= nmin;
            for (int q = 0; q < 3; ++q) {
                if (it > mom[q] / 2 || (three && it > 1)) continue;
                RC(iteration(q, it, three));
                if (next3) RC(iteration(q, it + 1, true));
            }
            if (!three) continue;
            FieldView<T> xin[3], xout[3], rhs[3];
            for (int q = 0; q < 3; ++q) {
                xin[q] = V(c->x[q][c->i2[q]]);
                xout[q] = V(c->x[q][dstb[q]]);
                rhs[q] = V(c->rhs[q]);
            }
            LCHK(launch_sor_pair3<T>(s, g, ph, xin, xout, rhs, V(c->mu), c->tiling));
            for (int q = 0; q < 3; ++q) {
                LCHK(launch_helmholtz_bc<T>(s, g, ph, q, xout[q], V(c->pres)));
                std::swap(srcb[q], dstb[q]);
            }
        }
    }
    // get_difference(x3, x3_old) of component q's finished loop; at it == 1
    // x3_old is zero
    auto difference = [&](int q) -> int {
        const int nz = q == 2 ? g.maxn : g.maxn - 1;
        if (dg && mom[q] / 2 >= 2)
            LCHK(launch_difference<T>(s, g, nz, V(c->x[q][srcb[q]]), V(c->x[q][dstb[q]]), red_of(c, 1 + q)));
        return TPLS_OK;
    };
    for (int q = 0; q < 3; ++q) {
        if (!lock)
            for (int it = 1; it <= mom[q] / 2; ++it) RC(iteration(q, it, false));
        RC(difference(q));
    }
    // ---- divergence and the pressure solve's set-up (main:889-899) ----
    const int* fin = pl.fin;
    return ps_setup<T>(c, V(c->x[0][fin[0]]), V(c->x[1][fin[1]]), V(c->x[2][fin[2]]), c->cg_r->view<double>(), 0,
                       static_cast<const T*>(c->inlet->ptr));
}

// The step up to the pressure solve's iterations: eager launches, or the
// replay of a graph captured for the same buffer roles.
template <typename T>
int step_pre(tpls_case* c)
{
    if (c->fail_next) {
        c->fail_next = 0;
        return fail(TPLS_ERR_ARG, "failure injected by tpls_test_fail_next_step");
    }
    const StepPlan pl = plan_step(c);
    if (c->use_mg) mg_coefficients(c);
    const bool warm = c->pre_warm;
    c->pre_warm = 1;
    if (c->prm.graphs && !c->timed && warm) {
        std::vector<const void*> key = {c->phi2, c->phi3, c->pra, c->prb, c->cphi[0], c->cphi[1], c->cphi[2]};
        // dim's first step starts its iterate from zero, later ones from phi2
        key.push_back(reinterpret_cast<const void*>(uintptr_t(c->prm.idm_dim && c->steps_done == 0)));
        key.push_back(reinterpret_cast<const void*>(uintptr_t(c->group_solve)));
        for (int q = 0; q < 3; ++q) {
            key.push_back(c->x[q][c->i2[q]]);
            for (int b = 0; b < 3; ++b) {
                key.push_back(c->x[q][b]);
                key.push_back(c->conv[q][b]);
                key.push_back(c->csf[q][b]);
            }
        }
        hipGraphExec_t ex = nullptr;
        for (const auto& pg : c->pre_graphs)
            if (pg.key == key) ex = pg.exec;
        if (!ex) {
            if (c->pre_graphs.size() >= 24) {  // bounded cache; the roles cycle with a short period
                for (auto& pg : c->pre_graphs) hipGraphExecDestroy(pg.exec);
                c->pre_graphs.clear();
            }
            RC(capture(c, &ex, [&]() { return enqueue_pre<T>(c, pl); }));
            c->pre_graphs.push_back({key, ex});
        }
        LCHK(hipGraphLaunch(ex, c->s));
    } else {
        RC(enqueue_pre<T>(c, pl));
    }
    c->phi3 = pl.phi3;
    c->pra = pl.pra;
    c->prb = pl.prb;
    for (int q = 0; q < 3; ++q) c->fin[q] = pl.fin[q];
    return TPLS_OK;
}

// The step up to the first CG chunk.
template <typename T>
int step_begin(tpls_case* c)
{
    c->group_solve = 0;
    RC(step_pre<T>(c));
    ps_reset(c);
    return ps_enqueue_chunk(c);
}

// The rest of the step once the pressure solve has finished.
template <typename T>
int step_end(tpls_case* c)
{
    const Grid& g = c->g;
    const Phys& ph = c->ph;
    hipStream_t s = c->s;
    const bool dg = c->prm.diagnostics != 0;
    auto V = [](DevArray* a) { return a->view<T>(); };
    const int* fin = c->fin;
    FieldView<T> u3 = V(c->x[0][fin[0]]), v3 = V(c->x[1][fin[1]]), w3 = V(c->x[2][fin[2]]);
    const T* inlet = static_cast<const T*>(c->inlet->ptr);
    // ---- copy-out and projection (main:914-918) ----
    LCHK(launch_pres_ghosts<T>(s, g, ph, c->pold->view<double>(), u3, inlet, V(c->pres)));
    LCHK(launch_uvw_pres<T>(s, g, ph, V(c->pres), inlet, u3, v3, w3));
    // ---- history rotation (main:2480-2509) ----
    for (int q = 0; q < 3; ++q) {
        DevArray* t = c->conv[q][0];
        c->conv[q][0] = c->conv[q][1];
        c->conv[q][1] = c->conv[q][2];
        c->conv[q][2] = t;
        t = c->csf[q][0];
        c->csf[q][0] = c->csf[q][1];
        c->csf[q][1] = c->csf[q][2];
        c->csf[q][2] = t;
        c->i3[q] = fin[q];
        c->i2[q] = fin[q];  // u2 <- u3
    }
    {
        DevArray* t = c->cphi[0];
        c->cphi[0] = c->cphi[1];
        c->cphi[1] = c->cphi[2];
        c->cphi[2] = t;
        std::swap(c->phi2, c->phi3);  // phi2 <- phi3
    }
    if (dg) {
        LCHK(hipMemcpyAsync(c->h_out, c->red_out, sizeof(double) * 4, hipMemcpyDeviceToHost, s));
        c->diag_pending = true;
    }
    c->steps_done += 1;
    return TPLS_OK;
}

// Parse the residual diagnostics once their copy has landed.
int finish_diag(tpls_case* c)
{
    if (!c->diag_pending) return TPLS_OK;
    LCHK(hipStreamSynchronize(c->s));
    c->diag.err_phi = c->h_out[0];
    c->diag.err_u = c->h_out[1];
    c->diag.err_v = c->h_out[2];
    c->diag.err_w = c->h_out[3];
    c->diag_pending = false;
    return TPLS_OK;
}

template <typename T>
int step_impl(tpls_case* c)
{
    int rc = step_begin<T>(c);
    for (bool fin = false; !rc && !fin;) {
        LCHK(hipEventSynchronize(c->ps_ev));
        rc = ps_poll(c, &fin);
    }
    if (!rc) rc = step_end<T>(c);
    if (!rc) rc = finish_diag(c);
    return rc;
}

// ---------------------------------------------------------------------------
// The pressure solves of a batch in lock-step: every CG iteration and every
// V-cycle kernel runs as one launch across the cases (group.hpp), on the
// group's own stream, instead of one launch per case on each case's stream.
// The small grids of a batch are launch-bound one by one; grouped, a level's
// sweep of 64 cases is one launch with the blocks of all of them.  Each case
// computes exactly what its own launches compute: the kernels of a case
// whose solve has converged return at once, as they do in a chunk launched
// past convergence.
// ---------------------------------------------------------------------------
int g_batch_grouped = 1;

struct SolveGroup {
    std::vector<uint64_t> key;        // the cases' uids, in batch order
    std::vector<double> coef;         // their level-0 coefficients
    GroupPlan vc, it[2], gather;      // first V-cycle; iteration by p parity; state gather
    GroupPlan hh;                     // this step's Helmholtz loops and solve set-up
    CgState* d_st = nullptr;          // gathered states, read back in one copy
    CgState* h_st = nullptr;
    int n = 0, cap = 0;
    bool any_mg = false;
    hipStream_t s = nullptr;
    hipEvent_t ev = nullptr, done = nullptr;
    ~SolveGroup()
    {
        if (s) (void)hipStreamSynchronize(s);
        if (d_st) (void)hipFree(d_st);
        if (h_st) (void)hipHostFree(h_st);
        if (ev) (void)hipEventDestroy(ev);
        if (done) (void)hipEventDestroy(done);
        if (s) (void)hipStreamDestroy(s);
    }
};

thread_local std::unique_ptr<SolveGroup> t_group;

std::vector<double> group_coef(const std::vector<tpls_case*>& cs)
{
    std::vector<double> v;
    for (tpls_case* c : cs) {
        v.push_back(c->ph.dx);
        v.push_back(c->ph.dy);
        v.push_back(c->ph.dz);
        if (c->use_mg) {
            const MgLevel& o = c->mg.L[0];
            v.insert(v.end(), {o.cx, o.cy, o.cz, o.diag});
        }
    }
    return v;
}

// The group for these cases, built (tables, buffers, stream) on first use
// and kept while the batch keeps its cases.
int group_for(const std::vector<tpls_case*>& cs, SolveGroup** out)
{
    std::vector<uint64_t> key;
    for (tpls_case* c : cs) key.push_back(c->uid);
    const std::vector<double> coef = group_coef(cs);
    if (t_group && t_group->key == key && t_group->coef == coef) {
        *out = t_group.get();
        return TPLS_OK;
    }
    // a new set of cases (serving re-forms its batch every step): the
    // stream, events and buffers stay, the tables are rebuilt
    if (!t_group) {
        auto G = std::make_unique<SolveGroup>();
        HIPCHK(hipStreamCreateWithFlags(&G->s, hipStreamNonBlocking));
        HIPCHK(hipEventCreateWithFlags(&G->ev, hipEventDisableTiming));
        HIPCHK(hipEventCreateWithFlags(&G->done, hipEventDisableTiming));
        t_group = std::move(G);
    }
    SolveGroup* G = t_group.get();
    // the tables are rewritten in place: nothing of a previous call (one that
    // failed part-way included) may still be reading them
    HIPCHK(hipStreamSynchronize(G->s));
    G->key.clear();  // invalid until the rebuild completes
    if (G->cap < int(cs.size())) {
        const int cap = std::max<int>(int(cs.size()), 2 * G->cap);
        if (G->d_st) (void)hipFree(G->d_st);
        if (G->h_st) (void)hipHostFree(G->h_st);
        G->d_st = nullptr;
        G->h_st = nullptr;
        G->cap = 0;
        if (hipMalloc(&G->d_st, sizeof(CgState) * cap) != hipSuccess) return fail(TPLS_ERR_NOMEM, "group states");
        if (hipHostMalloc(&G->h_st, sizeof(CgState) * cap) != hipSuccess) return fail(TPLS_ERR_NOMEM, "pinned");
        G->cap = cap;
    }
    G->n = int(cs.size());
    for (GroupPlan* P : {&G->vc, &G->it[0], &G->it[1], &G->gather}) P->clear();
    std::vector<MgGroupCase> mg;
    std::vector<CgGroupCase> cg;
    std::vector<const CgState*> src;
    for (tpls_case* c : cs) {
        const bool z32 = c->use_mg && c->mg.fp32;
        CgGroupCase k{};
        k.g = c->g;
        k.ph = c->ph;
        k.x = c->pold->view<double>();
        k.r = c->cg_r->view<double>();
        k.pb[0] = c->cg_p0->view<double>();
        k.pb[1] = c->cg_p1->view<double>();
        k.z32 = z32 ? 1 : 0;
        k.z = c->use_mg && !z32 ? c->cg_z->view<double>() : k.r;
        if (z32) k.zf = c->mg.F[0].x;
        k.use_mz = c->use_mg ? 1 : 0;
        k.part = c->red_part;
        k.st = c->cg;
        cg.push_back(k);
        if (c->use_mg) mg.push_back(MgGroupCase{&c->mg, c->g, c->red_part, c->cg});
        src.push_back(c->cg);
    }
    G->any_mg = !mg.empty();
    if (mg_group_vcycle(G->vc, mg)) return fail(TPLS_ERR_ARG, "a multigrid hierarchy the grouped solve cannot run");
    for (int cur = 0; cur < 2; ++cur) {
        cg_group_iteration(G->it[cur], cg, cur);
        mg_group_vcycle(G->it[cur], mg);
    }
    cg_group_gather(G->gather, src, G->d_st);
    for (GroupPlan* P : {&G->vc, &G->it[0], &G->it[1], &G->gather}) HIPCHK(P->upload());
    G->key = key;
    G->coef = coef;
    *out = G;
    return TPLS_OK;
}

// Solve the pressure systems of the cases whose step_pre has been enqueued
// (each with group_solve set), then enqueue their step_end on their own
// streams.  A failure fails every case of the group.
int group_solve(const std::vector<tpls_case*>& cs, int mode)
{
    SolveGroup* G = nullptr;
    RC(group_for(cs, &G));
    // mode 2: the Helmholtz loops and the set-up, in this step's buffer roles
    G->hh.clear();
    std::vector<HhGroupCase<double>> h64;
    std::vector<HhGroupCase<float>> h32;
    std::vector<SetupGroupCase<double>> s64;
    std::vector<SetupGroupCase<float>> s32;
    bool diag = false;
    auto fill = [&](tpls_case* c, auto& hv, auto& sv, auto zero) {
        using T = decltype(zero);
        HhGroupCase<T> h{};
        h.g = c->g;
        h.ph = c->ph;
        for (int q = 0; q < 3; ++q) {
            for (int b = 0; b < 3; ++b) h.x[q][b] = c->x[q][b]->view<T>();
            h.i2[q] = c->i2[q];
            h.rhs[q] = c->rhs[q]->view<T>();
        }
        h.mu = c->mu->view<T>();
        h.pres = c->pres->view<T>();
        h.mom[0] = c->prm.mom_u;
        h.mom[1] = c->prm.mom_v;
        h.mom[2] = c->prm.mom_w;
        h.tl = c->tiling;
        const bool dg = c->prm.diagnostics != 0;
        diag = diag || dg;
        h.part = dg ? c->red_part : nullptr;
        h.out = c->red_out + 1;
        hv.push_back(h);
        SetupGroupCase<T> su{};
        su.g = c->g;
        su.ph = c->ph;
        su.u3 = c->x[0][c->fin[0]]->view<T>();
        su.v3 = c->x[1][c->fin[1]]->view<T>();
        su.w3 = c->x[2][c->fin[2]]->view<T>();
        su.inlet = static_cast<const T*>(c->inlet->ptr);
        su.rhs_p = c->cg_r->view<double>();
        su.r = c->cg_r->view<double>();
        su.x = c->pold->view<double>();
        su.part = c->red_part;
        su.st = c->cg;
        su.rtol = c->prm.ksp_rtol;
        su.abstol = c->prm.ksp_abstol;
        su.dtol = c->prm.ksp_dtol;
        su.maxits = c->prm.ksp_maxits;
        sv.push_back(su);
    };
    if (mode == 2) {
        for (tpls_case* c : cs) {
            if (c->prec == TPLS_FP32) fill(c, h32, s32, 0.0f);
            else fill(c, h64, s64, 0.0);
        }
        RC(hh_group_loops<double>(G->hh, h64, diag) ? fail(TPLS_ERR_HIP, "grouped Helmholtz loops") : TPLS_OK);
        RC(hh_group_loops<float>(G->hh, h32, diag) ? fail(TPLS_ERR_HIP, "grouped Helmholtz loops") : TPLS_OK);
        cg_group_setup<double>(G->hh, s64);
        cg_group_setup<float>(G->hh, s32);
        LCHK(hipStreamSynchronize(G->s));  // the previous step's tables are rewritten
        LCHK(G->hh.upload());
    }
    for (tpls_case* c : cs) {
        LCHK(hipEventRecord(c->ps_ev, c->s));
        LCHK(hipStreamWaitEvent(G->s, c->ps_ev, 0));
    }
    LCHK(G->hh.run(G->s));
    LCHK(G->vc.run(G->s));
    long first = 1, launched = 0, maxits = 1;
    for (tpls_case* c : cs) {
        first = std::max<long>(first, c->use_mg ? std::max(4, c->last_its) : std::max(8, c->last_its - 2));
        maxits = std::max<long>(maxits, c->prm.ksp_maxits);
    }
    int cur = 0;
    for (long chunk = first;;) {
        chunk = std::min<long>(chunk, std::max<long>(1, maxits - launched));
        for (long q = 0; q < chunk; ++q) {
            LCHK(G->it[cur].run(G->s));
            cur ^= 1;
        }
        launched += chunk;
        LCHK(G->gather.run(G->s));
        LCHK(hipMemcpyAsync(G->h_st, G->d_st, sizeof(CgState) * G->n, hipMemcpyDeviceToHost, G->s));
        LCHK(hipEventRecord(G->ev, G->s));
        LCHK(hipEventSynchronize(G->ev));
        bool all = true;
        for (int q = 0; q < G->n; ++q)
            if (!G->h_st[q].done && launched < cs[q]->prm.ksp_maxits) all = false;
        if (all) break;
        chunk = G->any_mg ? 2 : 16;
    }
    LCHK(hipEventRecord(G->done, G->s));
    for (int q = 0; q < G->n; ++q) {
        tpls_case* c = cs[q];
        *c->h_cg = G->h_st[q];
        c->diag.ksp_its = G->h_st[q].its;
        c->diag.ksp_reason = G->h_st[q].reason;
        c->diag.ksp_rnorm = G->h_st[q].dp;
        c->last_its = G->h_st[q].its;
        LCHK(hipStreamWaitEvent(c->s, G->done, 0));
    }
    return TPLS_OK;
}

}  // namespace

// ---------------------------------------------------------------------------
// Per-routine drop-in entry points: host arrays in the reference's shapes are
// staged through per-thread device buffers, the same kernels run, results are
// copied back.  Arrays a routine only partly writes are uploaded first so
// that the elements the reference leaves untouched keep the caller's values.
// ---------------------------------------------------------------------------
namespace {

struct DropCtx {
    hipStream_t s = nullptr;
    std::vector<void*> buf;
    std::vector<size_t> cap;
    double* red_part = nullptr;
    double* red_out = nullptr;
    double* h_out = nullptr;
    // device copies of attached host arrays already bound in the current call
    std::vector<void*> mused;
    // cached device case for tpls_pressure_solve
    tpls_case* pcase = nullptr;
    int pl = 0, pm = 0, pn = 0;
    ~DropCtx()
    {
        if (pcase) tpls_case_destroy(pcase);
        for (void* p : buf) hipFree(p);
        hipFree(red_part);
        hipFree(red_out);
        if (h_out) hipHostFree(h_out);
        if (s) hipStreamDestroy(s);
    }
};

thread_local DropCtx g_drop;

int drop_init()
{
    DropCtx& d = g_drop;
    d.mused.clear();
    if (d.s) return TPLS_OK;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return fail(TPLS_ERR_NOGPU, "no HIP device");
    HIPCHK(hipStreamCreateWithFlags(&d.s, hipStreamNonBlocking));
    HIPCHK(hipMalloc(&d.red_part, sizeof(double) * 3 * REDUCE_BLOCKS));
    HIPCHK(hipMalloc(&d.red_out, sizeof(double) * 8));
    HIPCHK(hipHostMalloc(&d.h_out, sizeof(double) * 8));
    return TPLS_OK;
}

// A staged array: contiguous device copy of a host Fortran array.  For a host
// array attached to the mirror the copy is the mirror's own (mirrored), which
// holds the array between calls; otherwise a per-thread staging buffer.
struct Stg {
    DevArray a;
    void* host = nullptr;
    bool out = false;
    int esz = 8;
    bool mirrored = false;
};

// mode: 0 = upload, 1 = zero-fill (whole array written by the routine)
int stage(int slot, const void* host, const Shape& sh, int esz, int mode, bool out, Stg& st)
{
    DropCtx& d = g_drop;
    const size_t nbytes = size_t(sh.count()) * esz;
    st.mirrored = false;
    void* dv = host && mirror::any() ? mirror::device_of(host, nbytes) : nullptr;
    const bool again = dv && std::find(d.mused.begin(), d.mused.end(), dv) != d.mused.end();
    if (dv && !again) {
        // an attached array, bound once per call: the kernels use its device
        // copy, made current (or zeroed, for an array the routine overwrites)
        if (mode == 0) RC(mirror::to_device(host, nbytes));
        else HIPCHK(hipMemsetAsync(dv, 0, nbytes, d.s));
        d.mused.push_back(dv);
        st.a.ptr = dv;
        st.a.sh = sh;
        st.a.pitch = sh.nx();
        st.a.esz = esz;
        st.host = const_cast<void*>(host);
        st.out = out;
        st.esz = esz;
        st.mirrored = true;
        return TPLS_OK;
    }
    if (again && mode != 0) return fail(TPLS_ERR_ARG, "an attached array bound twice in one call must be read twice");
    if ((int)d.buf.size() <= slot) {
        d.buf.resize(slot + 1, nullptr);
        d.cap.resize(slot + 1, 0);
    }
    const size_t bytes = size_t(sh.count()) * esz;
    if (d.cap[slot] < bytes) {
        if (d.buf[slot]) HIPCHK(hipFree(d.buf[slot]));
        d.buf[slot] = nullptr;
        d.cap[slot] = 0;
        if (hipMalloc(&d.buf[slot], bytes) != hipSuccess) return fail(TPLS_ERR_NOMEM, "staging allocation");
        d.cap[slot] = bytes;
    }
    st.a.ptr = d.buf[slot];
    st.a.sh = sh;
    st.a.pitch = sh.nx();
    st.a.esz = esz;
    st.host = const_cast<void*>(host);
    st.out = out;
    st.esz = esz;
    if (again) {
        // the second binding of an attached array in one call (a routine that
        // reads and rewrites it through two buffers): a private device copy
        HIPCHK(hipMemcpyAsync(st.a.ptr, dv, bytes, hipMemcpyDeviceToDevice, d.s));
    } else if (mode == 0) {
        // host memory that may share pages with attached arrays is made
        // current and readable before the copy engine reads it
        if (host && mirror::any()) RC(mirror::host_access(host, bytes, false));
        HIPCHK(hipMemcpyAsync(st.a.ptr, host, bytes, hipMemcpyHostToDevice, d.s));
    } else {
        HIPCHK(hipMemsetAsync(st.a.ptr, 0, bytes, d.s));
    }
    return TPLS_OK;
}

// Results back to the host arrays: copied down, or, for attached arrays,
// left in (or copied into) the mirror's device copy, whose host pages then
// fetch on first access.
int unstage(std::initializer_list<Stg*> outs)
{
    DropCtx& d = g_drop;
    Stg* kept[16];
    int nkept = 0;
    for (Stg* st : outs) {
        const size_t bytes = size_t(st->a.sh.count()) * st->esz;
        void* dv = st->mirrored ? st->a.ptr : mirror::any() ? mirror::device_of(st->host, bytes) : nullptr;
        if (dv) {
            if (nkept == 16) return fail(TPLS_ERR_ARG, "too many outputs");
            if (dv != st->a.ptr) HIPCHK(hipMemcpyAsync(dv, st->a.ptr, bytes, hipMemcpyDeviceToDevice, d.s));
            kept[nkept++] = st;
            continue;
        }
        if (mirror::any()) RC(mirror::host_access(st->host, bytes, true));
        HIPCHK(hipMemcpyAsync(st->host, st->a.ptr, bytes, hipMemcpyDeviceToHost, d.s));
    }
    HIPCHK(hipStreamSynchronize(d.s));
    for (int q = 0; q < nkept; ++q) RC(mirror::written(kept[q]->host, size_t(kept[q]->a.sh.count()) * kept[q]->esz));
    return TPLS_OK;
}

// The y-halo refresh a routine ends with (wrap_rows), on the mirror's device
// copy when the device holds data the host has not seen.
int wrap_dropin(double* h, const Shape& sh, int maxm)
{
    const size_t bytes = size_t(sh.count()) * sizeof(double);
    void* dv = mirror::any() ? mirror::device_of(h, bytes) : nullptr;
    RowPairs rp{};
    const int P = maxm - 1;
    for (int j = sh.ly; j <= sh.uy; ++j) {
        if (j >= 1 && j <= P) continue;
        if (rp.n == MAX_ROW_PAIRS) {
            dv = nullptr;
            break;
        }
        rp.dst[rp.n] = j - sh.ly;
        rp.src[rp.n] = (j < 1 ? j + P : j - P) - sh.ly;
        ++rp.n;
    }
    if (!dv || rp.n == 0 || mirror::host_current(h, bytes)) {
        wrap_rows(h, sh, maxm);
        return TPLS_OK;
    }
    RC(mirror::to_device(h, bytes));
    HIPCHK(launch_copy_rows(g_drop.s, static_cast<double*>(dv), sh.nx(), sh.ny(), sh.nz(), 0, int(sh.nx()), rp));
    HIPCHK(hipStreamSynchronize(g_drop.s));
    return mirror::written(h, bytes);
}

// A case array from, or to, a host array that may be attached: the mirror's
// device copy is copied on the device instead of over the host link.
int put_any(tpls_case* c, DevArray* a, const double* host)
{
    const size_t bytes = size_t(a->sh.count()) * sizeof(double);
    void* dv = a->esz == 8 && mirror::any() ? mirror::device_of(host, bytes) : nullptr;
    if (!dv) return put_array(c, a, host);
    RC(mirror::to_device(host, bytes));
    const size_t row = size_t(a->sh.nx()) * sizeof(double);
    HIPCHK(hipMemcpy2DAsync(a->ptr, size_t(a->pitch) * sizeof(double), dv, row, row,
                            size_t(a->sh.ny() * a->sh.nz()), hipMemcpyDeviceToDevice, c->s));
    HIPCHK(hipStreamSynchronize(c->s));
    return TPLS_OK;
}

int get_any(tpls_case* c, const DevArray* a, double* host)
{
    const size_t bytes = size_t(a->sh.count()) * sizeof(double);
    void* dv = a->esz == 8 && mirror::any() ? mirror::device_of(host, bytes) : nullptr;
    if (!dv) return get_array(c, a, host);
    const size_t row = size_t(a->sh.nx()) * sizeof(double);
    HIPCHK(hipMemcpy2DAsync(dv, row, a->ptr, size_t(a->pitch) * sizeof(double), row,
                            size_t(a->sh.ny() * a->sh.nz()), hipMemcpyDeviceToDevice, c->s));
    HIPCHK(hipStreamSynchronize(c->s));
    return mirror::written(host, bytes);
}

int check_decomp(int sx, int ex, int sy, int ey, int ex_max)
{
    if (sx != 1 || sy != 1 || ex != ex_max)
        return fail(TPLS_ERR_DECOMP, "single-rank decomposition required: sx = sy = 1, ex = ex_max = maxl-1");
    if (ey < 4 || ex < 4 || (ey % 2) != 0)
        return fail(TPLS_ERR_DECOMP, "ey = maxm-1 must be even and >= 4; ex >= 4");
    return TPLS_OK;
}

Phys phys_d(double dx, double dy, double dz, double dt)
{
    Phys p{};
    p.dx = dx;
    p.dy = dy;
    p.dz = dz;
    p.dt = dt;
    return p;
}


}  // namespace

extern "C" {

int tpls_wenoappoxconvection(const double* phi, const double* u, const double* v, const double* w, int sx, int ex,
                             int sy, int ey, int l, int n, double* dfx, double* dfy, double* dfz)
{
    RC(drop_init());
    RC(check_decomp(sx, ex, sy, ey, l - 1));
    if (!phi || !u || !v || !w || !dfx || !dfy || !dfz) return fail(TPLS_ERR_ARG, "null array");
    const int maxl = l, maxm = ey + 1, maxn = n;
    Grid g{maxl, maxm, maxn, maxm - 1};
    Stg a_phi, a_u, a_v, a_w, a_x, a_y, a_z;
    RC(stage(0, phi, {-2, maxl + 2, -2, maxm + 2, 0, maxn}, 8, 0, false, a_phi));
    RC(stage(1, u, shape_of(G_UV, maxl, maxm, maxn), 8, 0, false, a_u));
    RC(stage(2, v, shape_of(G_UV, maxl, maxm, maxn), 8, 0, false, a_v));
    RC(stage(3, w, shape_of(G_W, maxl, maxm, maxn), 8, 0, false, a_w));
    RC(stage(4, dfx, shape_of(G_UV, maxl, maxm, maxn), 8, 0, true, a_x));
    RC(stage(5, dfy, shape_of(G_UV, maxl, maxm, maxn), 8, 0, true, a_y));
    RC(stage(6, dfz, shape_of(G_W, maxl, maxm, maxn), 8, 0, true, a_z));
    HIPCHK(launch_weno_faces<double>(g_drop.s, g, a_phi.a.view<double>(), a_u.a.view<double>(), a_v.a.view<double>(),
                                     a_w.a.view<double>(), a_x.a.view<double>(), a_y.a.view<double>(),
                                     a_z.a.view<double>()));
    return unstage({&a_x, &a_y, &a_z});
}

int tpls_do_advect_phi(double* phi2, double* u2, double* v2, double* w2, double* conv2_phi, int sx, int ex, int sy,
                       int ey, int maxl, int maxn, int ex_max, double dx, double dy, double dz)
{
    RC(drop_init());
    RC(check_decomp(sx, ex, sy, ey, ex_max));
    if (ex_max != maxl - 1) return fail(TPLS_ERR_DECOMP, "ex_max must be maxl-1");
    if (!phi2 || !u2 || !v2 || !w2 || !conv2_phi) return fail(TPLS_ERR_ARG, "null array");
    const int maxm = ey + 1;
    Grid g{maxl, maxm, maxn, maxm - 1};
    Stg a_phi, a_u, a_v, a_w, a_c;
    RC(stage(0, phi2, shape_of(G_P, maxl, maxm, maxn), 8, 0, false, a_phi));
    RC(stage(1, u2, shape_of(G_UV, maxl, maxm, maxn), 8, 0, false, a_u));
    RC(stage(2, v2, shape_of(G_UV, maxl, maxm, maxn), 8, 0, false, a_v));
    RC(stage(3, w2, shape_of(G_W, maxl, maxm, maxn), 8, 0, false, a_w));
    RC(stage(4, conv2_phi, shape_of(G_P, maxl, maxm, maxn), 8, 1, true, a_c));
    HIPCHK(launch_advect_phi<double>(g_drop.s, g, phys_d(dx, dy, dz, 0), a_phi.a.view<double>(), a_u.a.view<double>(),
                                     a_v.a.view<double>(), a_w.a.view<double>(), a_c.a.view<double>()));
    RC(unstage({&a_c}));
    // advect_phi.f90:317-322 refresh the y halos of u2, v2, w2 in place
    RC(wrap_dropin(u2, shape_of(G_UV, maxl, maxm, maxn), maxm));
    RC(wrap_dropin(v2, shape_of(G_UV, maxl, maxm, maxn), maxm));
    RC(wrap_dropin(w2, shape_of(G_W, maxl, maxm, maxn), maxm));
    return TPLS_OK;
}

int tpls_do_tpls(double* phi2, double* phi3, double* c0, double* c1, double* c2, int sx, int ex, int sy, int ey,
                 int maxn, int ex_max, int max_it, double height, double* err2, double dx, double dz, double dt)
{
    RC(drop_init());
    RC(check_decomp(sx, ex, sy, ey, ex_max));
    if (!phi2 || !phi3 || !c0 || !c1 || !c2 || !err2) return fail(TPLS_ERR_ARG, "null array");
    const int maxl = ex_max + 1, maxm = ey + 1;
    Grid g{maxl, maxm, maxn, maxm - 1};
    const Shape sp = shape_of(G_P, maxl, maxm, maxn);
    Stg a2, a3, b0, b1, b2, aD, aS, aI, ra, rb;
    RC(stage(0, phi2, sp, 8, 0, false, a2));
    RC(stage(1, phi3, sp, 8, 0, true, a3));
    RC(stage(2, c0, sp, 8, 0, false, b0));
    RC(stage(3, c1, sp, 8, 0, false, b1));
    RC(stage(4, c2, sp, 8, 0, false, b2));
    RC(stage(5, nullptr, sp, 8, 1, false, aD));
    RC(stage(6, nullptr, sp, 8, 1, false, aS));
    RC(stage(7, nullptr, sp, 4, 1, false, aI));
    RC(stage(8, nullptr, sp, 8, 1, false, ra));
    RC(stage(9, nullptr, sp, 8, 1, false, rb));
    hipStream_t s = g_drop.s;
    Phys ph = phys_d(dx, dx, dz, dt);
    ph.height = height;
    HIPCHK(launch_tpls_update<double>(s, g, ph, a2.a.view<double>(), b0.a.view<double>(), b1.a.view<double>(),
                                      b2.a.view<double>(), a3.a.view<double>()));
    HIPCHK(launch_phi_bc<double>(s, g, dz, height, a3.a.view<double>(), 1));
    HIPCHK(launch_levelset_status<double>(s, g, dx, a3.a.view<double>(), aD.a.view<double>(), aI.a.view<int>(),
                                          aS.a.view<double>()));
    Stg* src = &a3;
    Stg* bufs[2] = {&ra, &rb};
    Reduce red{g_drop.red_part, g_drop.red_out};
    for (int it = 1; it <= max_it; ++it) {
        Stg* dst = bufs[(it - 1) & 1];
        HIPCHK(launch_levelset_iter<double>(s, g, dx, a3.a.view<double>(), aD.a.view<double>(), aI.a.view<int>(),
                                            aS.a.view<double>(), src->a.view<double>(), dst->a.view<double>()));
        HIPCHK(launch_phi_bc<double>(s, g, dx, height, dst->a.view<double>(), 1));
        if (it == max_it) HIPCHK(launch_levelset_err<double>(s, g, src->a.view<double>(), dst->a.view<double>(), red));
        src = dst;
    }
    if (max_it > 0) {
        src->host = phi3;
        src->out = true;
        RC(unstage({src}));
        HIPCHK(hipMemcpyAsync(g_drop.h_out, g_drop.red_out, sizeof(double), hipMemcpyDeviceToHost, s));
        HIPCHK(hipStreamSynchronize(s));
        *err2 = g_drop.h_out[0];
    } else {
        RC(unstage({&a3}));
    }
    RC(wrap_dropin(phi3, sp, maxm));
    return TPLS_OK;
}

int tpls_dim(int maxl, int maxn, double dx, double dy, double dz, double dt, double epn, double Pe, const double* phi,
             const double* phiOld, const double* conv2_phi, double* ConvPrevPhi, double* phiNew, int sx, int ex, int sy,
             int ey, double height, int max_iteration_dim, double* res)
{
    RC(drop_init());
    RC(check_decomp(sx, ex, sy, ey, maxl - 1));
    if (!phi || !phiOld || !conv2_phi || !ConvPrevPhi || !phiNew) return fail(TPLS_ERR_ARG, "null array");
    if (max_iteration_dim < 0 || !(epn > 0.0) || !(Pe > 0.0))
        return fail(TPLS_ERR_ARG, "dim needs epn > 0, Pe > 0 and max_iteration_dim >= 0");
    const int maxm = ey + 1;
    Grid g{maxl, maxm, maxn, maxm - 1};
    const Shape sp = shape_of(G_P, maxl, maxm, maxn);
    std::vector<double> prof(maxn + 1);
    dim_inlet(maxn, dz, height, epn, prof.data());
    DimCoef dc;
    dim_coefficients(dx, dy, dz, dt, epn, Pe, &dc);
    Stg a_phi, a_po, a_c2, a_cp, a_pn, a_pn2, a_d2, a_src, a_prof;
    RC(stage(0, phi, sp, 8, 0, false, a_phi));
    RC(stage(1, phiOld, sp, 8, 0, false, a_po));
    RC(stage(2, conv2_phi, sp, 8, 0, false, a_c2));
    RC(stage(3, ConvPrevPhi, sp, 8, 0, true, a_cp));
    // the fused sweeps alternate between two copies of the initial iterate
    RC(stage(4, phiNew, sp, 8, 0, false, a_pn));
    RC(stage(8, phiNew, sp, 8, 0, false, a_pn2));
    RC(stage(5, nullptr, sp, 8, 1, false, a_d2));
    RC(stage(6, nullptr, sp, 8, 1, false, a_src));
    RC(stage(7, prof.data(), Shape{0, maxn, 0, 0, 0, 0}, 8, 0, false, a_prof));
    hipStream_t s = g_drop.s;
    HIPCHK(launch_dim_core<double>(s, g, phys_d(dx, dy, dz, dt), dc, max_iteration_dim, 0.0, a_phi.a.view<double>(),
                                   a_po.a.view<double>(), a_c2.a.view<double>(), a_cp.a.view<double>(),
                                   a_pn.a.view<double>(), a_pn2.a.view<double>(), a_d2.a.view<double>(),
                                   a_src.a.view<double>(), static_cast<const double*>(a_prof.a.ptr),
                                   res ? g_drop.red_out : nullptr, 1));
    Stg* last = (max_iteration_dim & 1) ? &a_pn2 : &a_pn;
    last->host = phiNew;
    last->out = true;
    RC(unstage({&a_cp, last}));
    if (res) {
        HIPCHK(hipMemcpyAsync(g_drop.h_out, g_drop.red_out, sizeof(double), hipMemcpyDeviceToHost, s));
        HIPCHK(hipStreamSynchronize(s));
        *res = g_drop.h_out[0];
    }
    // cahn_hilliard_solver.f90:340-341 refresh phiNew's y halos (the inlet and
    // outlet columns of those rows are set from the interior rows here)
    RC(wrap_dropin(phiNew, sp, maxm));
    return TPLS_OK;
}

int tpls_get_levelset_status(const double* phi, double* D, int* ist, double* sg, int sx, int ex, int sy, int ey,
                             int n, double h)
{
    RC(drop_init());
    RC(check_decomp(sx, ex, sy, ey, ex));
    if (!phi || !D || !ist || !sg) return fail(TPLS_ERR_ARG, "null array");
    const int maxl = ex + 1, maxm = ey + 1;
    Grid g{maxl, maxm, n, maxm - 1};
    const Shape sp = shape_of(G_P, maxl, maxm, n);
    Stg a, aD, aI, aS;
    RC(stage(0, phi, sp, 8, 0, false, a));
    RC(stage(1, D, sp, 8, 1, true, aD));
    RC(stage(2, ist, sp, 4, 1, true, aI));
    RC(stage(3, sg, sp, 8, 1, true, aS));
    HIPCHK(launch_levelset_status<double>(g_drop.s, g, h, a.a.view<double>(), aD.a.view<double>(), aI.a.view<int>(),
                                          aS.a.view<double>()));
    return unstage({&aD, &aI, &aS});
}

int tpls_do_levelset_iteration(const double* phi, double* pr, const double* D, const int* ist, const double* sg,
                               int sx, int ex, int sy, int ey, int n, int ex_max, double h, double height,
                               double* error_val)
{
    RC(drop_init());
    RC(check_decomp(sx, ex, sy, ey, ex_max));
    if (!phi || !pr || !D || !ist || !sg || !error_val) return fail(TPLS_ERR_ARG, "null array");
    const int maxl = ex + 1, maxm = ey + 1;
    Grid g{maxl, maxm, n, maxm - 1};
    const Shape sp = shape_of(G_P, maxl, maxm, n);
    Stg a, r0, r1, aD, aI, aS;
    RC(stage(0, phi, sp, 8, 0, false, a));
    RC(stage(1, pr, sp, 8, 0, false, r0));
    RC(stage(2, pr, sp, 8, 0, true, r1));
    RC(stage(3, D, sp, 8, 0, false, aD));
    RC(stage(4, ist, sp, 4, 0, false, aI));
    RC(stage(5, sg, sp, 8, 0, false, aS));
    hipStream_t s = g_drop.s;
    HIPCHK(launch_levelset_iter<double>(s, g, h, a.a.view<double>(), aD.a.view<double>(), aI.a.view<int>(),
                                        aS.a.view<double>(), r0.a.view<double>(), r1.a.view<double>()));
    HIPCHK(launch_phi_bc<double>(s, g, h, height, r1.a.view<double>(), 1));
    Reduce red{g_drop.red_part, g_drop.red_out};
    HIPCHK(launch_levelset_err<double>(s, g, r0.a.view<double>(), r1.a.view<double>(), red));
    RC(unstage({&r1}));
    HIPCHK(hipMemcpyAsync(g_drop.h_out, g_drop.red_out, sizeof(double), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    *error_val = g_drop.h_out[0];
    return TPLS_OK;
}

int tpls_get_csf(double* fx, double* fy, double* fz, const double* fi, int sx, int ex, int sy, int ey, int n,
                 int ex_max, double dx, double dy, double dz, double scap, double sw)
{
    RC(drop_init());
    RC(check_decomp(sx, ex, sy, ey, ex_max));
    if (!fx || !fy || !fz || !fi) return fail(TPLS_ERR_ARG, "null array");
    const int maxl = ex + 1, maxm = ey + 1;
    Grid g{maxl, maxm, n, maxm - 1};
    const Shape sp = shape_of(G_P, maxl, maxm, n);
    Stg a, bx, by, bz;
    RC(stage(0, fi, sp, 8, 0, false, a));
    RC(stage(1, fx, sp, 8, 0, true, bx));
    RC(stage(2, fy, sp, 8, 0, true, by));
    RC(stage(3, fz, sp, 8, 0, true, bz));
    Phys ph = phys_d(dx, dy, dz, 0);
    ph.scap = scap;
    ph.sw = sw;
    HIPCHK(launch_csf<double>(g_drop.s, g, ph, a.a.view<double>(), bx.a.view<double>(), by.a.view<double>(),
                              bz.a.view<double>()));
    HIPCHK(launch_csf_bc<double>(g_drop.s, g, bx.a.view<double>(), by.a.view<double>(), bz.a.view<double>()));
    return unstage({&bx, &by, &bz});
}

static int momentum_dropin(int mode, double* o_u, double* o_v, double* o_w, const double* visc, const double* u2,
                           const double* v2, const double* w2, int sx, int sy, int ex, int ey, int maxn, int ex_max,
                           double dx, double dy, double dz)
{
    RC(drop_init());
    RC(check_decomp(sx, ex, sy, ey, ex_max));
    if (!o_u || !o_v || !o_w || !visc || !u2 || !v2 || !w2) return fail(TPLS_ERR_ARG, "null array");
    const int maxl = ex + 1, maxm = ey + 1;
    Grid g{maxl, maxm, maxn, maxm - 1};
    const Shape suv = shape_of(G_UV, maxl, maxm, maxn), sw = shape_of(G_W, maxl, maxm, maxn);
    Stg au, av, aw, amu, ou, ov, ow;
    RC(stage(0, u2, suv, 8, 0, false, au));
    RC(stage(1, v2, suv, 8, 0, false, av));
    RC(stage(2, w2, sw, 8, 0, false, aw));
    RC(stage(3, visc, shape_of(G_AUG, maxl, maxm, maxn), 8, 0, false, amu));
    // the reference zeroes the whole output array first
    RC(stage(4, o_u, suv, 8, 1, true, ou));
    RC(stage(5, o_v, suv, 8, 1, true, ov));
    RC(stage(6, o_w, sw, 8, 1, true, ow));
    Phys ph = phys_d(dx, dy, dz, 0);
    Stg* outs[3] = {&ou, &ov, &ow};
    for (int q = 0; q < 3; ++q) {
        FieldView<double> o = outs[q]->a.view<double>();
        HIPCHK(launch_momentum<double>(g_drop.s, g, ph, q, mode, au.a.view<double>(), av.a.view<double>(),
                                       aw.a.view<double>(), amu.a.view<double>(), o, o, o, o, o, o, o, o, o));
    }
    return unstage({&ou, &ov, &ow});
}

int tpls_get_diffusion_all(double* du, double* dv, double* dw, const double* visc, const double* u2, const double* v2,
                           const double* w2, int sx, int sy, int ex, int ey, int maxn, int ex_max, double dx,
                           double dy, double dz)
{
    return momentum_dropin(2, du, dv, dw, visc, u2, v2, w2, sx, sy, ex, ey, maxn, ex_max, dx, dy, dz);
}

int tpls_get_conv_all(double* cu, double* cv, double* cw, const double* visc, const double* u2, const double* v2,
                      const double* w2, int sx, int sy, int ex, int ey, int maxn, int ex_max, double dx, double dy,
                      double dz)
{
    return momentum_dropin(1, cu, cv, cw, visc, u2, v2, w2, sx, sy, ex, ey, maxn, ex_max, dx, dy, dz);
}

int tpls_get_viscosity(double* visc, const double* phi, int ex, int ey, int sx, int sy, int maxn, double sw, double Re,
                       double mu_minus, double mu_plus)
{
    RC(drop_init());
    RC(check_decomp(sx, ex, sy, ey, ex));
    if (!visc || !phi) return fail(TPLS_ERR_ARG, "null array");
    const int maxl = ex + 1, maxm = ey + 1;
    Grid g{maxl, maxm, maxn, maxm - 1};
    const Shape sp = shape_of(G_P, maxl, maxm, maxn);
    Stg a, m;
    RC(stage(0, phi, sp, 8, 0, false, a));
    RC(stage(1, visc, sp, 8, 0, true, m));
    Phys ph{};
    ph.sw = sw;
    ph.Re = Re;
    ph.mu_minus = mu_minus;
    ph.mu_plus = mu_plus;
    HIPCHK(launch_viscosity<double>(g_drop.s, g, ph, a.a.view<double>(), m.a.view<double>(), 1, ex));
    return unstage({&m});
}

static int jacobi_dropin(int comp, double* x, const double* xo, const double* R, const double* visc, double dx,
                         double dy, double dz, double dt, int ex, int ey, int sx, int sy, int maxn, int ex_max)
{
    RC(drop_init());
    RC(check_decomp(sx, ex, sy, ey, ex_max));
    if (!x || !xo || !R || !visc) return fail(TPLS_ERR_ARG, "null array");
    const int maxl = ex + 1, maxm = ey + 1;
    Grid g{maxl, maxm, maxn, maxm - 1};
    const Shape sh = shape_of(comp == 2 ? G_W : G_UV, maxl, maxm, maxn);
    Stg a, ao, ar, am;
    RC(stage(0, x, sh, 8, 0, true, a));
    RC(stage(1, xo, sh, 8, 0, false, ao));
    RC(stage(2, R, sh, 8, 0, false, ar));
    RC(stage(3, visc, shape_of(G_AUG, maxl, maxm, maxn), 8, 0, false, am));
    Phys ph = phys_d(dx, dy, dz, dt);
    HIPCHK(launch_jacobi<double>(g_drop.s, g, ph, comp, ao.a.view<double>(), ar.a.view<double>(), am.a.view<double>(),
                                 a.a.view<double>(), false));
    // do_jacobi_v / do_jacobi_w close with their x and wall conditions
    // (jacobi_iteration_allflux.f90:178-189, 257-279); do_jacobi_u does not
    if (comp != 0)
        HIPCHK(launch_helmholtz_bc<double>(g_drop.s, g, ph, comp, a.a.view<double>(), a.a.view<double>()));
    return unstage({&a});
}

int tpls_do_jacobi_u(double* u3, const double* uo, const double* R, const double* visc, double dx, double dy,
                     double dz, double dt, int ex, int ey, int sx, int sy, int maxn, int ex_max)
{
    return jacobi_dropin(0, u3, uo, R, visc, dx, dy, dz, dt, ex, ey, sx, sy, maxn, ex_max);
}
int tpls_do_jacobi_v(double* v3, const double* vo, const double* R, const double* visc, double dx, double dy,
                     double dz, double dt, int ex, int ey, int sx, int sy, int maxn, int ex_max)
{
    return jacobi_dropin(1, v3, vo, R, visc, dx, dy, dz, dt, ex, ey, sx, sy, maxn, ex_max);
}
int tpls_do_jacobi_w(double* w3, const double* wo, const double* R, const double* visc, double dx, double dy,
                     double dz, double dt, int ex, int ey, int sx, int sy, int maxn, int ex_max)
{
    return jacobi_dropin(2, w3, wo, R, visc, dx, dy, dz, dt, ex, ey, sx, sy, maxn, ex_max);
}

int tpls_get_difference(const double* ua, const double* ub, int ex, int ey, int ez, int sx, int sy, int sz,
                        double* diff)
{
    RC(drop_init());
    if (!ua || !ub || !diff) return fail(TPLS_ERR_ARG, "null array");
    if (sx != 1 || sy != 1 || sz != 0) return fail(TPLS_ERR_DECOMP, "single-rank decomposition required");
    const int maxl = ex + 1, maxm = ey + 1;
    Grid g{maxl, maxm, ez + 2, maxm - 1};
    const Shape sh{0, maxl, 0, maxm, 0, ez};
    Stg a, b;
    RC(stage(0, ua, sh, 8, 0, false, a));
    RC(stage(1, ub, sh, 8, 0, false, b));
    Reduce red{g_drop.red_part, g_drop.red_out};
    HIPCHK(launch_difference<double>(g_drop.s, g, ez + 1, a.a.view<double>(), b.a.view<double>(), red));
    HIPCHK(hipMemcpyAsync(g_drop.h_out, g_drop.red_out, sizeof(double), hipMemcpyDeviceToHost, g_drop.s));
    HIPCHK(hipStreamSynchronize(g_drop.s));
    *diff = g_drop.h_out[0];
    return TPLS_OK;
}

static int sor_dropin(int comp, double* x, const double* R, const double* visc, double dx, double dy, double dz,
                      double dt, int ex, int ey, int sx, int sy, int maxn, int ex_max, int flg, int it)
{
    RC(drop_init());
    RC(check_decomp(sx, ex, sy, ey, ex_max));
    if (!x || !R || !visc) return fail(TPLS_ERR_ARG, "null array");
    const int maxl = ex + 1, maxm = ey + 1;
    Grid g{maxl, maxm, maxn, maxm - 1};
    const Shape sh = shape_of(comp == 2 ? G_W : G_UV, maxl, maxm, maxn);
    Stg a, ar, am;
    RC(stage(0, x, sh, 8, 0, true, a));
    RC(stage(1, R, sh, 8, 0, false, ar));
    RC(stage(2, visc, shape_of(G_AUG, maxl, maxm, maxn), 8, 0, false, am));
    HIPCHK(launch_sor<double>(g_drop.s, g, phys_d(dx, dy, dz, dt), comp, a.a.view<double>(), ar.a.view<double>(),
                              am.a.view<double>(), flg, it));
    return unstage({&a});
}

int tpls_do_sor_u(double* u3, const double* R, const double* visc, double dx, double dy, double dz, double dt, int ex,
                  int ey, int sx, int sy, int maxn, int ex_max, int flg, int it)
{
    return sor_dropin(0, u3, R, visc, dx, dy, dz, dt, ex, ey, sx, sy, maxn, ex_max, flg, it);
}
int tpls_do_sor_v(double* v3, const double* R, const double* visc, double dx, double dy, double dz, double dt, int ex,
                  int ey, int sx, int sy, int maxn, int flg, int it)
{
    return sor_dropin(1, v3, R, visc, dx, dy, dz, dt, ex, ey, sx, sy, maxn, ex, flg, it);
}
int tpls_do_sor_w(double* w3, const double* R, const double* visc, double dx, double dy, double dz, double dt, int ex,
                  int ey, int sx, int sy, int maxn, int flg, int it)
{
    return sor_dropin(2, w3, R, visc, dx, dy, dz, dt, ex, ey, sx, sy, maxn, ex, flg, it);
}

int tpls_get_source_pres(double* rhs, const double* u3, const double* v3, const double* w3, int ex, int ey, int sx,
                         int sy, int maxn, double dx, double dy, double dz, double dt)
{
    RC(drop_init());
    RC(check_decomp(sx, ex, sy, ey, ex));
    if (!rhs || !u3 || !v3 || !w3) return fail(TPLS_ERR_ARG, "null array");
    const int maxl = ex + 1, maxm = ey + 1;
    Grid g{maxl, maxm, maxn, maxm - 1};
    Stg r, au, av, aw;
    RC(stage(0, rhs, shape_of(G_P, maxl, maxm, maxn), 8, 1, true, r));
    RC(stage(1, u3, shape_of(G_UV, maxl, maxm, maxn), 8, 0, false, au));
    RC(stage(2, v3, shape_of(G_UV, maxl, maxm, maxn), 8, 0, false, av));
    RC(stage(3, w3, shape_of(G_W, maxl, maxm, maxn), 8, 0, false, aw));
    HIPCHK(launch_source_pres<double>(g_drop.s, g, phys_d(dx, dy, dz, dt), au.a.view<double>(), av.a.view<double>(),
                                      aw.a.view<double>(), r.a.view<double>()));
    return unstage({&r});
}

int tpls_get_uvw_pres(double* u3, double* v3, double* w3, const double* p, int ex, int ey, int sx, int sy, int maxn,
                      int ex_max, double dx, double dy, double dz, double dt, const double* u_inlet)
{
    RC(drop_init());
    RC(check_decomp(sx, ex, sy, ey, ex_max));
    if (!u3 || !v3 || !w3 || !p || !u_inlet) return fail(TPLS_ERR_ARG, "null array");
    const int maxl = ex + 1, maxm = ey + 1;
    Grid g{maxl, maxm, maxn, maxm - 1};
    Stg au, av, aw, ap, ai;
    RC(stage(0, u3, shape_of(G_UV, maxl, maxm, maxn), 8, 0, true, au));
    RC(stage(1, v3, shape_of(G_UV, maxl, maxm, maxn), 8, 0, true, av));
    RC(stage(2, w3, shape_of(G_W, maxl, maxm, maxn), 8, 0, true, aw));
    RC(stage(3, p, shape_of(G_P, maxl, maxm, maxn), 8, 0, false, ap));
    RC(stage(4, u_inlet, shape_of(G_INLET, maxl, maxm, maxn), 8, 0, false, ai));
    HIPCHK(launch_uvw_pres<double>(g_drop.s, g, phys_d(dx, dy, dz, dt), ap.a.view<double>(),
                                   static_cast<const double*>(ai.a.ptr), au.a.view<double>(), av.a.view<double>(),
                                   aw.a.view<double>()));
    return unstage({&au, &av, &aw});
}

int tpls_pressure_solve(double* pres, const double* RHS_p, const double* u3, const double* u_inlet, int maxl, int maxm,
                        int maxn, double dx, double dy, double dz, double dt, double rtol, double abstol, double dtol,
                        int maxits, int* its, double* rnorm, int* reason)
{
    RC(drop_init());
    if (!pres || !RHS_p || !u3 || !u_inlet || !its || !rnorm || !reason) return fail(TPLS_ERR_ARG, "null argument");
    DropCtx& d = g_drop;
    if (!d.pcase || d.pl != maxl || d.pm != maxm || d.pn != maxn) {
        if (d.pcase) tpls_case_destroy(d.pcase);
        d.pcase = nullptr;
        tpls_params prm{};
        prm.dx = dx;
        prm.dy = dy;
        prm.dz = dz;
        prm.dt = dt;
        prm.Re = 1;
        prm.mom_u = prm.mom_v = prm.mom_w = 2;
        RC(tpls_case_create(&prm, maxl, maxm, maxn, &d.pcase));
        d.pl = maxl;
        d.pm = maxm;
        d.pn = maxn;
    }
    tpls_case* c = d.pcase;
    c->ph = phys_of(c->prm);
    c->ph.dx = dx;
    c->ph.dy = dy;
    c->ph.dz = dz;
    c->ph.dt = dt;
    c->prm.ksp_rtol = rtol;
    c->prm.ksp_abstol = abstol;
    c->prm.ksp_dtol = dtol;
    c->prm.ksp_maxits = maxits;
    // the case's rhs[0..2] hold u3 (uv grid) and RHS_p (p grid, double)
    RC(put_any(c, c->x[0][0], u3));
    RC(put_any(c, c->pold, pres));
    RC(put_any(c, c->cg_p0, RHS_p));
    std::vector<double> inl(u_inlet, u_inlet + (maxn - 1));
    RC(put_array(c, c->inlet, inl.data()));
    FieldView<double> u = c->x[0][0]->view<double>();
    RC(pressure_solve<double>(c, u, u, u, c->cg_p0->view<double>(), 1, static_cast<const double*>(c->inlet->ptr)));
    HIPCHK(launch_pres_ghosts<double>(c->s, c->g, c->ph, c->pold->view<double>(), u,
                                      static_cast<const double*>(c->inlet->ptr), c->pres->view<double>()));
    RC(get_any(c, c->pres, pres));
    RC(wrap_dropin(pres, shape_of(G_P, maxl, maxm, maxn), maxm));
    *its = c->diag.ksp_its;
    *rnorm = c->diag.ksp_rnorm;
    *reason = c->diag.ksp_reason;
    return TPLS_OK;
}

// Several cases advanced together: each case runs on its own stream, and
// every phase of the step is enqueued for all cases before the host waits on
// any of them, so the GPU overlaps the cases' kernels.  Each case computes
// exactly what tpls_step computes for it alone.
int tpls_batch_step_status(tpls_case** cases, int ncases, int nsteps, int* status)
{
    if (!cases || ncases < 0 || nsteps < 0) return fail(TPLS_ERR_ARG, "bad argument");
    for (int q = 0; q < ncases; ++q) {
        if (!cases[q]) return fail(TPLS_ERR_ARG, "null case");
        for (int r = 0; r < q; ++r)
            if (cases[r] == cases[q]) return fail(TPLS_ERR_ARG, "a case appears twice in the batch");
    }
    std::vector<int> st(ncases, TPLS_OK);
    std::vector<char> open(ncases);
    int first_bad = -1;
    std::string first_msg;
    // a case whose step fails stops here: its stream is drained, it is marked
    // broken, and the other cases go on with their steps
    auto drop = [&](int q, int rc) {
        tpls_case* c = cases[q];
        st[q] = rc;
        open[q] = 0;
        c->broken = 1;
        c->broken_msg = g_err;
        if (first_bad < 0) {
            first_bad = q;
            first_msg = g_err;
        }
        (void)hipStreamSynchronize(c->s);
        (void)hipGetLastError();
    };
    for (int q = 0; q < ncases; ++q)
        if (cases[q]->broken) {
            fail(TPLS_ERR_ARG, "the case failed in an earlier step: " + cases[q]->broken_msg);
            drop(q, TPLS_ERR_ARG);
        }
    int grouped = ncases > 1 ? g_batch_grouped : 0;  // one case gains nothing from grouping
    for (int q = 0; q < ncases; ++q)
        if (cases[q]->timed) grouped = 0;  // kernel timing brackets per-case launches
    for (int n = 0; n < nsteps; ++n) {
        if (grouped) {
            std::vector<tpls_case*> cs;
            std::vector<int> qs;
            for (int q = 0; q < ncases; ++q) {
                if (st[q]) continue;
                tpls_case* c = cases[q];
                c->group_solve = grouped;
                const int rc = c->prec == TPLS_FP32 ? step_pre<float>(c) : step_pre<double>(c);
                if (rc) {
                    drop(q, rc);
                    continue;
                }
                cs.push_back(c);
                qs.push_back(q);
            }
            if (cs.empty()) continue;
            int rc = group_solve(cs, grouped);
            if (rc) {
                const std::string msg = g_err;
                for (int q : qs) {
                    g_err = msg;
                    drop(q, rc);
                }
                continue;
            }
            for (size_t k = 0; k < cs.size(); ++k) {
                tpls_case* c = cs[k];
                rc = c->prec == TPLS_FP32 ? step_end<float>(c) : step_end<double>(c);
                if (rc) drop(qs[k], rc);
            }
            continue;
        }
        int left = 0;
        for (int q = 0; q < ncases; ++q) {
            if (st[q]) continue;
            tpls_case* c = cases[q];
            const int rc = c->prec == TPLS_FP32 ? step_begin<float>(c) : step_begin<double>(c);
            if (rc) {
                drop(q, rc);
                continue;
            }
            open[q] = 1;
            ++left;
        }
        // serve whichever case's CG chunk has landed, so that no finished case
        // waits behind a slower one for its next chunk; when none has, block
        // on the oldest open one
        for (int next = 0; left > 0;) {
            bool served = false;
            for (int q = 0; q < ncases; ++q) {
                if (!open[q]) continue;
                tpls_case* c = cases[q];
                const hipError_t e = hipEventQuery(c->ps_ev);
                if (e == hipErrorNotReady) continue;
                served = true;
                int rc = e == hipSuccess ? TPLS_OK : fail(TPLS_ERR_HIP, std::string("hipEventQuery: ") + hipGetErrorString(e));
                bool fin = false;
                if (!rc) rc = ps_poll(c, &fin);
                if (!rc && fin) rc = c->prec == TPLS_FP32 ? step_end<float>(c) : step_end<double>(c);
                if (rc) {
                    drop(q, rc);
                    --left;
                } else if (fin) {
                    open[q] = 0;
                    --left;
                }
            }
            if (!served && left > 0) {
                while (!open[next]) next = (next + 1) % ncases;
                const hipError_t e = hipEventSynchronize(cases[next]->ps_ev);
                if (e != hipSuccess) {
                    drop(next, fail(TPLS_ERR_HIP, std::string("hipEventSynchronize: ") + hipGetErrorString(e)));
                    --left;
                }
                next = (next + 1) % ncases;
            }
        }
    }
    for (int q = 0; q < ncases; ++q)
        if (!st[q]) {
            const int rc = finish_diag(cases[q]);
            if (rc) drop(q, rc);
        }
    if (status)
        for (int q = 0; q < ncases; ++q) status[q] = st[q];
    if (first_bad >= 0) return fail(st[first_bad], "case " + std::to_string(first_bad) + ": " + first_msg);
    return TPLS_OK;
}

int tpls_batch_step(tpls_case** cases, int ncases, int nsteps)
{
    return tpls_batch_step_status(cases, ncases, nsteps, nullptr);
}

int tpls_batch_set_grouped(int mode)
{
    if (mode < 0 || mode > 2) return fail(TPLS_ERR_ARG, "grouping mode must be 0, 1 or 2");
    g_batch_grouped = mode;
    return TPLS_OK;
}

// Test hook, not part of the public interface: the case's next step fails
// before it launches anything (host-side), so that failure isolation can be
// tested without a GPU fault.
int tpls_test_fail_next_step(tpls_case* c)
{
    if (!c) return fail(TPLS_ERR_ARG, "null case");
    c->fail_next = 1;
    return TPLS_OK;
}

// Test hook, not part of the public interface: z = M r, one V-cycle of the
// pressure solve's multigrid preconditioner on an nx x ny x nz reduced
// operator (r, z flat with x fastest, no mean removal), for comparison with
// the oracle's orc_mg_apply.  Returns the number of levels or a status < 0.
int tpls_test_mg_apply(int nx, int ny, int nz, double cx, double cy, double cz, double diag, int nu, const double* r,
                       double* z)
{
    RC(drop_init());
    if (!r || !z || nx < 1 || ny < 2 || nz < 1 || (ny % 2) || nu < 1) return fail(TPLS_ERR_ARG, "bad argument");
    MgHierarchy h{};
    h.nu = nu;
    h.nlev = mg_plan(nx, ny, nz, cx, cy, cz, diag, h.L);
    const MgLevel& top = h.L[h.nlev - 1];
    if (long(top.nx) * top.ny * top.nz > MG_COARSE_MAX) return fail(TPLS_ERR_ARG, "coarsest level too large");
    hipStream_t s = g_drop.s;
    std::vector<void*> mem;
    double* spec = nullptr;
    RC(upload_spectral(top, s, &spec));
    if (spec) mem.push_back(spec);
    h.spec = spec;
    auto release = [&]() {
        hipStreamSynchronize(s);
        for (void* p : mem) hipFree(p);
    };
    DevArray lev[MG_MAXLEV][2];
    for (int l = 0; l < h.nlev; ++l) {
        MgLevel& o = h.L[l];
        FieldView<double>* v[2] = {&o.x, &o.b};
        for (int q = 0; q < 2; ++q) {
            DevArray& a = lev[l][q];
            a.sh = Shape{0, o.nx + 1, 0, o.ny + 1, 0, o.nz + 1};
            a.esz = 8;
            a.pitch = pitch_for(a.sh.nx(), 8);
            if (hipMalloc(&a.ptr, a.bytes()) != hipSuccess) {
                release();
                return fail(TPLS_ERR_NOMEM, "hipMalloc");
            }
            mem.push_back(a.ptr);
            if (hipMemsetAsync(a.ptr, 0, a.bytes(), s) != hipSuccess) {
                release();
                return fail(TPLS_ERR_HIP, "memset");
            }
            *v[q] = a.view<double>();
        }
    }
    const Grid g{nx + 1, ny + 1, nz + 1, ny};
    const long nb = std::max<long>(cg_num_blocks(g), (long)REDUCE_BLOCKS);
    double* part = nullptr;
    CgState* st = nullptr;
    if (hipMalloc(&part, sizeof(double) * 3 * nb) == hipSuccess) mem.push_back(part);
    if (hipMalloc(&st, sizeof(CgState)) == hipSuccess) mem.push_back(st);
    if (!part || !st) {
        release();
        return fail(TPLS_ERR_NOMEM, "hipMalloc");
    }
    CgState init{};
    const DevArray& b0 = lev[0][1];
    const DevArray& x0 = lev[0][0];
    const size_t row = size_t(nx) * 8;
    int rc = TPLS_OK;
    auto plane = [&](const DevArray& a, int k) {
        return static_cast<char*>(a.ptr) + (size_t(1) + size_t(a.pitch) + size_t(k) * a.pitch * a.sh.ny()) * 8;
    };
    if (hipMemcpyAsync(st, &init, sizeof init, hipMemcpyHostToDevice, s) != hipSuccess) rc = fail(TPLS_ERR_HIP, "copy");
    for (int k = 1; k <= nz && !rc; ++k)
        if (hipMemcpy2DAsync(plane(b0, k), b0.pitch * 8, r + size_t(k - 1) * nx * ny, row, row, ny,
                             hipMemcpyHostToDevice, s) != hipSuccess)
            rc = fail(TPLS_ERR_HIP, "copy in");
    if (!rc && launch_mg_vcycle(s, g, h, Reduce{part, nullptr}, st) != hipSuccess) rc = fail(TPLS_ERR_HIP, "launch");
    for (int k = 1; k <= nz && !rc; ++k)
        if (hipMemcpy2DAsync(z + size_t(k - 1) * nx * ny, row, plane(x0, k), x0.pitch * 8, row, ny,
                             hipMemcpyDeviceToHost, s) != hipSuccess)
            rc = fail(TPLS_ERR_HIP, "copy out");
    if (!rc && hipStreamSynchronize(s) != hipSuccess) rc = fail(TPLS_ERR_HIP, "sync");
    release();
    return rc ? rc : h.nlev;
}

}  // extern "C"

// Grouped launches: one kernel launch that runs the same kernel for many
// cases at once, each case on its own blocks with its own arguments.
//
// A kernel body is written once, as Op::run(args, blk), where blk gives the
// block's coordinates within that case's own grid (and, for the grid-stride
// kernels, that grid's block count).  The plain kernel calls it with its own
// blockIdx; the grouped kernel k_group<Op> finds, for its flat block index,
// the case whose block range holds it in a device table of GroupItems and
// calls it with the case's arguments and local coordinates.  Every case
// therefore computes exactly what its own launch computes.
#pragma once

#include <hip/hip_runtime.h>

#include <cstring>
#include <functional>
#include <vector>

#include "tpls_kernels.hpp"

namespace tpls {

struct Blk {
    int x, y, z;  // block coordinates in the case's own grid
    int n;        // that grid's extent in x (the block count of a 1-D grid)
};

__device__ __forceinline__ Blk blk_here()
{
    return Blk{int(blockIdx.x), int(blockIdx.y), int(blockIdx.z), int(gridDim.x)};
}

// One case's share of a grouped launch: its kernel arguments, its first
// block in the launch and its own grid.
template <class A>
struct GroupItem {
    A a;
    int start, gx, gy, gz;
};

// Case ranges start at multiples of 8 blocks: the hardware deals blocks to
// the 8 XCDs round-robin by flat index, so a case's local block b runs on
// XCD b % 8 as in its own launch, which the kernels' XCD-aware tile order
// (xcd_tile) assumes.  Blocks of the padding return at once.
constexpr int GROUP_ALIGN = 8;

// The item whose range holds flat block b: every wavefront reads 64 of the
// sorted starts at a time (one coalesced load) and counts the ones <= b.
__device__ __forceinline__ int group_find(const int* starts, int n, int b)
{
    const int lane = int(__lane_id());
    for (int base = 0;; base += 64) {
        const int idx = base + lane;
        const bool le = idx < n && starts[idx] <= b;
        const unsigned long long m = __ballot(le);
        if (m != ~0ull || base + 64 >= n) return base + __popcll(m) - 1;
    }
}

template <class Op>
__global__ void __launch_bounds__(Op::THREADS) k_group(const GroupItem<typename Op::Args>* tab, const int* starts,
                                                       int n)
{
    const int b = int(blockIdx.x);
    const GroupItem<typename Op::Args>& it = tab[group_find(starts, n, b)];
    const int lb = b - it.start;
    if (lb >= it.gx * it.gy * it.gz) return;
    const typename Op::Args a = it.a;
    Op::run(a, Blk{lb % it.gx, (lb / it.gx) % it.gy, lb / (it.gx * it.gy), it.gx});
}

// The tables of a sequence of grouped launches, packed into one device
// buffer, and the launches themselves in order.  Built on the host once for
// a set of cases, uploaded once, run as often as needed.
class GroupPlan {
public:
    ~GroupPlan() { release(); }
    GroupPlan() = default;
    GroupPlan(const GroupPlan&) = delete;
    GroupPlan& operator=(const GroupPlan&) = delete;

    void clear()
    {
        host_.clear();
        steps_.clear();
    }

    // Append one grouped launch of Op over the items (in case order).
    template <class Op>
    void add(const std::vector<GroupItem<typename Op::Args>>& items)
    {
        if (items.empty()) return;
        using Item = GroupItem<typename Op::Args>;
        const size_t off = (host_.size() + 255) / 256 * 256;
        host_.resize(off + items.size() * sizeof(Item));
        std::memcpy(host_.data() + off, items.data(), items.size() * sizeof(Item));
        // the starts again, packed, for the lookup
        const size_t soff = (host_.size() + 255) / 256 * 256;
        host_.resize(soff + items.size() * sizeof(int));
        for (size_t q = 0; q < items.size(); ++q)
            std::memcpy(host_.data() + soff + q * sizeof(int), &items[q].start, sizeof(int));
        const int n = int(items.size());
        const int blocks = items.back().start + items.back().gx * items.back().gy * items.back().gz;
        steps_.push_back([off, soff, n, blocks](hipStream_t s, const char* dev) {
            hipLaunchKernelGGL(k_group<Op>, dim3(blocks), Op::block(), 0, s,
                               reinterpret_cast<const Item*>(dev + off), reinterpret_cast<const int*>(dev + soff), n);
            return hipGetLastError();
        });
    }

    // Copy the tables to the device (synchronously); reallocates as needed.
    hipError_t upload()
    {
        if (host_.empty()) return hipSuccess;
        if (cap_ < host_.size()) {
            release();
            hipError_t e = hipMalloc(&dev_, host_.size());
            if (e != hipSuccess) {
                dev_ = nullptr;
                return e;
            }
            cap_ = host_.size();
        }
        return hipMemcpy(dev_, host_.data(), host_.size(), hipMemcpyHostToDevice);
    }

    hipError_t run(hipStream_t s) const
    {
        for (const auto& f : steps_)
            if (hipError_t e = f(s, dev_)) return e;
        return hipSuccess;
    }

    size_t launches() const { return steps_.size(); }

private:
    void release()
    {
        if (dev_) (void)hipFree(dev_);
        dev_ = nullptr;
        cap_ = 0;
    }
    std::vector<char> host_;
    std::vector<std::function<hipError_t(hipStream_t, const char*)>> steps_;
    char* dev_ = nullptr;
    size_t cap_ = 0;
};

// Items of one grouped launch: add() appends a case with its grid.
template <class A>
struct GroupItems {
    std::vector<GroupItem<A>> v;
    int blocks = 0;
    void add(const A& a, dim3 g)
    {
        if (g.x * g.y * g.z == 0) return;
        blocks = (blocks + GROUP_ALIGN - 1) / GROUP_ALIGN * GROUP_ALIGN;
        v.push_back(GroupItem<A>{a, blocks, int(g.x), int(g.y), int(g.z)});
        blocks += int(g.x * g.y * g.z);
    }
};

// ---- the grouped pressure solve ----
// One case of a grouped V-cycle: its hierarchy, its level-0 grid, the
// partials buffer of its reductions and its solver state.
struct MgGroupCase {
    const MgHierarchy* h;
    Grid g;
    double* part;
    CgState* st;
};
// Append the launches of one V-cycle (launch_mg_vcycle) for every case, in
// lock-step: each kernel and level once across the cases of one precision
// and sweep count.  0, or -1 for a hierarchy launch_mg_vcycle would refuse.
int mg_group_vcycle(GroupPlan& P, const std::vector<MgGroupCase>& cs);

// One case of a grouped CG iteration (launch_cg_p then launch_cg_x): its
// grid and spacings, x (the unknown), r, the two p buffers, the
// preconditioned residual z (FP64, or FP32 from an FP32 V-cycle when z32),
// whether z is M r (use_mz), its reduction partials and its solver state.
struct CgGroupCase {
    Grid g;
    Phys ph;
    FieldView<double> x, r, pb[2];
    int z32;
    FieldView<double> z;
    FieldView<float> zf;
    int use_mz;
    double* part;
    CgState* st;
};
// Append one lock-step CG iteration with p buffers pb[cur] (old) and
// pb[cur ^ 1] (new).  Returns 0.
int cg_group_iteration(GroupPlan& P, const std::vector<CgGroupCase>& cs, int cur);
// Append the copy of every src state into dst[0 .. n-1] (one device array,
// for one read-back).
void cg_group_gather(GroupPlan& P, const std::vector<const CgState*>& src, CgState* dst);

// One case's Helmholtz outer loops (main:758-887) for a grouped step: the
// three buffers of each component, u2's buffer among them (i2), the
// right-hand sides, the augmented viscosity, the pressure of the u outlet
// condition and the outer iteration counts; part / out receive the
// get_difference diagnostics (part null: none).
template <typename T>
struct HhGroupCase {
    Grid g;
    Phys ph;
    FieldView<T> x[3][3];
    int i2[3];
    FieldView<T> rhs[3], mu, pres;
    int mom[3];
    Tiling tl;    // the case's launch shapes
    double* part;
    double* out;  // out[q] receives component q's get_difference
};
// Append every launch of the loops of all three components, in lock-step
// across the cases (cases with fewer iterations drop out of later slots).
template <typename T>
int hh_group_loops(GroupPlan& P, const std::vector<HhGroupCase<T>>& cs, bool diag);
// get_difference of a[k], b[k] for the cases of `on` into on[k]->out[q]
// (nz planes: maxn for w, maxn-1 otherwise).
template <typename T>
void group_difference(GroupPlan& P, const std::vector<const HhGroupCase<T>*>& on, const std::vector<FieldView<T>>& a,
                      const std::vector<FieldView<T>>& b, int q);

// One case's pressure-solve set-up (ps_setup without its V-cycle): state
// reset, b from div(u3) with the inlet flux, b - mean, r = b - A x.
template <typename T>
struct SetupGroupCase {
    Grid g;
    Phys ph;
    FieldView<T> u3, v3, w3;
    const T* inlet;
    FieldView<double> rhs_p, r, x;
    double* part;
    CgState* st;
    double rtol, abstol, dtol;
    int maxits;
};
template <typename T>
void cg_group_setup(GroupPlan& P, const std::vector<SetupGroupCase<T>>& cs);

}  // namespace tpls

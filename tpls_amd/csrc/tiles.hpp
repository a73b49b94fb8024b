// Tiling of the pressure unknowns (i in 1..maxl-1, j in 1..P, k in 1..maxn-1)
// shared by the conjugate-gradient and multigrid kernels: tiles of
// BX x BY columns x CG_KC levels (one wavefront spans a 512-byte row
// segment; the level march reuses z neighbours in registers), walked
// grid-stride by at most CG_BLOCKS blocks so that a reduction has at most
// CG_BLOCKS partials.
#pragma once

#include "tpls_kernels.hpp"

namespace tpls {

constexpr int CG_KC = 4;
constexpr int CG_BLOCKS = 4096;

struct Tiles {
    int nbx, nby, nbz, ntiles, nblocks;
};

inline Tiles cg_tiles(const Grid& g)
{
    Tiles t;
    t.nbx = (g.maxl - 1 + BX - 1) / BX;
    t.nby = (g.P + BY - 1) / BY;
    t.nbz = (g.maxn - 1 + CG_KC - 1) / CG_KC;
    t.ntiles = t.nbx * t.nby * t.nbz;
    t.nblocks = t.ntiles < CG_BLOCKS ? t.ntiles : CG_BLOCKS;
    return t;
}

// Blocks of a 1-D launch are dealt to the 8 XCDs round-robin; this order
// gives each XCD a contiguous range of tiles, so tiles that share
// neighbour rows and planes run under one L2.  A bijection on [0, nblk).
__device__ __forceinline__ int xcd_tile(int b, int nblk)
{
    const int per = nblk >> 3;
    return b < (per << 3) ? (b & 7) * per + (b >> 3) : b;
}

// Blocks of `kernel` (static LDS only) that the device can hold at once: its
// CU count times the blocks of `threads` lanes one CU holds.  Read once per
// kernel instantiation (the caller keeps it in a static), on the first device
// it is launched on.
template <typename K>
inline int resident_blocks(K kernel, int threads)
{
    int dev = 0, cus = 0, per = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
        hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, kernel, threads, 0) != hipSuccess)
        return 0;
    return cus * per;
}

// Levels per block of a sweep that marches up k: nxy column tiles times nbz
// chunks of ks levels run in rounds of `slots` resident blocks, a round
// lasting about ks levels plus the pipeline's fill (three levels).  The chunk
// count that minimises rounds x (ks + 3) keeps the last round from running
// nearly empty.  slots <= 0 (unknown) falls back to `dflt` levels.
inline int march_chunk(int nxy, int nz, int slots, int dflt)
{
    if (slots <= 0 || nz <= 0) return dflt;
    int best_ks = dflt;
    long best = -1;
    for (int nbz = 1; nbz <= nz; ++nbz) {
        const int ks = (nz + nbz - 1) / nbz;
        if ((nz + ks - 1) / ks != nbz) continue;  // this ks is a smaller chunk count's
        const long rounds = (long(nxy) * nbz + slots - 1) / slots;
        const long cost = rounds * (ks + 3);
        if (best < 0 || cost < best) {
            best = cost;
            best_ks = ks;
        }
    }
    return best_ks;
}

// Decode tile t into the first interior cell (i, j, k0) of this thread.
__device__ __forceinline__ void tile_origin(const Tiles& tl, int t, int& i, int& j, int& k0)
{
    const int bx = t % tl.nbx;
    const int r = t / tl.nbx;
    const int by = r % tl.nby;
    const int bz = r / tl.nby;
    i = bx * BX + threadIdx.x + 1;
    j = by * BY + threadIdx.y + 1;
    k0 = bz * CG_KC + 1;
}

// Lanes of a fused red-black pass over a tile of TX x TY interior columns and
// its one-cell ring: RPR cells of the first colour in each of the TY+2 rows,
// BPR cells of the second in each of the TY interior rows; each colour's
// lanes rounded up to whole wavefronts, so that a wavefront is of one colour.
template <int TX, int TY>
struct RbLanes {
    static constexpr int RPR = (TX + 2) / 2, BPR = TX / 2;
    static constexpr int RC = (TY + 2) * RPR, BC = TY * BPR;
    static constexpr int RL = (RC + 63) / 64 * 64, BL = (BC + 63) / 64 * 64, NT = RL + BL;
};

}  // namespace tpls

// HIP kernels of the lava CA step and the Potts/Metropolis sweep, written for gfx950 (CDNA4).
//
// CA step (one synchronous step of engine/src/ca/LavaCA.hpp) = two stream-ordered launches for lava-front
// simulations:
//   ca_descent : clears the counts of the previous step's tile list, does the previous step's finish (step counters
//                -> histories, stop rule, LavaCA.hpp:1206-1229), and computes D(n) of every lava cell
//                (LavaCA.hpp:421-429, 877-888) of the staged tiles whose D(n) can have changed, one block per tile
//                staged in LDS.
//   ca_rules   : one block per staged tile (tile + 1-cell halo in LDS); the rules of every cell read the
//                start-of-step state only (type and temperature copies of the step's parity); the tile's next state
//                goes into the other copies (types that differ, every lava cell's temperature), flowIds and step
//                stamps of changed cells in place -- so no scatter pass follows (the reference applies its change
//                list after the rules, LavaCA.hpp:1052-1066); counters, tile flags, and the next step's tile list
//                (the tile itself and the neighbours it holds lava next to: the reference's "may change" set,
//                CaRules.hpp:69-79, at tile granularity) (LavaCA.hpp:963-1048, 1109-1202).
// The finish of the last step of a run of steps is one more launch (ca_finish).  Simulations that need every tile
// visited (ca.sweep = full) or classic mode add ca_front (a scan of every tile) before and ca_finish after each step.
// Every kernel skips the entries of a simulation whose stop rule fired, so the host can enqueue many steps without
// waiting.
//
// MC checkerboard sweep (engine/src/mc/PottsMC.cpp:393-514): one launch per colour over the colour's list
// of domain cells; Metropolis and mobility decisions are exact integer compares against host-built 53-bit
// thresholds (no exp on the device).
//
// Temperatures use only IEEE float subtract / compare (built with -ffp-contract=off).
#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <type_traits>

#include "vca_internal.h"

namespace vca {

// The simulation table (see BatchDev).  In the constant address space, uniform reads of an entry are scalar
// loads kept in SGPRs, like kernel arguments.
__constant__ CaSim g_sims[kSimTable];

namespace {

__device__ inline bool wrap_axis(int& v, int n, int per) {
    if (static_cast<unsigned>(v) < static_cast<unsigned>(n)) return true;
    if (!per) return false;
    v = v < 0 ? v + n : v - n;
    return true;
}

// Wave reductions by DPP row shifts (lane i takes lane i - 1, - 2, - 4, - 8 of its 16-lane row, no LDS round
// trip) and one v_readlane per row; every lane of the wave must be active.  The result is wave-uniform.
template <int CTRL>
__device__ inline uint32_t dpp_row_shr(uint32_t v) {
    return static_cast<uint32_t>(__builtin_amdgcn_update_dpp(0, static_cast<int>(v), CTRL, 0xF, 0xF, false));
}
template <int CTRL>
__device__ inline void dpp_add64(uint32_t& lo, uint32_t& hi) {
    const uint32_t l2 = dpp_row_shr<CTRL>(lo), h2 = dpp_row_shr<CTRL>(hi);
    const uint32_t nl = lo + l2;
    hi = hi + h2 + (nl < lo ? 1u : 0u);
    lo = nl;
}
__device__ inline unsigned long long wave_sum(unsigned long long v) {
    uint32_t lo = static_cast<uint32_t>(v), hi = static_cast<uint32_t>(v >> 32);
    dpp_add64<0x111>(lo, hi);  // row_shr:1
    dpp_add64<0x112>(lo, hi);  // row_shr:2
    dpp_add64<0x114>(lo, hi);  // row_shr:4
    dpp_add64<0x118>(lo, hi);  // row_shr:8
    unsigned long long t = 0;
#pragma unroll
    for (int r = 15; r < 64; r += 16)
        t += static_cast<unsigned long long>(static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(lo), r))) |
             (static_cast<unsigned long long>(static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(hi), r))) << 32);
    return t;
}

// threadIdx.x as a value the compiler cannot hoist out of a loop: per-lane index arithmetic derived from it is
// recomputed where it is used (a few instructions) instead of being kept live -- and spilled -- across a loop.
// The lane comes from two v_mbcnt instructions in an asm block (no memory access), the wave's first thread from
// a scalar register, so no vector register holds the thread index across the loop either.
__device__ inline unsigned opaque_tid() {
    unsigned lane;
    asm volatile("v_mbcnt_lo_u32_b32 %0, -1, 0\n\tv_mbcnt_hi_u32_b32 %0, -1, %0" : "=v"(lane));
    return (static_cast<unsigned>(__builtin_amdgcn_readfirstlane(static_cast<int>(threadIdx.x))) & ~63u) | lane;
}
// The lane index alone (no thread index needed, so nothing is kept live for it).
__device__ inline unsigned opaque_lane() {
    unsigned lane;
    asm volatile("v_mbcnt_lo_u32_b32 %0, -1, 0\n\tv_mbcnt_hi_u32_b32 %0, -1, %0" : "=v"(lane));
    return lane;
}
// 0 that the compiler cannot hoist out of a loop (a hoisted constant held in a vector register across the loop
// gets spilled).  One v_mov in an asm block.
__device__ inline uint32_t opaque_zero() {
    uint32_t z;
    asm volatile("v_mov_b32 %0, 0" : "=v"(z));
    return z;
}

// p (a global-memory pointer), held in scalar registers from here on: the compiler may not sink the load of a
// pointer into each of a row loop's branches (a scalar load and its wait per row).  An empty asm block; the result
// is typed as global memory again, so that its loads stay global (not flat) loads.
template <class T>
using global_ptr = __attribute__((address_space(1))) T*;
template <class T>
__device__ inline global_ptr<T> pinned(T* p) {
    asm volatile("" : "+s"(p));
    return (global_ptr<T>)p;
}

// v of lane `lane` (uniform) into scalar registers.
__device__ inline uint64_t readlane64(uint64_t v, int lane) {
    const uint32_t lo = static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(static_cast<uint32_t>(v)), lane));
    const uint32_t hi = static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(static_cast<uint32_t>(v >> 32)), lane));
    return static_cast<uint64_t>(lo) | (static_cast<uint64_t>(hi) << 32);
}

// Set bits of m below the calling lane (v_mbcnt: no lane mask is built, so none is kept in registers).
__device__ inline uint32_t mbcnt64(unsigned long long m) {
    return __builtin_amdgcn_mbcnt_hi(static_cast<uint32_t>(m >> 32),
                                      __builtin_amdgcn_mbcnt_lo(static_cast<uint32_t>(m), 0u));
}

// Sub-list of tile t (its index within its simulation; with kSub = 1 there is one): t plus its higher 5-bit groups,
// modulo kSub (a power of two up to 32).  Every aligned run of kSub consecutive tiles covers each sub-list once (the
// added groups are constant over it), so a simulation of n tiles holds at most ceil(n / kSub) entries of a sub-list;
// and since every bit of t takes part, a front that covers only some tile columns still spreads over all sub-lists
// (the consumers walk them in an interleaved order).
__device__ inline uint32_t sub_of(uint32_t t) {
    return (t + (t >> 5) + (t >> 10) + (t >> 15) + (t >> 20)) & (kSub - 1);
}

// Sub-list j of tile list (par, shape) and its count (each count on a 64-byte line of its own).
__device__ inline uint32_t* list_count(const BatchDev& B, int par, int shape, int j) {
    return &B.ctrl->nactive[par][shape][j][0];
}
// Count of descent list (par, shape) (each on a 64-byte line of its own).
__device__ inline uint32_t* desc_count(const BatchDev& B, int par, int shape) {
    return B.desc_count + (par * kShapes + shape) * 16;
}
__device__ inline uint64_t* sub_list(const BatchDev& B, int par, int shape, int j) {
    return B.active + static_cast<size_t>((par * kShapes + shape) * kSub + j) * B.subcap;
}

// Appends ent of every lane with add (all 64 lanes call) to sub-list (par, shape, sub_of(tile)): the lanes are
// grouped by sub-list, the lowest lane of each group reserves the group's slots, all groups with one atomic
// instruction (its lanes address different counts), and each lane reads its group's base from the leader.
__device__ inline void wave_list_append(const BatchDev& B, int par, int shape, bool add, uint32_t tile, uint64_t ent) {
    const unsigned long long m = __ballot(add);
    if (m == 0) return;
    const int lane = static_cast<int>(opaque_lane());
    const uint32_t j = sub_of(tile);
    int lead = lane;
    uint32_t size = 0, off = 0;
    for (unsigned long long rest = m; rest;) {
        const int l0 = static_cast<int>(__builtin_ctzll(rest));
        const uint32_t j0 = static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(j), l0));
        const unsigned long long grp = __ballot(add && j == j0);
        if (add && j == j0) {
            lead = l0;
            off = mbcnt64(grp);
        }
        if (lane == l0) size = static_cast<uint32_t>(__popcll(grp));
        rest &= ~grp;
    }
    uint32_t base = 0;
    if (add && lead == lane) base = atomicAdd(list_count(B, par, shape, static_cast<int>(j)), size);
    base = static_cast<uint32_t>(__builtin_amdgcn_ds_bpermute(lead << 2, static_cast<int>(base)));
    if (add) sub_list(B, par, shape, static_cast<int>(j))[base + off] = ent;
}

// Wave-aggregated slot reservation: every lane of the wave must call this (pred may differ per lane).  The
// leader's reserved base reaches the other lanes by v_readlane (no second LDS round trip).
__device__ inline uint32_t wave_append(bool pred, uint32_t* counter) {
    const unsigned long long mask = __ballot(pred);
    const uint32_t below = mbcnt64(mask);
    uint32_t base = 0;
    if (mask != 0) {
        const int leader = __ffsll(static_cast<long long>(mask)) - 1;
        if (pred && below == 0) base = atomicAdd(counter, static_cast<uint32_t>(__popcll(mask)));
        base = static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(base), leader));
    }
    return base + below;
}

// K appends of one wave (slot[k] for pred[k]) with one atomic: all 64 lanes must be active.
template <int K>
__device__ inline void wave_append_k(const bool (&pred)[K], uint32_t* counter, uint32_t (&slot)[K]) {
    unsigned long long m[K];
    uint32_t tot = 0;
#pragma unroll
    for (int k = 0; k < K; ++k) {
        m[k] = __ballot(pred[k]);
        tot += static_cast<uint32_t>(__popcll(m[k]));
    }
    uint32_t base = 0;
    if (tot != 0) {
        if (opaque_lane() == 0) base = atomicAdd(counter, tot);
        base = static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(base), 0));
    }
#pragma unroll
    for (int k = 0; k < K; ++k) {
        slot[k] = base + mbcnt64(m[k]);
        base += static_cast<uint32_t>(__popcll(m[k]));
    }
}

// Reduces N per-thread values over the block (a sum, or a max where bit i of max_mask is set) and folds the
// block's results into row[0..N) with one atomic each (every thread of the block calls).
template <int N>
__device__ inline void block_fold(unsigned long long (&v)[N], uint32_t max_mask, unsigned long long* row) {
    __shared__ unsigned long long red[N][kBlock / 64];
#pragma unroll
    for (int i = 0; i < N; ++i) {
        const bool mx = (max_mask >> i) & 1u;
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) {
            const unsigned long long y = __shfl_xor(v[i], o);
            v[i] = mx ? (y > v[i] ? y : v[i]) : v[i] + y;
        }
    }
    const int w = static_cast<int>(threadIdx.x >> 6);
    if ((threadIdx.x & 63u) == 0) {
#pragma unroll
        for (int i = 0; i < N; ++i) red[i][w] = v[i];
    }
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma unroll
        for (int i = 0; i < N; ++i) {
            const bool mx = (max_mask >> i) & 1u;
            unsigned long long r = red[i][0];
            for (int k = 1; k < kBlock / 64; ++k) r = mx ? (red[i][k] > r ? red[i][k] : r) : r + red[i][k];
            if (r == 0) continue;
            if (mx)
                atomicMax(&row[i], r);
            else
                atomicAdd(&row[i], r);
        }
    }
}

// Variant v's mask by selects, not by a dynamic index into the table (which would force a register copy of
// the table into private memory).
__device__ inline uint32_t nb_mask(const NbTable& nb, uint32_t v) {
    uint32_t m = nb.mask[0];
    m = v == 1 ? nb.mask[1] : m;
    m = v == 2 ? nb.mask[2] : m;
    m = v == 3 ? nb.mask[3] : m;
    m = v == 4 ? nb.mask[4] : m;
    m = v == 5 ? nb.mask[5] : m;
    return m;
}

// ------------------------------------------------------------------------------------------------
// CA helpers
// ------------------------------------------------------------------------------------------------

struct TileCoord {
    int x0, y0, z0;
};
__device__ inline TileCoord tile_origin(const CaParams& P, uint32_t t) {
    const uint32_t r = div_magic(t, P.div_ntx_m, P.div_ntx_p);
    const int tx = static_cast<int>(t - r * static_cast<uint32_t>(P.ntx));
    const uint32_t tz_ = div_magic(r, P.div_nty_m, P.div_nty_p);
    const int ty = static_cast<int>(r - tz_ * static_cast<uint32_t>(P.nty));
    const int tz = static_cast<int>(tz_);
    return TileCoord{tx * kTileX, ty * P.ty, tz * P.tz};
}
// Lane q < 27 (q != 13) of the calling wave marks neighbour tile q of tile (tx, ty, tz) as next to lava (periodic
// axes wrap, tiles past an absorbing edge do not exist).
__device__ inline void mark_near(const CaParams& P, const CaDev& D, int tx, int ty, int tz, int q) {
    if (q >= 27 || q == 13) return;
    int qx = tx + q % 3 - 1, qy = ty + (q / 3) % 3 - 1, qz = tz + q / 9 - 1;
    if (!wrap_axis(qx, P.ntx, P.per[0]) || !wrap_axis(qy, P.nty, P.per[1]) || !wrap_axis(qz, P.ntz, P.per[2])) return;
    D.tile_near[static_cast<uint32_t>(qx) +
                static_cast<uint32_t>(P.ntx) * (static_cast<uint32_t>(qy) + static_cast<uint32_t>(P.nty) * static_cast<uint32_t>(qz))] = 1;
}

// Directions q = (dx+1) + 3(dy+1) + 9(dz+1) toward which the cell at (lx, ly, lz) of its tile lies on the boundary:
// per axis 0 always, -1 at the first cell, +1 at the last real cell (xm / ym / zm, which is short of the tile's edge
// in a partial tile).  Bit 13 (the tile itself) is clear.
__device__ inline uint32_t boundary_dirs(int lx, int ly, int lz, int xm, int ym, int zm) {
    const uint32_t ax = 2u | (lx == 0 ? 1u : 0u) | (lx == xm ? 4u : 0u);
    const uint32_t ay = 2u | (ly == 0 ? 1u : 0u) | (ly == ym ? 4u : 0u);
    const uint32_t az = 2u | (lz == 0 ? 1u : 0u) | (lz == zm ? 4u : 0u);
    const uint32_t ysel = ((ay & 1u) ? 0x007u : 0u) | ((ay & 2u) ? 0x038u : 0u) | ((ay & 4u) ? 0x1C0u : 0u);
    const uint32_t xy = (ax * 0x49u) & ysel;  // 0x49: the 3 bits of ax repeated at i + 3j, j = 0..2
    const uint32_t zsel = ((az & 1u) ? 0x1FFu : 0u) | ((az & 2u) ? 0x3FE00u : 0u) | ((az & 4u) ? 0x7FC0000u : 0u);
    return ((xy * 0x40201u) & zsel) & ~(1u << 13);  // 0x40201: xy repeated at 9k, k = 0..2
}
// Lane q < 27 of the calling wave: when bit q of bnd is set, tell neighbour tile q of tile (tx, ty, tz) that this
// tile holds lava next to it (byte 26 - q of the neighbour: the direction back to this tile).
__device__ inline void mark_halo(const CaParams& P, const CaDev& D, int tx, int ty, int tz, int q, uint32_t bnd) {
    if (q >= 27 || !((bnd >> q) & 1u)) return;
    int qx = tx + q % 3 - 1, qy = ty + (q / 3) % 3 - 1, qz = tz + q / 9 - 1;
    if (!wrap_axis(qx, P.ntx, P.per[0]) || !wrap_axis(qy, P.nty, P.per[1]) || !wrap_axis(qz, P.ntz, P.per[2])) return;
    const uint32_t n = static_cast<uint32_t>(qx) + static_cast<uint32_t>(P.ntx) *
                                                       (static_cast<uint32_t>(qy) + static_cast<uint32_t>(P.nty) * static_cast<uint32_t>(qz));
    D.tile_halo[static_cast<size_t>(n) * kHaloBytes + static_cast<uint32_t>(26 - q)] = 1;
}
// Lane q < 27 of wave 0 (all 64 lanes call): when bit q of dirs is set, lists tile q of the 3 x 3 x 3 block around tile
// (tx, ty, tz) (q = 13: the tile itself) for the simulation's step `step` in list copy par -- once per step, by the
// tile's stamp (listed): the first lane to stamp a tile appends it.  Periodic axes wrap, tiles past an absorbing edge
// do not exist.  One list atomic instruction per call.
__device__ inline void push_next(const CaParams& P, const CaDev& D, const BatchDev& B, int tx, int ty, int tz, int q,
                                 uint32_t dirs, uint32_t slot, uint32_t step, int par, int shape) {
    bool add = false;
    uint32_t n = 0;
    if (q < 27 && ((dirs >> q) & 1u)) {
        int qx = tx + q % 3 - 1, qy = ty + (q / 3) % 3 - 1, qz = tz + q / 9 - 1;
        if (wrap_axis(qx, P.ntx, P.per[0]) && wrap_axis(qy, P.nty, P.per[1]) && wrap_axis(qz, P.ntz, P.per[2])) {
            n = static_cast<uint32_t>(qx) +
                static_cast<uint32_t>(P.ntx) * (static_cast<uint32_t>(qy) + static_cast<uint32_t>(P.nty) * static_cast<uint32_t>(qz));
            add = atomicExch(&D.listed[n], step) != step;
        }
    }
    wave_list_append(B, par, shape, add, n, (static_cast<uint64_t>(step) << 32) | (slot << kTileBits) | n);
}
// Tile layers below a tile that D(n) of its cells reads: drops reach kMaxDrop + 1 cells below its lowest layer.
template <int TZ>
constexpr int kDescentLayers = (kMaxDrop + 1 + TZ - 1) / TZ;

// After a step in which the tile (tx, ty, tz) changed a cell's type: every tile whose D(n) reads it -- the 3 x 3
// columns around it, from its layer up kDescentLayers tile layers -- that can hold lava after the step into the
// descent list of the next step (list copy par, entries for `step`), each once (the `dirtied` stamp).  A tile holds
// lava after the step only if it held lava before it or had a cell born in it (then it changed a type and lists
// itself), so its lava flag, read before or after its own rules of this step wrote it, selects every such tile.
// Thread q of the block takes tile q of that set (all threads call: one list atomic per wave).
template <int TZ>
__device__ inline void push_descent(const CaParams& P, const CaDev& D, const BatchDev& B, int tx, int ty, int tz, int q,
                                    uint32_t slot, uint32_t step, int par, int shape) {
    static_assert(9 * (kDescentLayers<TZ> + 1) <= kBlock, "one thread per tile of the set");
    bool add = false;
    uint32_t n = 0;
    if (q < 9 * (kDescentLayers<TZ> + 1)) {
        int qx = tx + q % 3 - 1, qy = ty + (q / 3) % 3 - 1;
        const int qz = tz + q / 9;
        if (qz < P.ntz && wrap_axis(qx, P.ntx, P.per[0]) && wrap_axis(qy, P.nty, P.per[1])) {
            n = static_cast<uint32_t>(qx) +
                static_cast<uint32_t>(P.ntx) * (static_cast<uint32_t>(qy) + static_cast<uint32_t>(P.nty) * static_cast<uint32_t>(qz));
            add = (q == 4 || (D.tile_flags[n] & kFlagLava)) && atomicExch(&D.dirtied[n], step) != step;
        }
    }
    const uint32_t at = wave_append(add, desc_count(B, par, shape));
    if (add)
        B.desc_list[static_cast<size_t>(par * kShapes + shape) * B.subcap + at] =
            (static_cast<uint64_t>(step) << 32) | (slot << kTileBits) | n;
}

__device__ inline unsigned wave_or(unsigned v) {
    v |= dpp_row_shr<0x111>(v);
    v |= dpp_row_shr<0x112>(v);
    v |= dpp_row_shr<0x114>(v);
    v |= dpp_row_shr<0x118>(v);
    return static_cast<unsigned>(__builtin_amdgcn_readlane(static_cast<int>(v), 15) |
                                 __builtin_amdgcn_readlane(static_cast<int>(v), 31) |
                                 __builtin_amdgcn_readlane(static_cast<int>(v), 47) |
                                 __builtin_amdgcn_readlane(static_cast<int>(v), 63));
}

// Cell k (0..3) of thread threadIdx.x inside the tile.
__device__ inline void tile_cell(const CaParams& P, const TileCoord& o, int k, int& x, int& y, int& z) {
    const int lx = static_cast<int>(threadIdx.x & 63u);
    const int row = static_cast<int>(threadIdx.x >> 6) + 4 * k;
    x = o.x0 + lx;
    y = o.y0 + row % P.ty;
    z = o.z0 + row / P.ty;
}

__device__ inline uint32_t col_index(const CaParams& P, int x, int y) {
    return static_cast<uint32_t>(x) + static_cast<uint32_t>(P.nx) * static_cast<uint32_t>(y);
}

// h % m for a small m >= 1 (m < 2^16) without the 64-bit division routine: 2^32 = q*m + r32 gives
// h = hi * 2^32 + lo == (hi mod m) * r32 + lo mod m  (mod m).
__device__ inline uint32_t mod_small(uint64_t h, uint32_t m) {
    const uint32_t r32 = (0u - m) % m;  // 2^32 mod m
    const uint32_t a = static_cast<uint32_t>(h >> 32) % m, b = static_cast<uint32_t>(h) % m;
    return (a * r32 + b) % m;
}
// The same value for m in [1, 2^16) from its reciprocal inv = 1 / m (double) and r32 = 2^32 mod m (both set by the
// host): a 32-bit quotient estimate trunc(x * inv) is exact or one too small (its relative error is below 2^-52, so
// it is off by less than 2^-20, and a quotient's fraction is 0 or at least 1 / m), fixed by one subtraction -- three
// multiplications instead of four integer divisions.
__device__ inline uint32_t mod32_inv(uint32_t x, uint32_t m, double inv) {
    const uint32_t q = static_cast<uint32_t>(static_cast<double>(x) * inv);
    const uint32_t r = x - q * m;
    return r >= m ? r - m : r;
}
__device__ inline uint32_t mod_small_inv(uint64_t h, uint32_t m, uint32_t r32, double inv) {
    const uint32_t a = mod32_inv(static_cast<uint32_t>(h >> 32), m, inv);
    const uint32_t b = mod32_inv(static_cast<uint32_t>(h), m, inv);
    return mod32_inv(a * r32 + b, m, inv);  // a * r32 + b < m * m < 2^32
}

// CandidateVote (engine/src/ca/CaRules.hpp:127-183) in registers for the common case, a cell whose candidates
// all carry one flowId: the winner is that flowId with the largest candidate temperature (most votes, and no tie
// to break).  A second flowId among the candidates sets overflow: such cells (where flows meet) are gathered again.
struct VoteFirst {
    int n = 0;             // candidates counted (all of flowId id)
    bool overflow = false;
    int32_t id = 0;
    float mt = -__builtin_inff();  // read only when n > 0: the first candidate's t replaces it (t is finite)
    // add(f, t) when ok, nothing otherwise.  The first candidate needs no case of its own: with n == 0 it
    // matches, takes id = f, and mt = t since -inf < t.
    __device__ void add_if(bool ok, int32_t f, float t, int /*k*/) {
        const bool match = n == 0 || f == id;
        const bool same = ok && match;
        overflow = overflow || (ok && !match);
        id = same ? f : id;
        mt = (same && mt < t) ? t : mt;  // std::max(a, b) == (a < b ? b : a)
        n += same ? 1 : 0;
    }
    __device__ void winner(int32_t& flow, float& temp) const {
        flow = id;
        temp = mt;
    }
};

// CandidateVote for candidates carrying at most two flowIds (where two flows meet): most votes wins; a tie between
// the two picks by the hash among them in ascending flowId order (engine/src/ca/CaRules.hpp:127-183); the
// winner's temperature is the largest of its candidates.  A third flowId sets overflow (the full vote decides).
struct VoteTwo {
    int n = 0, n2 = 0;     // candidates of id / of id2
    bool overflow = false;
    int32_t id = 0, id2 = 0;
    float mt = -__builtin_inff(), mt2 = -__builtin_inff();  // as in VoteFirst: an empty slot matches any f
    __device__ void add_if(bool ok, int32_t f, float t, int /*k*/) {
        const bool m1 = n == 0 || f == id;
        const bool m2 = n2 == 0 || f == id2;
        const bool one = ok && m1;
        const bool two = ok && !m1 && m2;
        overflow = overflow || (ok && !m1 && !m2);
        id = one ? f : id;
        mt = (one && mt < t) ? t : mt;  // std::max(a, b) == (a < b ? b : a)
        n += one ? 1 : 0;
        id2 = two ? f : id2;
        mt2 = (two && mt2 < t) ? t : mt2;
        n2 += two ? 1 : 0;
    }
    __device__ void winner(uint64_t h_index, int32_t& flow, float& temp) const {
        bool take2 = n2 > n;
        if (n2 > 0 && n2 == n) {
            const uint32_t pick = mod_small(splitmix64(h_index ^ kSaltTieBreak), 2u);
            take2 = (pick == 1u) == (id2 > id);  // pick 0: the smaller flowId
        }
        flow = take2 ? id2 : id;
        temp = take2 ? mt2 : mt;
    }
};

// Whether any candidate the rules accept carries a candidate temperature at the solidus or above (gathered over the
// union of the variants: a cell without one is not born in any step that keeps the types and lowers no bound).
struct VoteAlive {
    float solidus;
    bool alive = false;
    __device__ void add_if(bool ok, int32_t, float t, int) { alive = alive || (ok && t >= solidus); }
};

// The accepted candidate positions of a cell (bit k: the neighbour at position k is a candidate), for the full vote
// of vote_positions.
struct VoteMask {
    uint32_t acc = 0;
    __device__ void add_if(bool ok, int32_t, float, int k) { acc |= (ok ? 1u : 0u) << k; }
};

}  // namespace

// ------------------------------------------------------------------------------------------------
// CA set-up kernels
// ------------------------------------------------------------------------------------------------

__global__ void __launch_bounds__(kBlock) ca_terrain_kernel(CaParams P, CaDev D, const int32_t* height) {
    const uint64_t n = static_cast<uint64_t>(P.nx) * P.ny * P.nz;
    const uint32_t plane = static_cast<uint32_t>(P.nx) * static_cast<uint32_t>(P.ny);
    for (uint64_t i = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; i < n;
         i += static_cast<uint64_t>(gridDim.x) * blockDim.x) {
        const uint32_t col = static_cast<uint32_t>(i % plane);
        const int z = static_cast<int>(i / plane);
        int h = height[col];
        h = h < P.nz ? h : P.nz;
        const uint8_t t = z < h ? kRock : kAir;
        D.type[0][i] = t;
        D.type[1][i] = t;
    }
}

// Tiles that hold the vent cell of some flowId (vent_cell[f]): the rules test a lava cell against its flow's vent only
// there.
__global__ void __launch_bounds__(kBlock) ca_vent_tiles_kernel(CaParams P, CaDev D) {
    const uint32_t f = blockIdx.x * kBlock + threadIdx.x;
    if (f >= static_cast<uint32_t>(P.nvent_slots)) return;
    const uint32_t i = D.vent_cell[f];
    if (i == 0xFFFFFFFFu) return;
    const uint32_t nx = static_cast<uint32_t>(P.nx), plane = nx * static_cast<uint32_t>(P.ny);
    const uint32_t z = i / plane, y = (i - z * plane) / nx, x = i - z * plane - y * nx;
    D.tile_vent[x / kTileX + static_cast<uint32_t>(P.ntx) * (y / static_cast<uint32_t>(P.ty) +
                                                             static_cast<uint32_t>(P.nty) * (z / static_cast<uint32_t>(P.tz)))] = 1;
}

__global__ void ca_seeds_kernel(CaDev D, const uint32_t* cells, const int32_t* flows, const float* temps, int n) {
    // Sequential in list order, like placeSeeds (LavaCA.hpp:215-227): a later seed in the same cell wins.
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    for (int k = 0; k < n; ++k) {
        const uint32_t i = cells[k];
        D.type[0][i] = kLava;
        D.type[1][i] = kLava;
        D.flow[i] = flows[k];
        D.temp[0][i] = temps[k];  // the state before step 0
        D.solid[i] = 0;
        D.arrival[i] = 0;
    }
}

// The stop rule before the first step (LavaCA.hpp:1252-1280): no lava (lava mode), no air (classic mode),
// or maxSteps 0.
__global__ void ca_init_stop_kernel(CaParams P, CaDev D) {
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    CaCtrl* c = D.ctrl;
    int32_t stop = 0;
    if (!P.classic && c->lava_total == 0)
        stop = 1;  // NoLava
    else if (P.classic && c->air_total == 0)
        stop = 2;  // NoAir
    else if (P.max_steps <= 0)
        stop = 4;  // MaxSteps
    c->stop = stop;
    c->prefix = hash_prefix(P.seed, 0);
}

// Tile flags and the initial LAVA / AIR counts (countInitialState, LavaCA.hpp:1252-1280).
__global__ void __launch_bounds__(kBlock) ca_init_flags_kernel(CaParams P, CaDev D) {
    __shared__ uint32_t s_bnd[kBlock / 64];
    __shared__ unsigned long long s_tot[2][kBlock / 64];
    // the block's lava and AIR counts, added to the totals once at the end (one atomic per block and counter:
    // per-wave atomics on the same two addresses serialised, 2.5 ms for a 512 x 512 x 256 grid)
    unsigned long long tl = 0, ta = 0;
    for (uint32_t t = blockIdx.x; t < P.ntiles; t += gridDim.x) {
        const TileCoord o = tile_origin(P, t);
        const int xm = min(kTileX - 1, P.nx - 1 - o.x0), ym = min(P.ty - 1, P.ny - 1 - o.y0),
                  zm = min(P.tz - 1, P.nz - 1 - o.z0);
        int lava = 0, air = 0;
        uint32_t bnd = 0;  // directions in which this tile holds lava on its boundary (mark_halo)
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            int x, y, z;
            tile_cell(P, o, k, x, y, z);
            if (x < P.nx && y < P.ny && z < P.nz) {
                const uint8_t ty = D.type[0][static_cast<uint32_t>(x) +
                                             static_cast<uint32_t>(P.nx) * (static_cast<uint32_t>(y) +
                                                                            static_cast<uint32_t>(P.ny) * z)];
                lava += ty == kLava;
                air += ty == kAir;
                if (ty == kLava) bnd |= boundary_dirs(x - o.x0, y - o.y0, z - o.z0, xm, ym, zm);
            }
        }
        bnd = wave_or(bnd);
        if ((threadIdx.x & 63u) == 0) s_bnd[threadIdx.x >> 6] = bnd;
        tl += wave_sum(static_cast<unsigned long long>(lava));
        ta += wave_sum(static_cast<unsigned long long>(air));
        const int any_lava = __syncthreads_or(lava);
        const int any_air = __syncthreads_or(air);
        if (threadIdx.x == 0) D.tile_flags[t] = static_cast<uint8_t>((any_lava ? kFlagLava : 0) | (any_air ? kFlagAir : 0));
        if (any_lava) {
            const int tx = o.x0 / kTileX, ty = o.y0 / P.ty, tz = o.z0 / P.tz;
            mark_near(P, D, tx, ty, tz, static_cast<int>(threadIdx.x));
            uint32_t all = 0;
            for (int w = 0; w < kBlock / 64; ++w) all |= s_bnd[w];  // written before the barriers above
            mark_halo(P, D, tx, ty, tz, static_cast<int>(threadIdx.x), all);
        }
        __syncthreads();  // s_bnd is rewritten by the next tile
    }
    if ((threadIdx.x & 63u) == 0) {
        s_tot[0][threadIdx.x >> 6] = tl;
        s_tot[1][threadIdx.x >> 6] = ta;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long l = 0, a = 0;
        for (int w = 0; w < kBlock / 64; ++w) {
            l += s_tot[0][w];
            a += s_tot[1][w];
        }
        if (l) atomicAdd(&D.ctrl->lava_total, l);
        if (a) atomicAdd(&D.ctrl->air_total, a);
    }
}

// ------------------------------------------------------------------------------------------------
// CA step kernels
// ------------------------------------------------------------------------------------------------

// Staged-tile lists of one step for the simulations that need a scan (classic mode, or ca.sweep = full), one lane per
// global tile.  A tile is visited when it holds lava (lava cools) or holds AIR next to a tile with lava: the
// reference's "may change" set (CaRules.hpp:69-79, LavaCA.hpp:844-871) at tile granularity, every tile with
// ca.sweep = full; a visited tile is staged when it or its 1-cell halo holds lava, the others settle here.  Lava-front
// simulations are skipped: their rules kernel pushed this step's tiles already.  One list per tile shape (copy of
// parity par); a block counts its entries by wave ballots and reserves them with one atomic per list (a device-wide
// counter is the contended resource: one atomic per block, not per wave).  The scan is latency-bound, so one tile per
// lane (a 512 x 512 x 256 grid of 65536 tiles: 8 per lane 19.5 us, 4 12.7, 2 9.3, 1 8.0 per step).
// Appends ent of every thread with code >= 0 to sub-list (par, code, sub_of(tile)), all threads of the block
// calling: positions within the block from LDS counts, one device atomic per non-empty sub-list of the block.
__device__ inline void block_list_append(const BatchDev& B, int par, int code, uint32_t tile, uint64_t ent) {
    __shared__ uint32_t s_cnt[kShapes * kSub], s_base[kShapes * kSub];
    static_assert(kShapes * kSub <= kBlock, "one thread per sub-list");
    if (threadIdx.x < kShapes * kSub) s_cnt[threadIdx.x] = 0u;
    __syncthreads();
    const int key = code >= 0 ? code * kSub + static_cast<int>(sub_of(tile)) : 0;
    uint32_t pos = 0;
    if (code >= 0) pos = atomicAdd(&s_cnt[key], 1u);
    __syncthreads();
    if (threadIdx.x < kShapes * kSub) {
        const uint32_t c = s_cnt[threadIdx.x];
        s_base[threadIdx.x] = c ? atomicAdd(list_count(B, par, threadIdx.x / kSub, threadIdx.x % kSub), c) : 0u;
    }
    __syncthreads();
    if (code >= 0) sub_list(B, par, code, key % kSub)[s_base[key] + pos] = ent;
}

__global__ void __launch_bounds__(kBlock) ca_front_kernel(BatchDev B, uint32_t shapes, int par) {
    const int lane = static_cast<int>(threadIdx.x & 63u);
    const uint32_t g = blockIdx.x * kBlock + threadIdx.x;
    const uint32_t slot = g < B.ntiles ? B.owner[g] : kNoSlot;
    // a wave's 64 tiles nearly always belong to one simulation: then its parameters are read once, as scalars
    const uint32_t s0 = static_cast<uint32_t>(__builtin_amdgcn_readfirstlane(static_cast<int>(slot)));
    const bool uniform = __ballot(slot != s0) == 0;
    uint64_t ent = 0;
    int code = -1;  // shape of the list the tile goes to, -1: none
    auto visit = [&](const CaSim& sim, uint32_t sl) -> bool {
        if (!sim.running || sim.push || sim.D.ctrl->stop) return false;
        const CaParams& P = sim.P;
        const uint32_t t = g - B.tile_base[sl];
        const uint8_t f = sim.D.tile_flags[t];
        const bool lava = !P.classic && (f & kFlagLava);
        // a neighbour tile with lava marked this tile when its own step ended (mark_near); lava tiles are visited
        // every step in lava mode, so their marks are renewed and this step's are cleared here
        uint8_t* near = sim.D.tile_near;
        const uint8_t nb = near[t];
        if (nb && !P.classic) near[t] = 0;
        // a tile that changed a cell's type in the previous step brings the other type copy up to date in this one
        const int32_t now = sim.D.ctrl->steps;
        const bool changed = now > 0 && sim.D.tile_tchg[t] == static_cast<uint32_t>(now);
        const bool act = P.dense || lava || changed || (P.classic && (f & kFlagLava) && (f & kFlagAir)) ||
                         ((f & kFlagAir) && nb);
        // a visited tile goes to the rules only when it or its 1-cell halo holds lava (its own flag, or a
        // neighbour's halo byte, mark_halo; the bytes are cleared for the next step in lava mode) or it changed in
        // the previous step.  Any other visited tile is settled here: none of its cells can change, its flags stay
        // as they are.
        bool stage = (f & kFlagLava) != 0 || changed;
        if (act) {
            uint4* hb = reinterpret_cast<uint4*>(sim.D.tile_halo + static_cast<size_t>(t) * kHaloBytes);
            bool any = stage;  // a tile with lava is staged anyway: its bytes are only cleared
            if (!stage) {
                const uint4 h0 = hb[0], h1 = hb[1];
                any = (h0.x | h0.y | h0.z | h0.w | h1.x | h1.y | h1.z | h1.w) != 0;
                stage = any;
            }
            if (any && !P.classic) {
                hb[0] = make_uint4(0, 0, 0, 0);
                hb[1] = make_uint4(0, 0, 0, 0);
            }
        }
        ent = (static_cast<uint64_t>(static_cast<uint32_t>(now)) << 32) | (sl << kTileBits) | t;
        code = act && stage ? sim.shape : -1;
        return act;
    };
    // visited tiles are counted here (staged or settled): one atomic per wave when the wave is one simulation's
    if (uniform) {
        if (s0 != kNoSlot) {
            const CaSim& sim = g_sims[B.sim_base + s0];
            const unsigned long long m = __ballot(visit(sim, s0));
            if (m && lane == 0)
                atomicAdd(&sim.D.part[(g >> 6) % kPartRows].tiles, static_cast<unsigned long long>(__popcll(m)));
        }
    } else if (slot != kNoSlot) {
        const CaSim& sim = g_sims[B.sim_base + slot];
        if (visit(sim, slot)) atomicAdd(&sim.D.part[(g - B.tile_base[slot]) % kPartRows].tiles, 1ULL);
    }
    (void)shapes;
    block_list_append(B, par, code, static_cast<uint32_t>(ent) & kTileMask, ent);
}

// finishStep (LavaCA.hpp:1206-1229) + frontCounts (LavaCA.hpp:865-871) of the simulation in slot `slot`: sums and
// clears its counter rows, writes the histories, advances the step count and applies the stop rule.  Every thread
// of the block calls (block-uniform slot).
__device__ __forceinline__ void finish_slot(const BatchDev& B, uint32_t slot) {
    __shared__ unsigned long long red[8][kBlock / 64];
    const CaSim& sim = g_sims[B.sim_base + slot];
    if (!sim.running) return;
    const CaParams& P = sim.P;
    const CaDev& D = sim.D;
    CaCtrl* c = D.ctrl;
    if (__builtin_amdgcn_readfirstlane(c->stop)) return;
    // the step and the next step's hash prefix as wave-uniform (scalar) values
    const int32_t t = __builtin_amdgcn_readfirstlane(c->steps);
    const uint64_t next_prefix = hash_prefix(P.seed, static_cast<uint64_t>(t) + 1u);
    unsigned long long v[8] = {};
    for (int r = static_cast<int>(threadIdx.x); r < kPartRows; r += kBlock) {
        CaPart* row = &D.part[r];
        v[0] += row->lava_after;
        v[1] += row->born;
        v[2] += row->changed;
        v[3] += row->tiles;
        v[4] += row->staged;
        v[5] += row->air_cand;
        v[6] += row->air_alive;
        v[7] += row->tchanged;
        *row = CaPart{0, 0, 0, 0, 0, 0, 0, 0};  // cleared for the next step
    }
    const int w = static_cast<int>(threadIdx.x >> 6);
#pragma unroll
    for (int q = 0; q < 8; ++q) {
        v[q] = wave_sum(v[q]);
        if ((threadIdx.x & 63u) == 0) red[q][w] = v[q];
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long sum[8] = {};
        for (int k = 0; k < kBlock / 64; ++k)
            for (int q = 0; q < 8; ++q) sum[q] += red[q][k];
        const unsigned long long lava_after = sum[0], born = sum[1], changed = sum[2], tiles = sum[3],
                                 staged = sum[4], air_cand = sum[5], air_alive = sum[6], tchanged = sum[7];
        const unsigned long long active = P.classic ? c->lava_total + born : lava_after;
        const unsigned long long air = c->air_total - born;
        D.hist_active[t] = active;
        D.hist_changed[t] = changed;
        c->steps = t + 1;
        c->prefix = next_prefix;
        c->lava_total = active;
        c->air_total = air;
        int32_t stop = 0;
        if (!P.classic && active == 0)
            stop = 1;  // NoLava
        else if (P.classic && air == 0)
            stop = 2;  // NoAir
        else if (changed == 0 && P.nb.variants == 1)
            stop = 3;  // NoChange
        else if (t + 1 >= P.max_steps)
            stop = 4;  // MaxSteps
        // quiet (below) with the host's hold on: the simulation waits for the host's fast-forward
        if (stop == 0 && sim.push && air_cand == 0 && c->hold) stop = kStopHeld;
        c->stop = stop;
        // Quiescence (lava mode): no AIR cell had LAVA at a position the rules read in this step, so none was born;
        // LAVA only turns into BASALT and AIR stays AIR, so no AIR cell can get a candidate in any later step either:
        // every later step only cools the lava cells (and keeps the vents at eruptTemp while they erupt), each on its
        // own.
        if (!P.classic && air_cand == 0 && c->quiet == 0) c->quiet = t + 1;
        // No type changed in this step and no AIR candidate could be born under any variant at these or lower
        // temperatures: the next steps keep the same types until a lava cell solidifies.
        c->qdead = (!P.classic && born == 0 && tchanged == 0 && air_alive == 0) ? t + 1 : 0;
        c->born_last = born != 0 ? 1 : 0;
        // lava-front simulations stage exactly the tiles they visit (no scan)
        c->tiles_visited += sim.push ? staged : tiles;
        c->tiles_staged += staged;
        c->changes_total += changed;  // every change is one record
    }
    __syncthreads();  // red is reused by the caller's next block-wide step
}

// A tile of 64 x TY x TZ cells and its 1-cell halo, as staged in LDS by the descent and rules kernels:
// R = LY * LZ halo rows of LX = 66 positions.  Wave w stages rows w, w + 4, ... with lane lx at halo column
// lx + 1 (x = x0 + lx); threads 2r and 2r + 1 stage the two edge columns of row r.
template <int TY, int TZ>
struct Halo {
    static constexpr int LX = kTileX + 2, LY = TY + 2, LZ = TZ + 2, LXY = LX * LY, N = LXY * LZ;
    static constexpr int R = LY * LZ;                                   // halo rows
    static constexpr int KR = (R + kBlock / 64 - 1) / (kBlock / 64);    // rows per wave
    static_assert(2 * R <= kBlock, "edge columns: one thread each");
    static_assert(TY * TZ == 16, "four waves x four rows per tile");
};
constexpr uint8_t kVoidLds = 0xFF;  // halo position with no cell (absorbing edge, or past a partial tile)
constexpr uint32_t kNoCell = 0xFFFFFFFFu;

// Global index of x = 0 in halo row r of the tile at o, or kNoCell when the row holds no cells.  Rows past the
// far side of a partial tile (y > ny, ...) neighbour no cell and hold none either.
// Coordinate v in [-1, n] of a halo position: its wrapped value, or -1 when it holds no cell (absorbing axis,
// or past the far side of a partial tile).  Selects only.
__device__ inline int halo_coord(int v, int n, int per) {
    const bool in = static_cast<unsigned>(v) < static_cast<unsigned>(n);
    const bool ok = v <= n && (in || per);
    const int w = in ? v : (v < 0 ? v + n : v - n);
    return ok ? w : -1;
}
template <int TY, int TZ>
__device__ inline uint32_t halo_row(const CaParams& P, const TileCoord& o, int r) {
    const int gy = halo_coord(o.y0 - 1 + r % Halo<TY, TZ>::LY, P.ny, P.per[1]);
    const int gz = halo_coord(o.z0 - 1 + r / Halo<TY, TZ>::LY, P.nz, P.per[2]);
    const uint32_t base = static_cast<uint32_t>(P.nx) * static_cast<uint32_t>(gy) +
                          static_cast<uint32_t>(P.nx) * static_cast<uint32_t>(P.ny) * static_cast<uint32_t>(gz);
    return (gy < 0 || gz < 0) ? kNoCell : base;
}
// Global x of halo x position gx in [-1, nx], or -1 when it holds no cell.
__device__ inline int halo_x(const CaParams& P, int gx) { return halo_coord(gx, P.nx, P.per[0]); }

// Halo of one tile in LDS with per-row lava masks: bit lx of lava[r] = halo column lx + 1 of row r holds
// LAVA; edge[2r] / edge[2r + 1] = the left / right edge column of row r holds LAVA.
template <int TY, int TZ>
struct HaloTypes {
    uint8_t type[Halo<TY, TZ>::N];
    uint64_t lava[Halo<TY, TZ>::R];
    uint8_t edge[2 * Halo<TY, TZ>::R];
};

// Halo rows of the tile at o, one per lane: lane j < R holds the global index of x = 0 in halo row j, or 0 when the
// row holds no cells (so that every address formed from a row base lies inside the grid); bit j of `ok` is set when
// it holds cells.  Rows are then read with v_readlane (scalar row index) instead of each wave redoing the row
// arithmetic in scalar code.
struct HaloRows {
    uint32_t base;
    uint64_t ok;
};
template <int TY, int TZ>
__device__ inline HaloRows halo_rows(const CaParams& P, const TileCoord& o) {
    const int lane = static_cast<int>(opaque_lane());
    const uint32_t b = lane < Halo<TY, TZ>::R ? halo_row<TY, TZ>(P, o, lane) : kNoCell;
    return {b != kNoCell ? b : 0u, static_cast<uint64_t>(__ballot(b != kNoCell))};
}
__device__ inline uint32_t row_of(uint32_t row_base, int r) {
    return static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(row_base), r));
}
// Lane `lane` of a wave-uniform 64-bit value (ballot result) as this lane's bool: the mask becomes the exec mask
// of a branch or the lane mask of a select directly.
__device__ inline bool lane_bit(uint64_t m) { return __builtin_amdgcn_inverse_ballot_w64(m); }
// v with the wave-uniform `bits` in lane k (k a compile-time constant in the callers' unrolled loops; lane = this
// thread's lane): one compare and one select.
__device__ inline uint32_t put_lane(uint32_t v, uint32_t bits, int k, int lane) { return lane == k ? bits : v; }

// What a step writes into the other type copy for the own cells of a tile, as two bit planes of one 64-bit mask per
// tile row (row = ly + TY * lz, bit lx) in 32-bit words w[plane][2 * row + (lx >> 5)]: plane 0 = the cell is
// written, plane 1 = with BASALT (else LAVA).
template <int TY, int TZ>
struct NextTypes {
    uint32_t w[2][2 * TY * TZ];
};

// Stages the types of the tile at o and its halo, row by row, with the row lava masks, and (with_spill)
// the spill heights of the LX x LY halo columns, or with OWN of the tile's own 64 x TY columns only (index
// lx + 64 * ly; 0 where a column holds no cells).  Every load of a thread is issued before its first LDS store, so a
// tile costs one memory latency.  Returns whether this thread saw lava; own_air |= this thread saw AIR among the
// tile's own cells; lane k of lava_lo / lava_hi gets the low / high word of the lava mask of the wave's row k (halo
// row w + 4k of wave w).
// The wave's row tests are three scalar masks built once by its lanes (lane k: row k), and every row is loaded
// without a guard from a valid address (a row without cells reads row base 0; a position without a cell is voided by
// a select), so no row costs exec-mask changes.
// With SYNC, N gets the previous step's changes that the other type copy `other` does not hold yet: when `changed`
// (the tile changed a cell's type in the previous step) the tile's own cells are read from that copy as well, else
// the copies agree on them.
template <int TY, int TZ, bool OWN = false, bool SYNC = false>
__device__ inline bool stage_rows(const CaParams& P, const CaDev& D, const uint8_t* type, const uint8_t* other,
                                  bool changed, NextTypes<TY, TZ>* N, const TileCoord& o, HaloTypes<TY, TZ>& S,
                                  bool with_spill, int32_t* s_spill, int& own_air, const HaloRows& rows, int gx,
                                  unsigned tid, uint32_t& lava_lo, uint32_t& lava_hi) {
    using H = Halo<TY, TZ>;
    constexpr int NS = OWN ? kTileX * TY : H::LXY;     // spill columns staged
    constexpr int KS = (NS + kBlock - 1) / kBlock;     // per thread
    const int t = static_cast<int>(tid);
    const int w = __builtin_amdgcn_readfirstlane(t >> 6);
    __builtin_assume(w >= 0 && w < kBlock / 64);
    const int lx = t & 63;
    // lane k < KR describes the wave's row k: whether it is a halo row, holds cells, is one of the tile's own rows,
    // and (SYNC, after a step that changed a type) is an own row inside the grid in y and z; its own-row index
    const int rk = w + 4 * lx;
    const bool lrow = lx < H::KR && rk < H::R;
    const int oly = rk % H::LY - 1, olz = rk / H::LY - 1;
    const bool lown = lrow && oly >= 0 && oly < TY && olz >= 0 && olz < TZ;
    const uint32_t okk = static_cast<uint32_t>(__ballot(lrow && ((rows.ok >> (rk & 63)) & 1u)));
    const uint32_t ownk = static_cast<uint32_t>(__ballot(lown));
    const uint32_t insk =
        (SYNC && changed) ? static_cast<uint32_t>(__ballot(lown && o.y0 + oly < P.ny && o.z0 + olz < P.nz)) : 0u;
    const bool xin = gx >= 0;
    const uint32_t gxc = xin ? static_cast<uint32_t>(gx) : 0u;
    uint32_t v[H::KR];  // 32-bit so that no load result is packed (and waited for) before the stores
    uint32_t ov[H::KR];
#pragma unroll
    for (int k = 0; k < H::KR; ++k) {
        v[k] = kVoidLds;
        ov[k] = kVoidLds;
        // lanes R .. 63 of rows.base hold 0 as well: a row index past R still reads a valid address
        if (4 * k + 3 < H::R || w + 4 * k < H::R) v[k] = type[row_of(rows.base, w + 4 * k) + gxc];
    }
    uint32_t e = kVoidLds;
    {
        // thread t < 2R: edge column (t & 1) of row t >> 1 (rows 0..31 in wave 0, 32..35 in wave 1)
        const int er = (t >> 1) & 63;
        const uint32_t base = static_cast<uint32_t>(__builtin_amdgcn_ds_bpermute(er << 2, static_cast<int>(rows.base)));
        if (t < 2 * H::R) {
            const int ex = halo_x(P, (t & 1) ? o.x0 + kTileX : o.x0 - 1);
            if (((rows.ok >> er) & 1u) && ex >= 0) e = type[base + static_cast<uint32_t>(ex)];
        }
    }
    int32_t sp[KS];
    if (with_spill) {
#pragma unroll
        for (int k = 0; k < KS; ++k) {
            const int q = t + k * kBlock;
            sp[k] = 0;
            if (q < NS) {
                if (OWN) {
                    const int sx = o.x0 + q % kTileX, sy = o.y0 + q / kTileX;
                    if (sx < P.nx && sy < P.ny)
                        sp[k] = D.spill[static_cast<uint32_t>(sx) + static_cast<uint32_t>(P.nx) * static_cast<uint32_t>(sy)];
                } else {
                    int sx = o.x0 - 1 + q % H::LX, sy = o.y0 - 1 + q / H::LX;
                    if (sx <= P.nx && sy <= P.ny && wrap_axis(sx, P.nx, P.per[0]) && wrap_axis(sy, P.ny, P.per[1]))
                        sp[k] = D.spill[static_cast<uint32_t>(sx) +
                                        static_cast<uint32_t>(P.nx) * static_cast<uint32_t>(sy)];
                }
            }
        }
    }
    if (SYNC && insk) {
#pragma unroll
        for (int k = 0; k < H::KR; ++k)
            if ((insk >> k) & 1u) ov[k] = other[row_of(rows.base, w + 4 * k) + gxc];
    }
    // lava rows OR-ed as masks; each row's masks go into lane k of the wave's mask registers
    const uint64_t xm = __ballot(xin);
    const uint64_t inx = __ballot(o.x0 + lx < P.nx);
    uint64_t any_lava = 0, air = 0;
    uint32_t lv_lo = 0, lv_hi = 0, n0_lo = 0, n0_hi = 0, n1_lo = 0, n1_hi = 0;
#pragma unroll
    for (int k = 0; k < H::KR; ++k) {
        const int r = w + 4 * k;
        if (4 * k + 3 < H::R || r < H::R) {
            const uint32_t val = lane_bit(((okk >> k) & 1u) ? xm : 0ull) ? v[k] : kVoidLds;
            S.type[lx + 1 + H::LX * r] = static_cast<uint8_t>(val);
            const uint64_t m = __ballot(val == kLava);
            lv_lo = put_lane(lv_lo, static_cast<uint32_t>(m), k, lx);
            lv_hi = put_lane(lv_hi, static_cast<uint32_t>(m >> 32), k, lx);
            any_lava |= m;
            if ((ownk >> k) & 1u) {
                air |= __ballot(val == kAir) & inx;
                if (SYNC && ((insk >> k) & 1u)) {
                    // own row inside the grid: the cells born (LAVA) or solidified (BASALT) in step t - 1; rows past
                    // the grid's far side hold wrapped halo cells
                    const uint64_t d = __ballot(val != ov[k]) & inx;
                    const uint64_t b = d & __ballot(val == kBasalt);
                    n0_lo = put_lane(n0_lo, static_cast<uint32_t>(d), k, lx);
                    n0_hi = put_lane(n0_hi, static_cast<uint32_t>(d >> 32), k, lx);
                    n1_lo = put_lane(n1_lo, static_cast<uint32_t>(b), k, lx);
                    n1_hi = put_lane(n1_hi, static_cast<uint32_t>(b >> 32), k, lx);
                }
            }
        }
    }
    if (lrow) S.lava[rk] = (static_cast<uint64_t>(lv_hi) << 32) | lv_lo;
    if (SYNC && lown) {
        const int row = oly + TY * olz;
        N->w[0][2 * row] = n0_lo;
        N->w[0][2 * row + 1] = n0_hi;
        N->w[1][2 * row] = n1_lo;
        N->w[1][2 * row + 1] = n1_hi;
    }
    if (air) own_air = 1;
    lava_lo = lv_lo;
    lava_hi = lv_hi;
    bool lava = any_lava != 0;
    if (t < 2 * H::R) {
        S.type[((t & 1) ? H::LX - 1 : 0) + H::LX * (t >> 1)] = static_cast<uint8_t>(e);
        S.edge[t] = e == kLava ? 1 : 0;
        lava = lava || e == kLava;
    }
    if (with_spill) {
#pragma unroll
        for (int k = 0; k < KS; ++k) {
            const int q = t + k * kBlock;
            if (q < NS) s_spill[q] = sp[k];
        }
    }
    return lava;
}

// dropDepth (LavaCA.hpp:404-414) of the AIR cell at halo position c, which sits in halo layer lz + 1 of the
// column col (global) with spill height spill_c.  Layers held in LDS are read there, deeper ones from global
// memory (one load per cell: reading several cells ahead measured slower, deep drops are rare).
template <int TY, int TZ>
__device__ int drop_depth_lds(const uint8_t* type, const uint8_t* s_type, int c, int lz, int z, uint32_t col,
                              uint32_t plane, int spill_c, bool has_spill) {
    constexpr int LXY = Halo<TY, TZ>::LXY;
    int zc = z;
    #pragma clang loop unroll(disable)
    for (int d = 0; d <= kMaxDrop; ++d, --zc) {
        // type(zc) is AIR here: the cell itself when d == 0, checked one iteration earlier otherwise
        if (zc == 0) return d;  // hasSupport: the floor
        const int l = lz - d;   // halo layer of zc - 1
        const uint8_t b = l >= 0 ? s_type[c - (d + 1) * LXY] : type[col + plane * static_cast<uint32_t>(zc - 1)];
        if (b == kRock || b == kBasalt) return d;
        if (b == kLava && has_spill && zc - 1 < spill_c) return d;
        if (b != kAir) return d + 1 <= kMaxDrop ? -1 : kMaxDrop;
    }
    return kMaxDrop;
}

// D(n) of every lava cell of the tiles holding lava (LavaCA.hpp:421-429, 877-888), from the LDS-staged tile:
// hasSupport of the cell, then the deepest dropDepth over its AIR side neighbours (union of the variants).
template <int TY, int TZ>
// The first kernel of a step: it clears the counts of the list copy the previous step read, does the previous step's
// finish in its first `fin_blocks` blocks (one slot each; 0 when the finish is not pending) and computes D(n) of the
// staged tiles of parity par that hold lava in the blocks after them, alongside the finish.
// With desc, the tile blocks walk the step's descent list (the tiles whose D(n) may have changed: each only checks
// that it holds lava), else every listed tile with the check of the tiles its D(n) reads.
__global__ void __launch_bounds__(kBlock) ca_descent_kernel(BatchDev B, int shape, int par, uint32_t fin_blocks,
                                                            int desc) {
    using H = Halo<TY, TZ>;
    __shared__ HaloTypes<TY, TZ> S;
    __shared__ int32_t s_spill[H::LXY];
    __shared__ uint16_t s_list[kTileCells];
    __shared__ uint32_t s_deep[kTileCells];  // lava cells with support and AIR side neighbours: cell | mask << 16
    __shared__ int32_t s_dd[kTileCells];     // their D(n), folded over the neighbours
    __shared__ uint32_t s_n, s_nd;
    const uint8_t* s_type = S.type;
    if (blockIdx.x == fin_blocks && threadIdx.x < kShapes * kSub) {
        // every kernel of the previous step is done with its list copy; the rules kernel of this step pushes the next
        // step's tiles there
        *list_count(B, par ^ 1, threadIdx.x / kSub, threadIdx.x % kSub) = 0u;
        if (B.desc_list && threadIdx.x < kShapes) *desc_count(B, par ^ 1, static_cast<int>(threadIdx.x)) = 0u;
    }
    if (blockIdx.x < fin_blocks) {
        finish_slot(B, blockIdx.x);
        return;
    }
    // tile block b walks sub-list b % kSub (the tile blocks are a multiple of kSub), or with desc the descent list
    const uint32_t tb = blockIdx.x - fin_blocks;
    const uint32_t stride = desc ? gridDim.x - fin_blocks : (gridDim.x - fin_blocks) / kSub;
    const uint32_t n = desc ? *desc_count(B, par, shape) : *list_count(B, par, shape, static_cast<int>(tb % kSub));
    const uint64_t* tiles = desc ? B.desc_list + static_cast<size_t>(par * kShapes + shape) * B.subcap
                                 : sub_list(B, par, shape, static_cast<int>(tb % kSub));
    // the block's first entry is read together with the list's length, each later one an iteration ahead
    const uint32_t s0 = desc ? tb : tb / kSub;
    uint64_t next = s0 < B.subcap ? tiles[s0] : 0ull;
    for (uint32_t s = s0; s < n; s += stride) {
        // uniform: one entry per block; low word slot << kTileBits | tile, high word the simulation's step
        const uint64_t ent = next;
        if (s + stride < n) next = tiles[s + stride];
        const uint32_t e = __builtin_amdgcn_readfirstlane(static_cast<uint32_t>(ent));
        const uint32_t now = __builtin_amdgcn_readfirstlane(static_cast<uint32_t>(ent >> 32));
        const CaSim& sim = g_sims[B.sim_base + (e >> kTileBits)];
        const CaParams& P = sim.P;
        const CaDev& D = sim.D;
        const uint32_t plane = static_cast<uint32_t>(P.nx) * static_cast<uint32_t>(P.ny);
        const bool has_spill = D.spill != nullptr;
        const uint32_t side = P.nb.side_mask;
        const uint32_t tile = e & kTileMask;
        const TileCoord o = tile_origin(P, tile);
        // D(n) of the tile's lava cells reads the types of the tile, its 8 horizontal neighbours and the tiles under
        // them down to kMaxDrop + 1 cells below the tile's lowest layer (support, side cells, their drop chains; spill
        // levels do not change).  When none of those tiles changed a cell's type in the previous step, the values
        // written then still hold.  Only tiles holding lava have D(n) to compute, and entries of a slot that left the
        // batch or of a simulation that stopped are skipped.
        uint32_t stop = 0, flags = kFlagLava;
        if (desc) {
            // a descent-list entry needs D(n) unless its simulation stopped or the tile holds no lava (rare: the list
            // takes lava tiles only): both flags are read together with the staging loads, which are safe for any
            // running simulation
            __syncthreads();  // the previous tile's LDS reads are done
            if (!sim.running || P.classic) continue;
            stop = D.ctrl->stop;
            flags = D.tile_flags[tile];
        } else {
            constexpr int NR = 9 * (kDescentLayers<TZ> + 1);
            const int tx = o.x0 / kTileX, ty = o.y0 / TY, tz = o.z0 / TZ;
            // the stop flag, the tile's flags and the marks of the tiles below are read together (the barrier also
            // orders the previous tile's LDS reads: every tile passes it)
            const bool go = sim.running && !P.classic && !__builtin_amdgcn_readfirstlane(D.ctrl->stop) &&
                            (__builtin_amdgcn_readfirstlane(D.tile_flags[tile]) & kFlagLava);
            bool dirty = false;
            for (int q = static_cast<int>(threadIdx.x); sim.running && !P.classic && q < NR; q += kBlock) {
                int qx = tx + q % 3 - 1, qy = ty + (q / 3) % 3 - 1;
                const int qz = tz - q / 9;
                if (qz >= 0 && wrap_axis(qx, P.ntx, P.per[0]) && wrap_axis(qy, P.nty, P.per[1]))
                    dirty = dirty || D.tile_tchg[static_cast<uint32_t>(qx) +
                                                 static_cast<uint32_t>(P.ntx) *
                                                     (static_cast<uint32_t>(qy) +
                                                      static_cast<uint32_t>(P.nty) * static_cast<uint32_t>(qz))] == now;
            }
            if (!__syncthreads_or(go && dirty)) continue;
        }
        int own_air = 0;
        const HaloRows rows = halo_rows<TY, TZ>(P, o);
        const int gx = halo_x(P, o.x0 + static_cast<int>(threadIdx.x & 63u));
        const uint8_t* type = D.type[now & 1u];
        uint32_t lava_lo, lava_hi;
        stage_rows<TY, TZ>(P, D, type, nullptr, false, static_cast<NextTypes<TY, TZ>*>(nullptr), o, S, has_spill,
                           s_spill, own_air, rows, gx, opaque_tid(), lava_lo, lava_hi);
        if (__builtin_amdgcn_readfirstlane(stop) || !(__builtin_amdgcn_readfirstlane(flags) & kFlagLava)) continue;
        if (threadIdx.x == 0) {
            s_n = 0;
            s_nd = 0;
        }
        __syncthreads();
        {
            // the tile's lava cells into a list (tile-local index lx + 64 * row)
            const int lx = static_cast<int>(threadIdx.x & 63u);
            const int w = static_cast<int>(threadIdx.x >> 6);
            bool lava[4];
            uint32_t slot[4];
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                const int row = w + 4 * k;
                const int rc = (row % TY + 1) + H::LY * (row / TY + 1);
                // positions past the grid's far side hold wrapped halo cells of a periodic partial tile
                const bool inside = o.x0 + lx < P.nx && o.y0 + row % TY < P.ny && o.z0 + row / TY < P.nz;
                lava[k] = inside && ((S.lava[rc] >> lx) & 1u);
            }
            wave_append_k(lava, &s_n, slot);
#pragma unroll
            for (int k = 0; k < 4; ++k)
                if (lava[k]) s_list[slot[k]] = static_cast<uint16_t>(lx + 64 * (w + 4 * k));
        }
        __syncthreads();
        const uint32_t nl = s_n;
        // D(n) = -1 without support, else the deepest dropDepth over the AIR side neighbours (0 without one).  Cells
        // with such neighbours go into a second list, whose (cell, neighbour) pairs are then spread over the threads
        // (one drop chain each instead of up to eight in a row) and folded with LDS maxima.
        auto cell_index = [&](int e) {
            const int lx = e & 63, row = e >> 6;
            return col_index(P, o.x0 + lx, o.y0 + row % TY) + plane * static_cast<uint32_t>(o.z0 + row / TY);
        };
        for (uint32_t j0 = 0; j0 < nl; j0 += kBlock) {
            const uint32_t j = j0 + threadIdx.x;
            uint32_t m = 0;
            bool deep = false;
            if (j < nl) {
                const int e = s_list[j];
                const int lx = e & 63, row = e >> 6;
                const int ly = row % TY, lz = row / TY;
                const int z = o.z0 + lz;
                const int c = (lx + 1) + H::LX * ((ly + 1) + H::LY * (lz + 1));
                bool sup = z == 0;  // hasSupport (LavaCA.hpp:390-399)
                if (!sup) {
                    const uint8_t b = s_type[c - H::LXY];
                    sup = b == kRock || b == kBasalt || (b == kLava && has_spill && z - 1 < s_spill[(lx + 1) + H::LX * (ly + 1)]);
                }
                if (sup) {
#pragma unroll
                    for (int i = 0; i < 8; ++i) {  // the dz == 0 positions q = 9 .. 17 but 13
                        const int q = 9 + i + (i >= 4 ? 1 : 0);
                        const int nb = c + (q % 3 - 1) + H::LX * ((q / 3) % 3 - 1);
                        // VOID (no cell) is not AIR
                        if (((side >> q) & 1u) && s_type[nb] == kAir) m |= 1u << i;
                    }
                }
                deep = m != 0;
                if (!deep) D.descent[cell_index(e)] = static_cast<int8_t>(sup ? 0 : -1);
                m = static_cast<uint32_t>(e) | (m << 16);
            }
            const uint32_t at = wave_append(deep, &s_nd);
            if (deep) {
                s_deep[at] = m;
                s_dd[at] = 0;
            }
        }
        __syncthreads();
        const uint32_t nd = s_nd;
        for (uint32_t p = threadIdx.x; p < 8 * nd; p += kBlock) {
            const uint32_t v = s_deep[p >> 3];
            const int i = static_cast<int>(p & 7u);
            if (!((v >> (16 + i)) & 1u)) continue;
            const int e = static_cast<int>(v & 0xFFFFu);
            const int lx = e & 63, row = e >> 6;
            const int ly = row % TY, lz = row / TY;
            const int x = o.x0 + lx, y = o.y0 + ly, z = o.z0 + lz;
            const int c = (lx + 1) + H::LX * ((ly + 1) + H::LY * (lz + 1));
            const int sc = (lx + 1) + H::LX * (ly + 1);
            const int q = 9 + i + (i >= 4 ? 1 : 0);
            const int dx = q % 3 - 1, dy = (q / 3) % 3 - 1;
            int qx = x + dx, qy = y + dy;
            wrap_axis(qx, P.nx, P.per[0]);
            wrap_axis(qy, P.ny, P.per[1]);
            const uint32_t qc = static_cast<uint32_t>(qx) + static_cast<uint32_t>(P.nx) * static_cast<uint32_t>(qy);
            const int dd = drop_depth_lds<TY, TZ>(type, s_type, c + dx + H::LX * dy, lz, z, qc, plane,
                                                  has_spill ? s_spill[sc + dx + H::LX * dy] : 0, has_spill);
            if (dd > 0) atomicMax(&s_dd[p >> 3], dd);
        }
        __syncthreads();
        for (uint32_t j = threadIdx.x; j < nd; j += kBlock)
            D.descent[cell_index(static_cast<int>(s_deep[j] & 0xFFFFu))] = static_cast<int8_t>(s_dd[j]);
    }
}

// Tile of the rules kernel in LDS: types of every halo position, flowId / temperature / D(n) of the LAVA
// positions, spill heights of the tile's own columns.
// F: the LDS type of flowIds (int16_t when every flowId of the batch fits, which keeps six tiles per CU)
template <int TY, int TZ, class F>
struct RulesTile {
    using H = Halo<TY, TZ>;
    HaloTypes<TY, TZ> S;
    uint16_t list[kTileCells];  // candidate cells of the tile (tile-local index lx + 64 * row)
    int8_t desc[H::N];
    F flow[H::N];
    float temp[H::N];
    int32_t spill[kTileX * TY];  // spill heights of the tile's own columns (lx + 64 * ly)
    // LAVA bits of halo row r over all LX = 66 positions (bit p = halo column p), as three 32-bit parts + 0
    uint32_t row66[H::R][4];
    // what the step writes into the other type copy (see NextTypes): the previous step's changes (staging), then this
    // step's on top
    NextTypes<TY, TZ> next;
};

// This step's change of own cell e (tile-local lx + 64 * row) into N (born: LAVA, else BASALT; a cell that
// solidifies may have been born in the previous step: BASALT wins).  LDS atomics: other cells share the words.
template <int TY, int TZ>
__device__ inline void next_type(NextTypes<TY, TZ>& N, int e, bool born) {
    const int q = 2 * (e >> 6) + ((e >> 5) & 1);
    const uint32_t bit = 1u << (e & 31);
    atomicOr(&N.w[0][q], bit);
    if (!born) atomicOr(&N.w[1][q], bit);
}

// Writes the tile's cells marked in N into the next step's type copy: after them the copy holds the tile's types
// after the step.  Thread tid takes cells tid + 256 k (row w + 4k, column lx), the cells it built candidates for.
template <int TY, int TZ>
__device__ inline void write_types(const CaParams& P, const TileCoord& o, const NextTypes<TY, TZ>& N, uint8_t* tyn,
                                   unsigned tid) {
    const int lx = static_cast<int>(tid & 63u), w = static_cast<int>(tid >> 6);
#pragma unroll
    for (int k = 0; k < kTileCells / kBlock; ++k) {
        const int row = w + 4 * k;
        const int q = 2 * row + (lx >> 5);
        if ((N.w[0][q] >> (lx & 31)) & 1u) {
            const uint32_t i = static_cast<uint32_t>(o.x0 + lx) +
                               static_cast<uint32_t>(P.nx) * (static_cast<uint32_t>(o.y0 + row % TY) +
                                                              static_cast<uint32_t>(P.ny) *
                                                                  static_cast<uint32_t>(o.z0 + row / TY));
            tyn[i] = ((N.w[1][q] >> (lx & 31)) & 1u) ? kBasalt : kLava;
        }
    }
}

// gatherLavaCandidates (LavaCA.hpp:465-519) / the classic vote (LavaCA.hpp:1165-1202) for the AIR cell at LDS
// position c; bits = the cell's neighbour positions (k as in vca_internal.h) that hold LAVA and that the rules
// read.  Only ponding below the staged layers and deep drop depths touch global memory.
// LDS offset of neighbour position k (< 27; dx = k % 3 - 1, dy = k / 3 % 3 - 1, dz = k / 9 - 1) in a halo of rows
// of LX positions and layers of LXY: k / 9 and r / 3 (r < 9) by multiply-shift.
template <int LX, int LXY>
__device__ inline int koff(int k) {
    const int dz = (k * 57) >> 9;
    const int r = k - 9 * dz;
    const int dy = (r * 11) >> 5;
    const int dx = r - 3 * dy;
    return dx + LX * dy + LXY * dz - (1 + LX + LXY);
}

template <int TY, int TZ, class F, class V>
__device__ void gather_lds(const CaParams& P, const uint8_t* type, const RulesTile<TY, TZ, F>& T, int c, int lz, int z,
                           uint32_t col, uint32_t plane, int spill_c, bool has_spill, uint32_t bits, V& v) {
    constexpr int LX = Halo<TY, TZ>::LX, LXY = Halo<TY, TZ>::LXY;
    // positions are taken two at a time so that their LDS reads are in flight together (the vote does not depend
    // on the order of its candidates)
    if (P.classic) {
        #pragma clang loop unroll(disable)
        while (bits) {
            const int k1 = __builtin_ctz(bits);
            bits &= bits - 1;
            const bool two = bits != 0;
            const int k2 = two ? __builtin_ctz(bits) : k1;
            bits &= bits - 1;
            const int32_t f1 = T.flow[c + koff<LX, LXY>(k1)], f2 = T.flow[c + koff<LX, LXY>(k2)];
            v.add_if(true, f1, 0.0f, k1);
            v.add_if(two, f2, 0.0f, k2);
        }
        return;
    }
    if (bits & (1u << 4)) {
        // tryPonding (LavaCA.hpp:443-461) into the cell from directly below, the only dz < 0 position in the
        // lava-mode mask
        bits &= ~(1u << 4);
        const int n = c - LXY;
        bool ok = !(z > spill_c || z == 0);
        if (ok && z - 1 > 0) {  // hasSupport(z - 1)
            const uint8_t b2 = lz >= 1 ? T.S.type[c - 2 * LXY] : type[col + plane * static_cast<uint32_t>(z - 2)];
            ok = b2 == kRock || b2 == kBasalt || (b2 == kLava && has_spill && z - 2 < spill_c);
        }
        if (ok) {
            const uint8_t h0 = T.S.type[n - 1], h1 = T.S.type[n + 1];
            bool closed = h0 != kAir && h0 != kVoidLds && h1 != kAir && h1 != kVoidLds;
            if (P.ny > 1) {
                const uint8_t h2 = T.S.type[n - LX], h3 = T.S.type[n + LX];
                closed = closed && h2 != kAir && h2 != kVoidLds && h3 != kAir && h3 != kVoidLds;
            }
            ok = closed;
        }
        v.add_if(ok, T.flow[n], T.temp[n] - P.cool_lat, 4);
    }
    // dropDepth of the cell, once, when a side or diagonal-up neighbour may need it (pure function of the cell)
    const int cdrop = (bits & ~(1u << 22)) ? drop_depth_lds<TY, TZ>(type, T.S.type, c, lz, z, col, plane, spill_c, has_spill)
                                           : -100;
    const bool diag = P.nb.diagonal_falls != 0;
    // k >= 18: dz > 0 (k == 22 directly above, a free fall); else dz == 0
    // (bitwise, not short-circuit: selects instead of exec-mask branches)
    auto accept = [&](int k, int d) {
        const bool up_ok = (cdrop == 0) & (d == 1);
        const bool side_ok = (cdrop == d) & !((d == 1) & diag);
        return (k == 22) | ((d >= 0) & (k >= 18 ? up_ok : side_ok));
    };
    #pragma clang loop unroll(disable)
    while (bits) {
        const int k1 = __builtin_ctz(bits);
        bits &= bits - 1;
        const bool two = bits != 0;
        const int k2 = two ? __builtin_ctz(bits) : k1;
        bits &= bits - 1;
        const int n1 = c + koff<LX, LXY>(k1), n2 = c + koff<LX, LXY>(k2);
        const int d1 = T.desc[n1], d2 = T.desc[n2];
        const int32_t f1 = T.flow[n1], f2 = T.flow[n2];
        const float t1 = T.temp[n1], t2 = T.temp[n2];
        v.add_if(accept(k1, d1), f1, t1 - (k1 >= 18 ? P.cool_fall : P.cool_lat), k1);
        v.add_if(two && accept(k2, d2), f2, t2 - (k2 >= 18 ? P.cool_fall : P.cool_lat), k2);
    }
}

// CandidateVote (engine/src/ca/CaRules.hpp:127-183) over any number of flowIds, for a cell whose candidates carry three
// or more (rare: where three flows meet).  acc = the accepted candidate positions; the candidate at position k carries
// the flowId and temperature of the LDS position c + koff(k), minus the cooling of its move (falls from above:
// coolingFall; side, diagonal-down and ponding moves: coolingLateral; classic mode: temperature 0).  Most candidates
// wins; a tie picks by the hash among the tied flowIds in ascending order; the temperature is the largest of the
// winner's candidates.  Loops over the positions instead of per-flowId tables (no private arrays).
template <int TY, int TZ, class F>
__device__ void vote_positions(const CaParams& P, const RulesTile<TY, TZ, F>& T, int c, uint32_t acc, uint64_t h_index,
                               int32_t& wf, float& wt) {
    constexpr int LX = Halo<TY, TZ>::LX, LXY = Halo<TY, TZ>::LXY;
    auto flow_at = [&](int k) { return static_cast<int32_t>(T.flow[c + koff<LX, LXY>(k)]); };
    auto count_of = [&](int32_t f) {
        int n = 0;
        for (uint32_t m = acc; m; m &= m - 1) n += flow_at(__builtin_ctz(m)) == f ? 1 : 0;
        return n;
    };
    // a position is a flowId's first when no lower accepted position carries the same flowId
    auto first_of_flow = [&](int k) {
        const int32_t f = flow_at(k);
        for (uint32_t m = acc & ((1u << k) - 1u); m; m &= m - 1)
            if (flow_at(__builtin_ctz(m)) == f) return false;
        return true;
    };
    int best = 0;
    for (uint32_t m = acc; m; m &= m - 1) {
        const int n = count_of(flow_at(__builtin_ctz(m)));
        best = n > best ? n : best;
    }
    uint32_t tied = 0;  // first positions of the flowIds with the most candidates
    for (uint32_t m = acc; m; m &= m - 1) {
        const int k = __builtin_ctz(m);
        if (first_of_flow(k) && count_of(flow_at(k)) == best) tied |= 1u << k;
    }
    const uint32_t ntied = static_cast<uint32_t>(__builtin_popcount(tied));
    const uint32_t pick = ntied > 1 ? mod_small(splitmix64(h_index ^ kSaltTieBreak), ntied) : 0u;
    int32_t win = 0;
    for (uint32_t m = tied; m; m &= m - 1) {
        const int32_t f = flow_at(__builtin_ctz(m));
        uint32_t rank = 0;  // tied flowIds below f
        for (uint32_t q = tied; q; q &= q - 1) rank += flow_at(__builtin_ctz(q)) < f ? 1u : 0u;
        if (rank == pick) win = f;
    }
    float mt = -__builtin_inff();
    for (uint32_t m = acc; m; m &= m - 1) {
        const int k = __builtin_ctz(m);
        if (flow_at(k) != win) continue;
        const float t = P.classic ? 0.0f : T.temp[c + koff<LX, LXY>(k)] - (k >= 18 ? P.cool_fall : P.cool_lat);
        mt = mt < t ? t : mt;  // std::max in the order of the positions
    }
    wf = win;
    wt = mt;
}

// Rules of every cell of the visited tiles (LavaCA.hpp:963-1048, 1109-1202).  The tile and a 1-cell halo
// are staged in LDS first (types for every position, then flowId / temperature / D(n) for LAVA positions,
// all independent coalesced loads), so an AIR cell folds its 26 neighbour types into a bitmask without
// branches and gathers its candidates from LDS.  Cells whose candidates carry more than one flowId are gathered
// again with two vote slots, three or more flowIds with the full vote (vote_positions).
// Register budget: as many waves per SIMD as tiles per CU fit in LDS (the 4 x 4 tile with 16-bit flowIds: six,
// 80 registers; the others five at most, 96 registers) -- more tiles in flight hide the staging latency better
// than the few registers the limit costs.
template <int TY, int TZ, class F>
constexpr int kRulesTilesPerCu = static_cast<int>(163840 / (sizeof(RulesTile<TY, TZ, F>) + 512));  // 160 KB of LDS
template <int TY, int TZ, class F>
constexpr int kRulesWaves = (TY == 4 && TZ == 4 && sizeof(F) == 2)
                                ? 6
                                : (kRulesTilesPerCu<TY, TZ, F> < 5 ? kRulesTilesPerCu<TY, TZ, F> : 5);

template <int TY, int TZ, class F>
__global__ void __launch_bounds__(kBlock, (kRulesWaves<TY, TZ, F>)) ca_rules_kernel(BatchDev B, int shape, int par) {
    using H = Halo<TY, TZ>;
    constexpr int LX = H::LX, LY = H::LY;
    __shared__ RulesTile<TY, TZ, F> T;
    __shared__ uint32_t s_nlist, s_nair;
    __shared__ uint32_t s_red[6][kBlock / 64];
    __shared__ uint32_t s_wflags[2];  // per-wave flag bytes: [0] staging, [1] end of tile
    __shared__ uint32_t s_bnd0[kBlock / 64];  // per wave: directions with LAVA on the tile's boundary at the step's start
    // block b walks sub-list b % kSub (the grid is a multiple of kSub); its first entry is read together with the
    // sub-list's length
    const uint32_t stride = gridDim.x / kSub, s0 = blockIdx.x / kSub;
    const uint64_t* tiles = sub_list(B, par, shape, static_cast<int>(blockIdx.x % kSub));
    uint64_t next = s0 < B.subcap ? tiles[s0] : 0ull;
    const uint32_t nact = *list_count(B, par, shape, static_cast<int>(blockIdx.x % kSub));
    // the wave's first thread, once (a scalar): inside the loop the thread index is rebuilt from it and the lane
    // (see opaque_tid), so no vector register holds threadIdx.x across the loop
    const unsigned wbase = static_cast<unsigned>(__builtin_amdgcn_readfirstlane(static_cast<int>(threadIdx.x))) & ~63u;
    for (uint32_t s = s0; s < nact; s += stride) {
        const unsigned tid = wbase | opaque_lane();
        // uniform: one entry per block; low word slot << kTileBits | tile, high word the simulation's step
        const uint32_t e = __builtin_amdgcn_readfirstlane(static_cast<uint32_t>(next));
        const int32_t t = __builtin_amdgcn_readfirstlane(static_cast<int32_t>(next >> 32));
        if (s + stride < nact) next = tiles[s + stride];
        const CaSim& sim = g_sims[B.sim_base + (e >> kTileBits)];
        const CaParams& P = sim.P;
        const CaDev& D = sim.D;
        const uint32_t tile = e & kTileMask;
        // hashPrefix(seed, t) (Rng.hpp:66-72), kept in the simulation's control block by the kernel that advanced it
        // to step t; read with the stop flag
        const uint64_t prefix = D.ctrl->prefix;
        const bool erupting = t < P.eruption_steps;
        const uint16_t stamp = static_cast<uint16_t>(t < 65535 ? t : 65535);
        const uint32_t plane = static_cast<uint32_t>(P.nx) * static_cast<uint32_t>(P.ny);
        const bool has_spill = !P.classic && D.spill != nullptr;
        // positions the rules read as lava sources: dz >= 0, plus directly below for ponding (lava mode)
        const uint32_t rules_mask =
            P.classic ? P.nb.any_mask
                      : ((P.nb.any_mask & ~0x1FFu) | (P.ponding ? (P.nb.any_mask & (1u << 4)) : 0u));
        uint32_t w_lava = 0, w_born = 0;  // wave-uniform counts: lava cells after the step, born cells
        uint32_t w_chg = 0;               // wave-uniform count: cells that changed type
        uint32_t w_temp = 0;              // wave-uniform count: lava cells whose temperature changed (no record)
        uint32_t w_alive = 0;             // wave-uniform count: AIR candidates not born that a later step could bear
        bool w_tchg = false;              // wave-uniform: a cell changed type
        const float* tcur = D.temp[t & 1];
        float* tnext = D.temp[(t + 1) & 1];
        const uint8_t* tyc = D.type[t & 1];
        uint8_t* tyn = D.type[(t + 1) & 1];
        const TileCoord o = tile_origin(P, tile);
        __syncthreads();  // the previous tile's LDS reads are done
        // an entry of a slot that left the batch, or of a simulation that stopped after the entry was pushed (its
        // state is final: the step is not run)
        if (!sim.running) continue;
        // the stop flag is read together with the staging loads (staging writes LDS only); a stopped simulation's
        // tile goes no further
        const int32_t stop = D.ctrl->stop;
        // cells were born in the last step: no quiescence to find in this one (every AIR candidate counts as alive)
        const int32_t born_last = D.ctrl->born_last;
        int own_air = 0;
        // lane j: base of halo row j; gx: this lane's column (x == nx is the wrapped x = 0 of a periodic partial
        // tile, -1 for no cell)
        const HaloRows rows = halo_rows<TY, TZ>(P, o);
        const int gx = halo_x(P, o.x0 + static_cast<int>(tid & 63u));
        // the tile changed a cell's type in the previous step: the other type copy is behind on it
        const bool changed = t > 0 && __builtin_amdgcn_readfirstlane(D.tile_tchg[tile]) == static_cast<uint32_t>(t);
        // a lava cell of the tile may be its flow's vent (the eruption rule's test needs the vent table)
        const bool vent_tile = erupting && __builtin_amdgcn_readfirstlane(D.tile_vent[tile]) != 0;
        uint32_t lava_lo, lava_hi;  // lane k: lava mask of the wave's halo row k
        const bool saw_lava = stage_rows<TY, TZ, true, true>(P, D, tyc, tyn, changed, &T.next, o, T.S, has_spill,
                                                             T.spill, own_air, rows, gx, tid, lava_lo, lava_hi);
        if (__builtin_amdgcn_readfirstlane(stop)) {
            // a held simulation keeps its tiles listed for the step it is at (the fast-forward reads them there)
            if (__builtin_amdgcn_readfirstlane(stop) == kStopHeld && tid < 64u)
                wave_list_append(B, par ^ 1, shape, (tid & 63u) == 0u, tile, (static_cast<uint64_t>(t) << 32) | e);
            continue;
        }
        if (tid == 0) {
            const uint32_t zero = opaque_zero();
            s_nlist = zero;
            s_nair = zero;
        }
        {
            // block OR of saw_lava (bit 0) and own_air (bit 1): one flag byte per wave
            const uint32_t f = (__ballot(saw_lava) ? 1u : 0u) | (__ballot(own_air) ? 2u : 0u);
            if ((tid & 63u) == 0)
                reinterpret_cast<uint8_t*>(&s_wflags[0])[tid >> 6] = static_cast<uint8_t>(f);
        }
        __syncthreads();
        const uint32_t stage_flags = s_wflags[0];
        if (!(stage_flags & 0x01010101u)) {
            // no lava in the tile or its halo: no cell changes; the tile keeps AIR iff it has AIR, and the next type
            // copy takes the previous step's changes
            const int any_air = (stage_flags & 0x02020202u) != 0;
            if (changed) write_types<TY, TZ>(P, o, T.next, tyn, tid);
            if (tid == 0) {
                D.tile_flags[tile] = static_cast<uint8_t>(any_air ? kFlagAir : 0);
                atomicAdd(&D.part[tile % kPartRows].staged, 1ULL);
            }
            continue;
        }
        int tile_lava = 0, tile_air = 0;
        {
            // flowId / temperature / D(n) of the LAVA positions, all loads in flight before the stores; each row's
            // lava mask (from the staging, in lane k) is the exec mask of its loads, and every row base is a valid
            // index, so no mask bit can address memory outside the grid
            const int w = __builtin_amdgcn_readfirstlane(static_cast<int>(tid >> 6));
            __builtin_assume(w >= 0 && w < kBlock / 64);
            const int lx = static_cast<int>(tid & 63u);
            const int t = static_cast<int>(tid);
            const uint32_t gxc = gx >= 0 ? static_cast<uint32_t>(gx) : 0u;
            int32_t fl[H::KR];
            float tp[H::KR];
            int32_t ds[H::KR];  // 32-bit: see stage_rows
            const global_ptr<const int32_t> flow_g = pinned(static_cast<const int32_t*>(D.flow));
            const global_ptr<const float> temp_g = pinned(tcur);
            const global_ptr<const int8_t> desc_g = pinned(static_cast<const int8_t*>(D.descent));
            auto load_rows = [&](auto lava_mode) {
#pragma unroll
                for (int k = 0; k < H::KR; ++k) {
                    fl[k] = 0;
                    tp[k] = 0.0f;
                    ds[k] = -1;
                    if (4 * k + 3 < H::R || w + 4 * k < H::R) {
                        const uint64_t m = static_cast<uint64_t>(row_of(lava_lo, k)) |
                                           (static_cast<uint64_t>(row_of(lava_hi, k)) << 32);
                        if (lane_bit(m)) {
                            const uint32_t g = row_of(rows.base, w + 4 * k) + gxc;
                            fl[k] = flow_g[g];
                            if (lava_mode) {
                                tp[k] = temp_g[g];
                                ds[k] = desc_g[g];
                            }
                        }
                    }
                }
            };
            if (P.classic)
                load_rows(std::false_type{});
            else
                load_rows(std::true_type{});
            int32_t efl = 0;
            float etp = 0.0f;
            int32_t eds = -1;
            const int er = (t >> 1) & 63;
            const uint32_t ebase = static_cast<uint32_t>(__builtin_amdgcn_ds_bpermute(er << 2, static_cast<int>(rows.base)));
            const int ex = halo_x(P, (t & 1) ? o.x0 + kTileX : o.x0 - 1);
            const bool elava = t < 2 * H::R && ((rows.ok >> er) & 1u) && ex >= 0 && T.S.edge[t];
            if (elava) {
                const uint32_t g = ebase + static_cast<uint32_t>(ex);
                efl = D.flow[g];
                if (!P.classic) {
                    etp = tcur[g];
                    eds = D.descent[g];
                }
            }
            {
                // candidate cells (while the lava loads above are in flight): LAVA, or AIR next to a LAVA position
                // the rules read -- the 3 x 3 halo rows around the cell's row dilated by one in x in classic mode;
                // in lava mode the rows of the cell's layer and the one above (dz >= 0) plus, with ponding, the cell
                // directly below.  Every other cell keeps its state.
                // lane j < 16 folds the halo rows around the tile's row j; each wave then reads its four rows' masks
                // from those lanes
                uint64_t near_j = 0;
                if (lx < 16) {
                    const int rc = (lx % TY + 1) + LY * (lx / TY + 1);
                    uint64_t u = 0;
                    uint32_t el = 0, er = 0;
#pragma unroll
                    for (int dz = -1; dz <= 1; ++dz)
#pragma unroll
                        for (int dy = -1; dy <= 1; ++dy) {
                            if (dz < 0 && !P.classic) continue;
                            const int rr = rc + dy + LY * dz;
                            u |= T.S.lava[rr];
                            el |= T.S.edge[2 * rr];
                            er |= T.S.edge[2 * rr + 1];
                        }
                    near_j = u | (u << 1) | (u >> 1) | static_cast<uint64_t>(el) | (static_cast<uint64_t>(er) << 63);
                    if (!P.classic && P.ponding) near_j |= T.S.lava[rc - LY];
                }
                // LAVA candidates fill the list from the front, AIR ones from the back: each wave of the
                // candidate loop then runs one rule, not both
                const int xm0 = min(kTileX - 1, P.nx - 1 - o.x0), ym0 = min(TY - 1, P.ny - 1 - o.y0),
                          zm0 = min(TZ - 1, P.nz - 1 - o.z0);
                uint32_t bnd0 = 0;  // directions in which the tile holds LAVA on its boundary at the step's start
                bool cl[4], ca[4];
                uint32_t sl[4], sa[4];
                // positions past the grid's far side hold wrapped halo cells of a periodic partial tile: bit k of cin
                // (from lane k) = the wave's own row k lies inside the grid in y and z
                const uint64_t inx = __ballot(o.x0 + lx < P.nx);
                const int crow = w + 4 * lx;
                const uint32_t cin =
                    static_cast<uint32_t>(__ballot(lx < 4 && o.y0 + crow % TY < P.ny && o.z0 + crow / TY < P.nz));
#pragma unroll
                for (int k = 0; k < 4; ++k) {
                    const int row = w + 4 * k;
                    const int rc = (row % TY + 1) + LY * (row / TY + 1);
                    const uint64_t near = readlane64(near_j, row);
                    const uint8_t ty = lane_bit(((cin >> k) & 1u) ? inx : 0ull) ? T.S.type[(lx + 1) + LX * rc] : kVoidLds;
                    cl[k] = ty == kLava;
                    ca[k] = ty == kAir && ((near >> lx) & 1u);
                    if (ty == kAir && !ca[k]) tile_air = 1;
                    if (cl[k]) bnd0 |= boundary_dirs(lx, row % TY, row / TY, xm0, ym0, zm0);
                }
                bnd0 = wave_or(bnd0);
                if (lx == 0) s_bnd0[w] = bnd0;
                wave_append_k(cl, &s_nlist, sl);
                wave_append_k(ca, &s_nair, sa);
#pragma unroll
                for (int k = 0; k < 4; ++k) {
                    const uint16_t cell = static_cast<uint16_t>(lx + 64 * (w + 4 * k));
                    if (cl[k]) T.list[sl[k]] = cell;
                    if (ca[k]) T.list[kTileCells - 1 - sa[k]] = cell;
                }
            }
            // every position of the rows: the rules read these arrays at LAVA positions only
#pragma unroll
            for (int k = 0; k < H::KR; ++k) {
                if (4 * k + 3 < H::R || w + 4 * k < H::R) {
                    const int q = lx + 1 + H::LX * (w + 4 * k);
                    T.flow[q] = static_cast<F>(fl[k]);
                    T.temp[q] = tp[k];
                    T.desc[q] = static_cast<int8_t>(ds[k]);
                }
            }
            if (elava) {
                const int q = ((t & 1) ? H::LX - 1 : 0) + H::LX * (t >> 1);
                T.flow[q] = static_cast<F>(efl);
                T.temp[q] = etp;
                T.desc[q] = static_cast<int8_t>(eds);
            }
            if (t < H::R) {
                const uint64_t m = T.S.lava[t];
                const uint64_t lo = (m << 1) | static_cast<uint64_t>(T.S.edge[2 * t]);
                const uint32_t hi = static_cast<uint32_t>(m >> 63) | (static_cast<uint32_t>(T.S.edge[2 * t + 1]) << 1);
                T.row66[t][0] = static_cast<uint32_t>(lo);
                T.row66[t][1] = static_cast<uint32_t>(lo >> 32);
                T.row66[t][2] = hi;
                T.row66[t][3] = 0;
            }
        }
        __syncthreads();  // list and lava data of the tile are in LDS
        // candidate slots: LAVA cells at [0, nlava), AIR cells from the next wave boundary on (list order reversed)
        const uint32_t nlava = s_nlist, abase = (nlava + 63u) & ~63u, nslots = abase + s_nair;
        // Lava front: the next step's tiles next to the lava the tile holds at the step's start (and the tile itself)
        // are listed now, by the last wave (the one with the fewest candidates), so that the list atomics overlap the
        // candidate loop; the end of the tile lists only what the step's new lava adds (a tile listed but left with
        // no lava in its halo settles in the next step).
        const uint32_t dirs0 = (s_bnd0[0] | s_bnd0[1] | s_bnd0[2] | s_bnd0[3]) | (nlava ? (1u << 13) : 0u);
        if (sim.push && dirs0 && (tid >> 6) == kBlock / 64 - 1) {
            const int tx = o.x0 / kTileX, tyy = o.y0 / TY, tz = o.z0 / TZ;
            push_next(P, D, B, tx, tyy, tz, static_cast<int>(tid & 63u), dirs0, e >> kTileBits, static_cast<uint32_t>(t) + 1u,
                      par ^ 1, shape);
        }
        // directions in which the tile holds lava on its boundary after the step (mark_halo): cells that stay or
        // become LAVA
        const int xm = min(kTileX - 1, P.nx - 1 - o.x0), ym = min(TY - 1, P.ny - 1 - o.y0),
                  zm = min(TZ - 1, P.nz - 1 - o.z0);
        uint32_t bnd = 0;
        for (uint32_t j0 = 0; j0 < nslots; j0 += kBlock) {
            const uint32_t j = j0 + tid;
            const bool live = j < nlava || (j >= abase && j < nslots);
            const int e = live ? T.list[j < nlava ? j : kTileCells - 1 - (j - abase)] : 0;
            const int lx = e & 63, row = e >> 6;
            const int ly = row % TY, lz = row / TY;
            const int x = o.x0 + lx, y = o.y0 + ly, z = o.z0 + lz;
            const int c = (lx + 1) + LX * ((ly + 1) + LY * (lz + 1));
            bool want = false, stay = false, newborn = false, solid = false, cooled = false, alive = false;
            uint32_t i = 0;
            if (live) {
                const uint32_t col = static_cast<uint32_t>(x) + static_cast<uint32_t>(P.nx) * static_cast<uint32_t>(y);
                i = col + plane * static_cast<uint32_t>(z);
                const uint8_t ty = T.S.type[c];
                if (ty == kLava) {
                    if (P.classic) {
                        tile_lava = 1;
                        bnd |= boundary_dirs(lx, ly, lz, xm, ym, zm);
                    } else {
                        // rule 2 (LavaCA.hpp:1129-1147, 973-994)
                        const int32_t f = T.flow[c];
                        float tc = T.temp[c];
                        const float before = tc;
                        if (vent_tile && f > 0 && f < P.nvent_slots && D.vent_cell[f] == i)
                            tc = P.erupt_temp;
                        else
                            tc = tc - P.cool_amb;
                        // the next step's temperature, written whether or not it changed (the other copy holds the
                        // value of two steps ago)
                        tnext[i] = tc;
                        if (tc < P.solidus) {
                            want = true;
                            solid = true;
                            D.solid[i] = stamp;  // no rule reads it
                        } else {
                            stay = true;
                            tile_lava = 1;
                            cooled = tc != before;
                        }
                    }
                } else if (ty == kAir) {
                    // LAVA among the 26 neighbours, three x positions at a time from the row masks
                    uint32_t lava_bits = 0;
                    const int rc = (ly + 1) + LY * (lz + 1);
                    // halo positions lx .. lx + 2 of each 66-position row word: one funnel shift of two of its
                    // three 32-bit parts
                    const uint32_t* wbase = &T.row66[0][0] + (lx >> 5);
                    const uint32_t sh = static_cast<uint32_t>(lx) & 31u;
#pragma unroll
                    for (int dz = -1; dz <= 1; ++dz)
#pragma unroll
                        for (int dy = -1; dy <= 1; ++dy) {
                            const int rr = rc + dy + LY * dz;
                            const uint32_t b3 = __builtin_amdgcn_alignbit(wbase[4 * rr + 1], wbase[4 * rr], sh) & 7u;
                            lava_bits |= b3 << (3 * (dy + 1) + 9 * (dz + 1));
                        }
                    lava_bits &= rules_mask & ~(1u << 13);
                    int32_t wf = 0;
                    float wt = 0.0f;
                    bool has = false;
                    const int spill_c = has_spill ? T.spill[lx + kTileX * ly] : 0;
                    if (lava_bits) {
                        // rule 1 (LavaCA.hpp:1115-1128) / classic (LavaCA.hpp:1175-1190)
                        const uint64_t h_index = splitmix64(prefix ^ i);
                        uint32_t mask = P.nb.mask[0];
                        if (P.nb.variants > 1)
                            mask = nb_mask(P.nb, mod_small_inv(splitmix64(h_index ^ kSaltNeighborhood),
                                                               static_cast<uint32_t>(P.nb.variants), P.nb.var_r32,
                                                               P.nb.var_inv));
                        const uint32_t bits = lava_bits & mask;
                        if (bits) {
                            VoteFirst v;
                            gather_lds(P, tyc, T, c, lz, z, col, plane, spill_c, has_spill, bits, v);
                            has = v.n > 0;
                            if (v.overflow) {
                                // a second flowId (flows meet): the gather again with two slots; a third one: the
                                // full vote over the accepted positions (rare: where three flows meet)
                                VoteTwo v2;
                                gather_lds(P, tyc, T, c, lz, z, col, plane, spill_c, has_spill, bits, v2);
                                if (!v2.overflow) {
                                    v2.winner(h_index, wf, wt);
                                } else {
                                    VoteMask vm;
                                    gather_lds(P, tyc, T, c, lz, z, col, plane, spill_c, has_spill, bits, vm);
                                    vote_positions(P, T, c, vm.acc, h_index, wf, wt);
                                }
                            } else if (has) {
                                v.winner(wf, wt);
                            }
                        }
                    }
                    if (has && (P.classic || wt >= P.solidus)) {
                        want = true;
                        // the cell is AIR at the step's start, so no rule of this step reads its flowId
                        D.flow[i] = wf;
                        D.arrival[i] = stamp;
                        if (!P.classic) tnext[i] = wt;  // read by the next step as the new lava cell's temperature
                        newborn = true;
                        tile_lava = 1;
                    } else {
                        tile_air = 1;
                        // could a later step bear it while no type changes (CaCtrl::qdead)?  Any source the rules
                        // accept under any variant with a candidate temperature at the solidus or above; unchecked
                        // (after a step with births) counts as yes
                        if (lava_bits && !P.classic) {
                            if (born_last) {
                                alive = true;
                            } else {
                                VoteAlive va{P.solidus};
                                gather_lds(P, tyc, T, c, lz, z, col, plane, spill_c, has_spill, lava_bits, va);
                                alive = va.alive;
                            }
                        }
                    }
                }
            }
            w_lava += static_cast<uint32_t>(__popcll(__ballot(stay || newborn)));
            w_born += static_cast<uint32_t>(__popcll(__ballot(newborn)));
            w_temp += static_cast<uint32_t>(__popcll(__ballot(cooled)));
            w_alive += static_cast<uint32_t>(__popcll(__ballot(alive)));
            w_chg += static_cast<uint32_t>(__popcll(__ballot(want)));
            w_tchg = w_tchg || __ballot(newborn || solid) != 0;
            if (stay || newborn) bnd |= boundary_dirs(lx, ly, lz, xm, ym, zm);
            // the new type goes into the next type copy at the tile's end (a born cell was not a change of the
            // previous step; a cell that solidifies may have been born then: the new type wins)
            if (want) next_type(T.next, e, newborn);
        }
        // tile flags, record counts and the tile's counters (row tile % kPartRows of its simulation); every
        // record is one changed cell
        {
            // bit 2: a cell of the tile changed type (born or solidified) -- tile_tchg, for the next step's D(n)
            const uint32_t f = (__ballot(tile_lava) ? static_cast<uint32_t>(kFlagLava) : 0u) |
                               (__ballot(tile_air) ? static_cast<uint32_t>(kFlagAir) : 0u) | (w_tchg ? 4u : 0u);
            bnd = wave_or(bnd);
            if ((tid & 63u) == 0) {
                s_red[0][tid >> 6] = w_lava;
                s_red[1][tid >> 6] = w_born;
                s_red[2][tid >> 6] = bnd;
                s_red[3][tid >> 6] = w_temp;
                s_red[4][tid >> 6] = w_chg;
                s_red[5][tid >> 6] = w_alive;
                reinterpret_cast<uint8_t*>(&s_wflags[1])[tid >> 6] = static_cast<uint8_t>(f);
            }
        }
        __syncthreads();
        // a tile that changed a cell's type lists the tiles whose D(n) reads it for the next step's descent
        if (B.desc_list && !P.classic && (s_wflags[1] & 0x04040404u))
            push_descent<TZ>(P, D, B, o.x0 / kTileX, o.y0 / TY, o.z0 / TZ, static_cast<int>(tid), e >> kTileBits,
                             static_cast<uint32_t>(t) + 1u, par ^ 1, shape);
        // a tile holding lava after the step lists the next step's tiles (lava front: itself and the neighbours it
        // holds lava next to) or marks its neighbours for the next step's scan (any lava: visit; lava next to each
        // neighbour: stage); a tile that changed a cell's type is listed too (it brings the other copy up to date)
        if ((s_wflags[1] & 0x01010101u) != 0) {
            const int tx = o.x0 / kTileX, tyy = o.y0 / TY, tz = o.z0 / TZ;
            const uint32_t bnd_all = s_red[2][0] | s_red[2][1] | s_red[2][2] | s_red[2][3];
            if (sim.push) {
                const uint32_t late = (bnd_all | (1u << 13)) & ~dirs0;
                if (late && tid < 64)
                    push_next(P, D, B, tx, tyy, tz, static_cast<int>(tid), late, e >> kTileBits,
                              static_cast<uint32_t>(t) + 1u, par ^ 1, shape);
            } else {
                mark_near(P, D, tx, tyy, tz, static_cast<int>(tid));
                mark_halo(P, D, tx, tyy, tz, static_cast<int>(tid), bnd_all);
            }
        }
        if (tid == 0) {
            const uint32_t wf = s_wflags[1];
            const uint32_t fl = (wf | wf >> 8 | wf >> 16 | wf >> 24) & 0xFFu;
            D.tile_flags[tile] = static_cast<uint8_t>(fl & (kFlagLava | kFlagAir));
            if (fl & 4u) D.tile_tchg[tile] = static_cast<uint32_t>(t) + 1u;
            uint32_t r0 = 0, r1 = 0, r3 = 0, nchg = 0, nalive = 0;
            for (int w = 0; w < kBlock / 64; ++w) {
                r0 += s_red[0][w];
                r1 += s_red[1][w];
                r3 += s_red[3][w];
                nchg += s_red[4][w];
                nalive += s_red[5][w];
            }
            CaPart* row = &D.part[tile % kPartRows];
            if (s_nair) atomicAdd(&row->air_cand, static_cast<unsigned long long>(s_nair));
            if (nalive) atomicAdd(&row->air_alive, static_cast<unsigned long long>(nalive));
            if (nchg) atomicAdd(&row->tchanged, static_cast<unsigned long long>(nchg));
            if (r0) atomicAdd(&row->lava_after, static_cast<unsigned long long>(r0));
            if (r1) atomicAdd(&row->born, static_cast<unsigned long long>(r1));
            // changed cells: every type change (born, solidified) and every lava cell whose temperature changed
            if (nchg + r3) atomicAdd(&row->changed, static_cast<unsigned long long>(nchg + r3));
            atomicAdd(&row->staged, 1ULL);
        }
        // last: a wait for the list atomics' returns above would also wait for these stores
        if (changed || (s_red[4][0] | s_red[4][1] | s_red[4][2] | s_red[4][3]))
            write_types<TY, TZ>(P, o, T.next, tyn, tid);
    }
}

// finishStep of every slot of the batch, one block per slot (after a step whose finish the next step's first kernel
// does not do: every step of a batch that needs ca_front, the last step of a run of steps otherwise).
__global__ void __launch_bounds__(kBlock) ca_finish_kernel(BatchDev B) { finish_slot(B, blockIdx.x); }

// Initial list of a lava-front simulation that joins a batch: every tile that holds lava or has a neighbour tile with
// lava (a superset of the tiles whose 1-cell halo holds lava: the extra tiles settle in the rules kernel), and every
// tile that changed a cell's type in the last step, for the simulation's current step, into list copy par.  One lane
// per tile, one list atomic per sub-list a block appends to.
__global__ void __launch_bounds__(kBlock) ca_seed_kernel(BatchDev B, uint32_t slot, int par) {
    const CaSim& sim = g_sims[B.sim_base + slot];
    if (!sim.running || !sim.push) return;
    const CaParams& P = sim.P;
    const CaDev& D = sim.D;
    const uint32_t t = blockIdx.x * kBlock + threadIdx.x;
    bool want = false;
    if (t < P.ntiles && !D.ctrl->stop) {
        // a tile that changed a cell's type in the last step brings the other type copy up to date in the next
        const int32_t now = D.ctrl->steps;
        want = now > 0 && D.tile_tchg[t] == static_cast<uint32_t>(now);
        const uint32_t r = div_magic(t, P.div_ntx_m, P.div_ntx_p);
        const int tx = static_cast<int>(t - r * static_cast<uint32_t>(P.ntx));
        const uint32_t tz_ = div_magic(r, P.div_nty_m, P.div_nty_p);
        const int ty = static_cast<int>(r - tz_ * static_cast<uint32_t>(P.nty)), tz = static_cast<int>(tz_);
        for (int q = 0; q < 27 && !want; ++q) {
            int qx = tx + q % 3 - 1, qy = ty + (q / 3) % 3 - 1, qz = tz + q / 9 - 1;
            if (!wrap_axis(qx, P.ntx, P.per[0]) || !wrap_axis(qy, P.nty, P.per[1]) || !wrap_axis(qz, P.ntz, P.per[2]))
                continue;
            want = (D.tile_flags[static_cast<uint32_t>(qx) + static_cast<uint32_t>(P.ntx) *
                                                             (static_cast<uint32_t>(qy) + static_cast<uint32_t>(P.nty) *
                                                                                              static_cast<uint32_t>(qz))] &
                    kFlagLava) != 0;
        }
    }
    block_list_append(B, par, want ? sim.shape : -1, t,
                      (static_cast<uint64_t>(static_cast<uint32_t>(D.ctrl->steps)) << 32) | (slot << kTileBits) | t);
}

// Removes the entries of the slots in `slots` from list copy par, every sub-list (one block; in place, order kept
// otherwise).
__global__ void __launch_bounds__(1024) ca_purge_kernel(BatchDev B, SlotMask slots, int par) {
    __shared__ uint32_t s_cnt[1024 / 64];
    for (int k = 0; k < kShapes * kSub; ++k) {
        const uint32_t n = *list_count(B, par, k / kSub, k % kSub);
        uint64_t* L = sub_list(B, par, k / kSub, k % kSub);
        uint32_t out = 0;
        for (uint32_t base = 0; base < n; base += 1024) {
            const uint32_t i = base + threadIdx.x;
            uint64_t v = 0;
            bool keep = false;
            if (i < n) {
                v = L[i];
                const uint32_t sl = static_cast<uint32_t>(v) >> kTileBits;  // < kMaxSlots
                keep = !((slots.bits[sl >> 5] >> (sl & 31u)) & 1u);
            }
            const unsigned long long m = __ballot(keep);
            const int w = static_cast<int>(threadIdx.x >> 6);
            if ((threadIdx.x & 63u) == 0) s_cnt[w] = static_cast<uint32_t>(__popcll(m));
            __syncthreads();  // also: the chunk's entries are read before any is written (positions < base + 1024)
            uint32_t pos = out, tot = 0;
            for (int q = 0; q < 1024 / 64; ++q) {
                pos += q < w ? s_cnt[q] : 0u;
                tot += s_cnt[q];
            }
            if (keep) L[pos + mbcnt64(m)] = v;
            out += tot;
            __syncthreads();  // s_cnt is rewritten by the next chunk
        }
        if (threadIdx.x == 0) *list_count(B, par, k / kSub, k % kSub) = out;
        __syncthreads();  // the count is read at the next list's start
    }
}

__global__ void __launch_bounds__(kBlock) ca_status_kernel(BatchDev B, int32_t* out) {
    for (uint32_t k = blockIdx.x * kBlock + threadIdx.x; k < B.nslots; k += gridDim.x * kBlock) {
        const CaSim& sim = g_sims[B.sim_base + k];
        out[4 * k] = sim.running ? sim.D.ctrl->steps : 0;
        const int32_t stop = sim.running ? sim.D.ctrl->stop : -1;
        out[4 * k + 1] = stop == kStopHeld ? 0 : stop;
        out[4 * k + 2] = sim.running ? sim.D.ctrl->quiet : 0;
        out[4 * k + 3] = sim.running ? sim.D.ctrl->qdead : 0;
    }
}

__global__ void __launch_bounds__(kBlock) ca_owner_kernel(uint32_t* owner, uint32_t base, uint32_t n, uint32_t slot) {
    for (uint32_t j = blockIdx.x * kBlock + threadIdx.x; j < n; j += gridDim.x * kBlock) owner[base + j] = slot;
}

// ------------------------------------------------------------------------------------------------
// Fast-forward over steps in which no cell can be born
// ------------------------------------------------------------------------------------------------
// While no cell can be born, a lava-mode step changes nothing but the lava cells' temperatures and types (rule 2,
// LavaCA.hpp:1129-1147): a vent cell of an erupting step is reset to eruptTemp, every other lava cell loses
// coolingAmbient, and a cell whose new temperature is below the solidus becomes BASALT in that step.  Cells do not
// interact, so those steps are each cell's own sequence of the same float operations; the histories follow from the
// steps in which each cell solidifies or starts / stops changing.

// Replays one lava cell (temperature T, vent: it is its flow's vent cell) from step t0 over at most h steps.  Returns
// the step (from 0) in which it solidifies, h when it does not; T = its temperature after the steps replayed (its
// last one when it solidifies).  flip(k, +1 / -1): from step t0 + k on the cell starts / stops counting as changed
// (its temperature changed, or it solidified) -- called only for k < h.
template <class F>
__device__ inline int ff_replay(float& T, bool vent, int t0, int h, const CaParams& P, F&& flip) {
    bool prev = false;
    #pragma clang loop unroll(disable)
    for (int k = 0; k < h; ++k) {
        const bool erupt = vent && t0 + k < P.eruption_steps;
        const float n = erupt ? P.erupt_temp : T - P.cool_amb;
        const bool sol = n < P.solidus;
        const bool ch = sol || n != T;
        if (ch != prev) {
            flip(k, ch ? 1 : -1);
            prev = ch;
        }
        T = n;
        if (sol) {
            if (k + 1 < h) flip(k + 1, -1);
            return k;
        }
        if (!ch) {
            // unchanged: it stays so until the eruption ends (a vent) or for good
            if (!erupt) return h;
            k = P.eruption_steps - t0 - 1;  // resumes at the first step after the eruption
        }
    }
    return h;
}

// hist[key] += val for every lane with has (all 64 lanes call), one device atomic per distinct key of the wave: many
// cells share a step (every lava cell starts changing in the first one), and atomics on one 64-byte line serialise.
__device__ inline void wave_hist_add(unsigned long long* hist, bool has, int key, long long val) {
    for (unsigned long long m = __ballot(has); m;) {
        const int l0 = static_cast<int>(__builtin_ctzll(m));
        const int k0 = __builtin_amdgcn_readlane(key, l0);
        const bool in = has && key == k0;
        m &= ~__ballot(in);
        const unsigned long long sum = wave_sum(in ? static_cast<unsigned long long>(val) : 0ull);
        if (static_cast<int>(opaque_lane()) == l0 && sum != 0) atomicAdd(&hist[k0], sum);
    }
}

// Sets up a fast-forward: the first step with a solidification is not known yet.
__global__ void ca_ff_init_kernel(CaDev D) { D.ctrl->ff_sol = INT_MAX; }

// The lava cells at the start of step t0, in the tiles of the simulation's entries of list copy par (every tile that
// holds lava is listed, and so is every tile that changed a type in step t0 - 1), one block per tile and a lane per
// cell, blocks walking the sub-lists as the step kernels do.  FINAL = false: difference counts over the steps t0 ..
// t0 + h - 1 into the history arrays (zeroed by the caller) -- hist_active[t0 + s] += 1 for a cell that solidifies in
// step t0 + s, hist_changed[t0 + k] += +1 / -1 where a cell starts / stops counting as changed -- and the first step
// with a solidification (CaCtrl::ff_sol).  FINAL = true: the state after the steps the scan kept (h = CaCtrl::steps -
// t0) into both type and both temperature copies (the other type copy is behind in the tiles that changed a type in
// step t0 - 1), the tiles' lava flags, and the step of a tile's last type change (tile_tchg) where a cell solidified in
// the last of those steps.
template <bool FINAL>
__global__ void __launch_bounds__(kBlock) ca_ff_tiles_kernel(BatchDev B, uint32_t slot, int shape, int par, CaParams P,
                                                             CaDev D, int32_t t0, int32_t h) {
    __shared__ int s_flags;
    const uint32_t lane = threadIdx.x & 63u;
    const int w = static_cast<int>(threadIdx.x >> 6);
    const uint8_t* cur = D.type[t0 & 1];
    const float* tcur = D.temp[t0 & 1];
    if (FINAL) h = __builtin_amdgcn_readfirstlane(D.ctrl->steps) - t0;
    const uint32_t plane = static_cast<uint32_t>(P.nx) * static_cast<uint32_t>(P.ny);
    const uint32_t stride = gridDim.x / kSub;
    const uint32_t n = *list_count(B, par, shape, static_cast<int>(blockIdx.x % kSub));
    const uint64_t* list = sub_list(B, par, shape, static_cast<int>(blockIdx.x % kSub));
    int sol_min = INT_MAX;
    for (uint32_t e = blockIdx.x / kSub; e < n; e += stride) {
        const uint32_t ent = static_cast<uint32_t>(list[e]);
        if ((ent >> kTileBits) != slot) continue;  // block-uniform
        const uint32_t tile = ent & kTileMask;
        const TileCoord o = tile_origin(P, tile);
        const int x = o.x0 + static_cast<int>(lane);
        bool left = false, last = false;
#pragma unroll
        for (int k = 0; k < kTileCells / kBlock; ++k) {  // wave-uniform: row w + 4 k
            const int row = w + 4 * k;
            const int y = o.y0 + row % P.ty, z = o.z0 + row / P.ty;
            const bool in = x < P.nx && y < P.ny && z < P.nz;
            const uint32_t i = static_cast<uint32_t>(x) + static_cast<uint32_t>(P.nx) * static_cast<uint32_t>(y) +
                               plane * static_cast<uint32_t>(z);
            const uint8_t ty = in ? cur[i] : kVoidLds;
            const bool lava = ty == kLava;
            if (FINAL && in && !lava) D.type[(t0 + 1) & 1][i] = ty;
            int s = h;
            float T = 0.0f;
            int fk[4] = {0, 0, 0, 0}, fd[4] = {0, 0, 0, 0}, nf = 0;
            if (lava) {
                const int32_t f = D.flow[i];
                const bool vent = f > 0 && f < P.nvent_slots && D.vent_cell[f] == i;
                T = tcur[i];
                // a cell starts / stops counting as changed at most four times (a vent: its first reset, the
                // eruption's constant temperature, the cooling after it, its end)
                s = ff_replay(T, vent, t0, h, P, [&](int kk, int d) {
                    if (!FINAL && nf < 4) {
                        fk[nf] = kk;
                        fd[nf] = d;
                    }
                    ++nf;
                });
            }
            if (!FINAL) {
                // added wave by wave: many cells share a step
#pragma unroll
                for (int q = 0; q < 4; ++q) wave_hist_add(D.hist_changed + t0, lava && q < nf, fk[q], fd[q]);
                wave_hist_add(D.hist_active + t0, lava && s < h, s, 1);
                if (lava && s < h) sol_min = min(sol_min, s);
            } else if (lava) {
                const bool solid = s < h;
                const uint8_t nt = solid ? kBasalt : kLava;
                D.type[0][i] = nt;
                D.type[1][i] = nt;
                D.temp[0][i] = T;
                D.temp[1][i] = T;
                if (solid) D.solid[i] = static_cast<uint16_t>(t0 + s);  // < maxSteps <= 65535
                left = left || !solid;
                last = last || (solid && s == h - 1);
            }
        }
        if (FINAL) {
            if (threadIdx.x == 0) s_flags = 0;
            __syncthreads();
            const uint32_t f = (__ballot(left) ? 1u : 0u) | (__ballot(last) ? 2u : 0u);
            if (lane == 0 && f) atomicOr(&s_flags, static_cast<int>(f));
            __syncthreads();
            if (threadIdx.x == 0) {
                const int fl = s_flags;
                D.tile_flags[tile] =
                    static_cast<uint8_t>((D.tile_flags[tile] & ~kFlagLava) | ((fl & 1) ? kFlagLava : 0));
                if (fl & 2) D.tile_tchg[tile] = static_cast<uint32_t>(t0 + h);
            }
            __syncthreads();  // s_flags is reset for the next tile
        }
    }
    if (!FINAL) {
        // the wave's smallest, one atomic per wave that saw a solidification
        for (int d = 32; d >= 1; d >>= 1) sol_min = min(sol_min, __shfl_xor(sol_min, d, 64));
        if (lane == 0 && sol_min != INT_MAX) atomicMin(&D.ctrl->ff_sol, sol_min);
    }
}

// Estimate of the steps from t0 to the first solidification (CaCtrl::ff_sol, set to INT_MAX by ca_ff_init first): per
// lava cell (temperature - solidus) / coolingAmbient, a vent while it erupts from the eruption's end on -- no replay,
// so that a fast-forward that would end within a few steps can be skipped cheaply.
__global__ void __launch_bounds__(kBlock) ca_ff_estimate_kernel(BatchDev B, uint32_t slot, int shape, int par, CaParams P,
                                                                CaDev D, int32_t t0) {
    const uint32_t lane = threadIdx.x & 63u;
    const int w = static_cast<int>(threadIdx.x >> 6);
    const uint8_t* cur = D.type[t0 & 1];
    const float* tcur = D.temp[t0 & 1];
    const uint32_t plane = static_cast<uint32_t>(P.nx) * static_cast<uint32_t>(P.ny);
    const uint32_t stride = gridDim.x / kSub;
    const uint32_t n = *list_count(B, par, shape, static_cast<int>(blockIdx.x % kSub));
    const uint64_t* list = sub_list(B, par, shape, static_cast<int>(blockIdx.x % kSub));
    const float rc = P.cool_amb > 0.0f ? 1.0f / P.cool_amb : 0.0f;
    float best = 1e9f;
    for (uint32_t e = blockIdx.x / kSub; e < n; e += stride) {
        const uint32_t ent = static_cast<uint32_t>(list[e]);
        if ((ent >> kTileBits) != slot) continue;
        const TileCoord o = tile_origin(P, ent & kTileMask);
        const int x = o.x0 + static_cast<int>(lane);
        // the four cells' type, flowId and temperature read together, then the vent test
        constexpr int K = kTileCells / kBlock;
        uint32_t ci[K];
        uint32_t ty[K];
        int32_t fl[K];
        float tp[K];
#pragma unroll
        for (int k = 0; k < K; ++k) {
            const int row = w + 4 * k;
            const int y = o.y0 + row % P.ty, z = o.z0 + row / P.ty;
            const bool in = x < P.nx && y < P.ny && z < P.nz;
            ci[k] = static_cast<uint32_t>(x) + static_cast<uint32_t>(P.nx) * static_cast<uint32_t>(y) +
                    plane * static_cast<uint32_t>(z);
            ty[k] = in ? cur[ci[k]] : kVoidLds;
            fl[k] = in ? D.flow[ci[k]] : 0;
            tp[k] = in ? tcur[ci[k]] : 0.0f;
        }
#pragma unroll
        for (int k = 0; k < K; ++k) {
            if (ty[k] != kLava) continue;
            const int32_t f = fl[k];
            const bool vent = t0 < P.eruption_steps && f > 0 && f < P.nvent_slots && D.vent_cell[f] == ci[k];
            const float est = vent ? static_cast<float>(P.eruption_steps - t0) + (P.erupt_temp - P.solidus) * rc
                                   : (tp[k] - P.solidus) * rc;
            best = rc > 0.0f ? fminf(best, est) : best;
        }
    }
    for (int d = 32; d >= 1; d >>= 1) best = fminf(best, __shfl_xor(best, d, 64));
    if (lane == 0 && best < 1e9f) atomicMin(&D.ctrl->ff_sol, static_cast<int>(fmaxf(0.0f, best)));
}

// Inclusive prefix sum over the 64 lanes of a wave.
__device__ inline unsigned long long wave_scan(unsigned long long v) {
    const int lane = static_cast<int>(threadIdx.x & 63u);
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const unsigned long long u = __shfl_up(v, static_cast<unsigned>(d), 64);
        if (lane >= d) v += u;
    }
    return v;
}

// One block: the histories of the steps t0 .. t0 + he - 1 from the difference counts (N = the lava cells at t0:
// active = N - cells solidified so far, changed = the running sum of the +1 / -1 flips), where he = h, or with
// until_solid the steps up to and including the first with a solidification; the first of those steps whose finish
// stops the run (finishStep's order, LavaCA.hpp:1206-1229: NoLava, NoChange for a deterministic neighbourhood,
// MaxSteps); and CaCtrl as the finish of the last step kept leaves it (when no stop rule fires, the simulation
// goes on from t0 + he with the next step counting as the first after a step without births).
__global__ void __launch_bounds__(kBlock) ca_ff_scan_kernel(CaParams P, CaDev D, int32_t t0, int32_t h,
                                                            int32_t until_solid) {
    __shared__ unsigned long long s_a[kBlock / 64], s_c[kBlock / 64], s_sum[kBlock / 64];
    __shared__ int s_first[kBlock / 64];
    __shared__ int s_stop_k, s_reason;
    __shared__ unsigned long long s_active;
    CaCtrl* c = D.ctrl;
    const unsigned long long N = c->lava_total;
    if (until_solid && c->ff_sol < h) h = c->ff_sol + 1;
    const int w = static_cast<int>(threadIdx.x >> 6), lane = static_cast<int>(threadIdx.x & 63u);
    unsigned long long carry_a = 0, carry_c = 0, changed_sum = 0;
    int done = 0;
    if (threadIdx.x == 0) {
        s_stop_k = h - 1;  // no stop rule fires: the last step kept
        s_reason = 0;
    }
    for (int base = 0; base < h && !done; base += kBlock) {
        const int k = base + static_cast<int>(threadIdx.x);
        const unsigned long long a = k < h ? D.hist_active[t0 + k] : 0ULL;
        const unsigned long long cc = k < h ? D.hist_changed[t0 + k] : 0ULL;
        const unsigned long long wa = wave_scan(a), wc = wave_scan(cc);
        if (lane == 63) {
            s_a[w] = wa;
            s_c[w] = wc;
        }
        __syncthreads();
        unsigned long long pa = carry_a + wa, pc = carry_c + wc;
        for (int q = 0; q < w; ++q) {
            pa += s_a[q];
            pc += s_c[q];
        }
        for (int q = 0; q < kBlock / 64; ++q) {
            carry_a += s_a[q];
            carry_c += s_c[q];
        }
        const unsigned long long active = N - pa, changed = pc;
        int reason = 0;
        if (k < h) {
            if (active == 0)
                reason = 1;  // NoLava
            else if (changed == 0 && P.nb.variants == 1)
                reason = 3;  // NoChange
            else if (t0 + k + 1 >= P.max_steps)
                reason = 4;  // MaxSteps
        }
        const uint64_t hit = __ballot(reason != 0);
        if (lane == 0) s_first[w] = hit ? base + 64 * w + static_cast<int>(__builtin_ctzll(hit)) : INT_MAX;
        __syncthreads();
        int kmin = INT_MAX;
        for (int q = 0; q < kBlock / 64; ++q) kmin = min(kmin, s_first[q]);
        if (k < h && k <= kmin) {
            D.hist_active[t0 + k] = active;
            D.hist_changed[t0 + k] = changed;
            changed_sum += changed;
        }
        if (k == kmin) {
            s_stop_k = k;
            s_reason = reason;
        }
        if (k == min(kmin, h - 1)) s_active = active;
        done = kmin != INT_MAX;
        __syncthreads();  // s_a, s_c, s_first are rewritten by the next chunk
    }
    changed_sum = wave_sum(changed_sum);
    if (lane == 0) s_sum[w] = changed_sum;
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long total = 0;
        for (int q = 0; q < kBlock / 64; ++q) total += s_sum[q];
        c->steps = t0 + s_stop_k + 1;
        c->prefix = hash_prefix(P.seed, static_cast<uint64_t>(t0 + s_stop_k + 1));
        c->stop = s_reason;
        c->lava_total = s_active;
        c->changes_total += total;
        if (c->ff_steps == 0) c->ff_from = t0;
        c->ff_steps += s_stop_k + 1;
        c->born_last = 0;
        c->qdead = 0;  // the types may have changed in the last step kept: the next finish decides again
    }
}

// A simulation that goes on after a fast-forward: its entries of list copy par (pushed for step t0) move to the
// step it now is at.
__global__ void __launch_bounds__(kBlock) ca_ff_relabel_kernel(BatchDev B, CaDev D, uint32_t slot, int shape, int par) {
    if (D.ctrl->stop) return;
    const uint32_t now = static_cast<uint32_t>(D.ctrl->steps);
    const uint32_t n = *list_count(B, par, shape, static_cast<int>(blockIdx.y));  // one sub-list per grid row
    uint64_t* list = sub_list(B, par, shape, static_cast<int>(blockIdx.y));
    for (uint32_t j = blockIdx.x * kBlock + threadIdx.x; j < n; j += gridDim.x * kBlock) {
        const uint64_t e = list[j];
        if (((static_cast<uint32_t>(e) >> kTileBits) & 0x3FFu) == slot)
            list[j] = (static_cast<uint64_t>(now) << 32) | (e & 0xFFFFFFFFull);
    }
}

// ------------------------------------------------------------------------------------------------
// MC kernels
// ------------------------------------------------------------------------------------------------

namespace {


__device__ inline void cell_xyz(const McParams& P, uint32_t g, int& x, int& y, int& z) {
    x = static_cast<int>(g % static_cast<uint32_t>(P.nx));
    const uint32_t r = g / static_cast<uint32_t>(P.nx);
    y = static_cast<int>(r % static_cast<uint32_t>(P.ny));
    z = static_cast<int>(r / static_cast<uint32_t>(P.ny));
}

// Spins of the 26 neighbours of cell g = (x, y, z) in one round trip: every read is issued before the first is
// used (no branch around a load: a position outside the mask or past an absorbing edge reads the cell itself
// and has use[k] false), each into a 32-bit register so that no result is packed -- and waited for -- early.
template <typename S>
__device__ inline void nb_spins(const McParams& P, const S* spin, uint32_t g, int x, int y, int z, uint32_t mask,
                                uint32_t (&sv)[27], bool (&use)[27]) {
    int wx[3], wy[3], wz[3];
    bool ox[3], oy[3], oz[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        wx[d] = x + d - 1;
        ox[d] = wrap_axis(wx[d], P.nx, P.per[0]);
        wy[d] = y + d - 1;
        oy[d] = wrap_axis(wy[d], P.ny, P.per[1]);
        wz[d] = z + d - 1;
        oz[d] = wrap_axis(wz[d], P.nz, P.per[2]);
    }
    // row and plane offsets of the three wrapped y and z values: a neighbour's index is then two adds
    uint32_t ry[3], pz[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        ry[d] = static_cast<uint32_t>(P.nx) * static_cast<uint32_t>(wy[d]);
        pz[d] = static_cast<uint32_t>(P.nx) * static_cast<uint32_t>(P.ny) * static_cast<uint32_t>(wz[d]);
    }
#pragma unroll
    for (int k = 0; k < 27; ++k) {
        if (k == 13) {
            use[k] = false;
            sv[k] = 0;
            continue;
        }
        const int dx = k % 3, dy = (k / 3) % 3, dz = k / 9;
        use[k] = ((mask >> k) & 1u) && ox[dx] && oy[dy] && oz[dz];
        const uint32_t q = static_cast<uint32_t>(wx[dx]) + ry[dy] + pz[dz];
        sv[k] = static_cast<uint32_t>(spin[use[k] ? q : g]);
    }
}

// The spins a Metropolis trial of cell g reads (its own and those of neighbourhood variant `var`), in one round trip.
template <typename S>
__device__ inline void mc_gather(const McParams& P, const S* spin, uint32_t g, int var, int32_t& cur,
                                 uint32_t (&sv)[27], bool (&use)[27]) {
    int x, y, z;
    cell_xyz(P, g, x, y, z);
    cur = static_cast<int32_t>(spin[g]);
    nb_spins(P, spin, g, x, y, z, P.nb.mask[var], sv, use);
}

// One Metropolis trial (PottsMC.cpp:403-443) on the gathered spins: h = hashIndex(prefix, g), accept is the table
// T[before][after] (kDownhill: dE <= 0, else accept iff (r >> 11) < T).
template <typename S>
__device__ inline void mc_decide(const McParams& P, S* spin, uint32_t g, uint64_t h, int32_t cur,
                                 const uint32_t (&sv)[27], const bool (&use)[27], const uint64_t* accept,
                                 unsigned long long& att, unsigned long long& acc, unsigned long long& up) {
    ++att;
    const uint64_t ho = splitmix64(h ^ kSaltMcOrient);
    const uint32_t qm = static_cast<uint32_t>(P.Q - 1);
    int32_t prop = 1 + static_cast<int32_t>(qm < 65536u ? mod_small(ho, qm) : ho % qm);
    if (prop >= cur) ++prop;
    int in_dom = 0, same_a = 0, same_b = 0;
#pragma unroll
    for (int k = 0; k < 27; ++k) {
        if (k == 13) continue;
        const int32_t sj = use[k] ? static_cast<int32_t>(static_cast<S>(sv[k])) : 0;
        in_dom += sj != 0;
        same_a += sj == cur;
        same_b += sj == prop;
    }
    const uint64_t thr = accept[(in_dom - same_a) * 27 + (in_dom - same_b)];
    bool ok;
    if (thr == kDownhill) {
        ok = true;
    } else {
        ok = thr != 0 && (splitmix64(h ^ kSaltMcAccept) >> 11) < thr;
        up += ok ? 1 : 0;
    }
    if (ok) {
        spin[g] = static_cast<S>(prop);
        ++acc;
    }
}

template <typename S>
__device__ inline void mc_trial(const McParams& P, S* spin, uint32_t g, uint64_t h, int var, const uint64_t* accept,
                                unsigned long long& att, unsigned long long& acc, unsigned long long& up) {
    int32_t cur;
    uint32_t sv[27];
    bool use[27];
    mc_gather(P, spin, g, var, cur, sv, use);
    mc_decide(P, spin, g, h, cur, sv, use, accept, att, acc, up);
}

}  // namespace

namespace {

// What a trial needs before it reads spins: the cell (list entry j), its hash, whether the mobility skip lets it
// try, and its neighbourhood variant.  Reads only the list and the thresholds, which no pass writes.
struct McPrep {
    uint32_t g;
    int32_t var;
    uint64_t h;
    bool go;
};

__device__ inline McPrep mc_prep(const McParams& P, const uint32_t* list, const uint64_t* mob_thr, uint32_t j,
                                 uint32_t hi, uint64_t prefix) {
    McPrep t{0u, 0, 0ull, false};
    if (j < hi) {
        t.g = list[j];
        t.h = splitmix64(prefix ^ t.g);
        t.go = mob_thr == nullptr || (splitmix64(t.h ^ kSaltMcMobility) >> 11) < mob_thr[j];
        t.var = P.nb.variants > 1 ? static_cast<int>(mod_small(splitmix64(t.h ^ kSaltNeighborhood),
                                                               static_cast<uint32_t>(P.nb.variants)))
                                  : 0;
    }
    return t;
}

}  // namespace

// One colour pass (PottsMC.cpp:450-514): one trial per list entry.
template <typename S>
__global__ void __launch_bounds__(kBlock) mc_color_kernel(McParams P, S* spin, const uint32_t* list,
                                                          const uint64_t* mob_thr, uint32_t count, uint64_t prefix,
                                                          const uint64_t* accept, unsigned long long* counters) {
    unsigned long long c[3] = {0, 0, 0};  // attempts, accepted, uphill
    for (uint32_t j = blockIdx.x * kBlock + threadIdx.x; j < count; j += gridDim.x * kBlock) {
        const McPrep t = mc_prep(P, list, mob_thr, j, count, prefix);
        if (t.go) mc_trial(P, spin, t.g, t.h, t.var, accept, c[0], c[1], c[2]);
    }
    block_fold(c, 0u, counters + (blockIdx.x % kMcRows) * kMcCounters);
}

namespace {

// Device-wide barrier of the persistent sweep kernel.  Blocks arrive in eight groups (blockIdx % 8, the XCD a block
// is dispatched to); the last arrival of a group arrives at the top counter, and the last top arrival publishes the
// generation `target` (barrier k of a launch: the generation read at the launch's start + k).  Every counter is reset
// by its last arrival before the generation moves on, so the state is ready for the next barrier and the next launch.
// Every write of the grid before the barrier is visible to every block after it: each arrival releases (one L2
// write-back), the chain of last arrivals acquires and releases, and a waiter polls the generation with relaxed
// loads (which read past the XCD's L2) and acquires once when it arrived -- an acquire per poll would invalidate the
// L2 under the blocks still working.  A wait ends after `limit` ticks of the constant-rate clock or when another
// block gave up: it then sets `abort` and returns false, and the caller stops -- so the kernel ends even if its
// blocks were not all resident (the host reports the flag as an error).
__device__ bool grid_barrier(McGridBar* bar, uint32_t target, uint64_t limit) {
    __shared__ int s_ok;
    __syncthreads();  // the block's writes are done before thread 0 releases them
    if (threadIdx.x == 0) {
        const uint32_t G = gridDim.x, grp = blockIdx.x & 7u, groups = G < 8u ? G : 8u;
        const uint32_t members = (G - grp + 7u) >> 3;
        if (__hip_atomic_fetch_add(&bar->group[grp][0], 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT) ==
            members - 1u) {
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
            __hip_atomic_store(&bar->group[grp][0], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            if (__hip_atomic_fetch_add(&bar->top[0], 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT) == groups - 1u) {
                __hip_atomic_store(&bar->top[0], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                __hip_atomic_store(&bar->gen[0], target, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
            }
        }
        int ok = 1;
        const uint64_t t0 = wall_clock64();
        while (__hip_atomic_load(&bar->gen[0], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != target) {
            if (__hip_atomic_load(&bar->abort[0], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0u ||
                wall_clock64() - t0 > limit) {
                __hip_atomic_store(&bar->abort[0], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                ok = 0;
                break;
            }
            __builtin_amdgcn_s_sleep(1);
        }
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
        s_ok = ok;
    }
    __syncthreads();
    return s_ok != 0;
}

// This block's share of colour pass `col`: list positions [lo, hi).
__device__ inline void mc_share(const McColours& C, int col, uint32_t chunk, uint32_t& lo, uint32_t& hi) {
    const uint32_t b = C.start[col], n = C.start[col + 1] - b;
    lo = b + static_cast<uint32_t>(static_cast<uint64_t>(n) * chunk / gridDim.x);
    hi = b + static_cast<uint32_t>(static_cast<uint64_t>(n) * (chunk + 1u) / gridDim.x);
}

}  // namespace

// nsweeps checkerboard sweeps in one launch (PottsMC.cpp:450-533): the colour passes of every sweep in order, a
// device-wide barrier between passes.  The grid is at most one block per CU (4 waves of fewer than 128 VGPRs, so four
// such blocks fit on a CU): every block is resident together, also beside the three other MC streams of the batched
// pipeline.  Block b takes
// the same share of every colour's list -- chunk order groups the blocks of one XCD on one stretch of the domain.
// The acceptance table sits in LDS.  The hashes of a trial depend only on (sweep, cell): the trials are those of
// one launch per colour.
template <typename S>
__global__ void __launch_bounds__(kBlock) mc_sweeps_kernel(McParams P, S* spin, const uint32_t* list,
                                                           const uint64_t* mob_thr, McColours C, uint64_t sweep0,
                                                           int32_t nsweeps, const uint64_t* accept,
                                                           unsigned long long* counters, McGridBar* bar, uint64_t limit) {
    __shared__ uint64_t s_acc[27 * 27];
    for (int k = threadIdx.x; k < 27 * 27; k += kBlock) s_acc[k] = accept[k];
    __syncthreads();
    const uint32_t G = gridDim.x, grp = blockIdx.x & 7u;
    uint32_t chunk = blockIdx.x >> 3;  // this block's share: groups in order, blocks of a group consecutive
    for (uint32_t q = 0; q < grp; ++q) chunk += (G - q + 7u) >> 3;
    // the generation before the first barrier (no barrier of this launch can complete before every block read it)
    uint32_t gen = __hip_atomic_load(&bar->gen[0], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    unsigned long long c[3] = {0, 0, 0};
    bool live = true;
    for (int32_t sw = 0; sw < nsweeps && live; ++sw) {
        const uint64_t prefix = hash_prefix(P.seed, sweep0 + static_cast<uint64_t>(sw));
        for (int32_t col = 0; col < C.count && live; ++col) {
            uint32_t lo, hi;
            mc_share(C, col, chunk, lo, hi);
            for (uint32_t j = lo + threadIdx.x; j < hi; j += kBlock) {
                const McPrep t = mc_prep(P, list, mob_thr, j, hi, prefix);
                if (t.go) mc_trial(P, spin, t.g, t.h, t.var, s_acc, c[0], c[1], c[2]);
            }
            if (sw + 1 < nsweeps || col + 1 < C.count) live = grid_barrier(bar, ++gen, limit);
        }
    }
    block_fold(c, 0u, counters + (blockIdx.x % kMcRows) * kMcCounters);
}

// mc.update = random (sweepTrials, PottsMC.cpp:312-391): inherently sequential, one lane.
template <typename S>
__global__ void mc_random_kernel(McParams P, S* spin, const uint32_t* list, const uint64_t* mob_thr, uint32_t count,
                                 uint64_t prefix, const uint64_t* accept, unsigned long long* counters) {
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    unsigned long long att = 0, acc = 0, up = 0;
    for (uint64_t a = 0; a < count; ++a) {
        const uint64_t h = splitmix64(prefix ^ a);
        const uint64_t kk = splitmix64(h ^ kSaltMcPick) % count;
        const uint32_t g = list[kk];
        if (mob_thr != nullptr && (splitmix64(h ^ kSaltMcMobility) >> 11) >= mob_thr[kk]) continue;
        const int var = P.nb.variants > 1
                            ? static_cast<int>(mod_small(splitmix64(splitmix64(prefix ^ g) ^ kSaltNeighborhood),
                                                         static_cast<uint32_t>(P.nb.variants)))
                            : 0;
        mc_trial(P, spin, g, h, var, accept, att, acc, up);
    }
    counters[0] += att;
    counters[1] += acc;
    counters[2] += up;
}

// initialize (PottsMC.cpp:104-143): spin = 1 + hashRand(seed, 0, g, kSaltMcInit) % Q in the domain.
template <typename S>
__global__ void __launch_bounds__(kBlock) mc_init_kernel(McParams P, S* spin, const uint32_t* list, uint32_t count) {
    const uint32_t j = blockIdx.x * kBlock + threadIdx.x;
    if (j >= count) return;
    const uint32_t g = list[j];
    const uint64_t p0 = hash_prefix(P.seed, 0);
    spin[g] = static_cast<S>(1 + splitmix64(splitmix64(p0 ^ g) ^ kSaltMcInit) % static_cast<uint64_t>(P.Q));
}

// totalEnergy (PottsMC.cpp:590-664): unlike pairs; forward half (k > 13) for deterministic neighbourhoods,
// full variant of step `prefix` for random ones.
template <typename S>
__global__ void __launch_bounds__(kBlock) mc_energy_kernel(McParams P, const S* spin, const uint32_t* list,
                                                           uint32_t count, uint64_t prefix, unsigned long long* out) {
    unsigned long long u = 0;  // unlike neighbour pairs counted from this cell
    const uint32_t j = blockIdx.x * kBlock + threadIdx.x;
    if (j < count) {
        const uint32_t g = list[j];
        int x, y, z;
        cell_xyz(P, g, x, y, z);
        const int32_t s = static_cast<int32_t>(spin[g]);
        uint32_t mask;
        if (P.nb.variants == 1) {
            mask = P.nb.mask[0] & ~((1u << 14) - 1u);
        } else {
            mask = nb_mask(P.nb, mod_small(splitmix64(splitmix64(prefix ^ g) ^ kSaltNeighborhood),
                                           static_cast<uint32_t>(P.nb.variants)));
        }
        uint32_t sv[27];
        bool use[27];
        nb_spins(P, spin, g, x, y, z, mask, sv, use);
        int in_dom = 0, same = 0;
#pragma unroll
        for (int k = 0; k < 27; ++k) {
            const int32_t sj = use[k] ? static_cast<int32_t>(static_cast<S>(sv[k])) : 0;
            in_dom += sj != 0;
            same += sj == s;
        }
        u = static_cast<unsigned long long>(in_dom - same);
    }
    unsigned long long v[1] = {u};
    block_fold(v, 0u, out + (blockIdx.x % kMcRows) * kMcSlots + 9);
}

// ---- grain labelling (Grains.cpp:33-150) by lock-free union-find: the larger root is always hooked under
// the smaller one, so every component's final root is its smallest global index, as in the reference.
namespace {
__device__ inline uint32_t uf_find(uint32_t* parent, uint32_t x) {
    volatile uint32_t* p = parent;
    uint32_t px = p[x];
    while (px != x) {
        const uint32_t gp = p[px];
        if (gp != px) p[x] = gp;  // path halving: gp is an ancestor of x in the same set
        x = px;
        px = gp;
    }
    return x;
}
__device__ inline void uf_unite(uint32_t* parent, uint32_t a, uint32_t b) {
    for (;;) {
        a = uf_find(parent, a);
        b = uf_find(parent, b);
        if (a == b) return;
        if (a > b) {
            const uint32_t t = a;
            a = b;
            b = t;
        }
        const uint32_t old = atomicCAS(&parent[b], b, a);
        if (old == b) return;
        b = old;
    }
}
}  // namespace

__global__ void __launch_bounds__(kBlock) mc_grain_init_kernel(const uint32_t* list, uint32_t count, uint32_t* parent,
                                                                uint32_t* size, uint32_t* hi) {
    const uint32_t j = blockIdx.x * kBlock + threadIdx.x;
    if (j >= count) return;
    const uint32_t g = list[j];
    parent[g] = g;
    size[g] = 0;
    if (hi) hi[g] = 0;
}

template <typename S>
__global__ void __launch_bounds__(kBlock) mc_grain_union_kernel(McParams P, const S* spin, const uint32_t* list,
                                                                 uint32_t count, uint32_t* parent) {
    const uint32_t j = blockIdx.x * kBlock + threadIdx.x;
    if (j >= count) return;
    const uint32_t g = list[j];
    int x, y, z;
    cell_xyz(P, g, x, y, z);
    const S s = spin[g];
    const uint32_t mask = P.nb.conn_mask & ~((1u << 14) - 1u);
    // the 13 forward neighbours' spins read together (no branch around a load: a position outside the mask or
    // past an absorbing edge reads the cell itself and is not used), then the unions
    uint32_t nq[13], sv[13];
    bool use[13];
#pragma unroll
    for (int j2 = 0; j2 < 13; ++j2) {
        const int k = 14 + j2;
        int qx = x + (k % 3 - 1), qy = y + ((k / 3) % 3 - 1), qz = z + (k / 9 - 1);
        use[j2] = ((mask >> k) & 1u) && wrap_axis(qx, P.nx, P.per[0]) && wrap_axis(qy, P.ny, P.per[1]) &&
                  wrap_axis(qz, P.nz, P.per[2]);
        nq[j2] = use[j2] ? static_cast<uint32_t>(qx) + static_cast<uint32_t>(P.nx) *
                                                         (static_cast<uint32_t>(qy) + static_cast<uint32_t>(P.ny) * qz)
                         : g;
        sv[j2] = static_cast<uint32_t>(spin[nq[j2]]);
    }
#pragma unroll
    for (int j2 = 0; j2 < 13; ++j2)
        if (use[j2] && static_cast<S>(sv[j2]) == s) uf_unite(parent, g, nq[j2]);
}

__global__ void __launch_bounds__(kBlock) mc_grain_flatten_kernel(const uint32_t* list, uint32_t count,
                                                                   uint32_t* parent, uint32_t* size, uint32_t* hi,
                                                                   const float* mobility, unsigned long long* out) {
    unsigned long long high = 0;
    const uint32_t j = blockIdx.x * kBlock + threadIdx.x;
    if (j < count) {
        const uint32_t g = list[j];
        const uint32_t r = uf_find(parent, g);
        atomicAdd(&size[r], 1u);
        if (hi != nullptr && mobility[g] >= 0.5f) {  // McStage.cpp:137 (float vs double 0.5: exact)
            atomicAdd(&hi[r], 1u);
            high = 1;
        }
    }
    unsigned long long v[1] = {high};
    block_fold(v, 0u, out + (blockIdx.x % kMcRows) * kMcSlots + 8);
}

__global__ void __launch_bounds__(kBlock) mc_grain_roots_kernel(const uint32_t* list, uint32_t count,
                                                                 uint32_t* parent, const uint32_t* size,
                                                                 const uint32_t* hi, unsigned long long* out) {
    // grains, largest, then count / cells / largest of the high and the low mobility class
    unsigned long long v[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    const uint32_t j = blockIdx.x * kBlock + threadIdx.x;
    if (j < count) {
        const uint32_t g = list[j];
        const uint32_t r = uf_find(parent, g);
        parent[g] = r;  // flatten for label download (same value any concurrent reader could see)
        if (r == g) {
            const unsigned long long sz = size[g];
            v[0] = 1;
            v[1] = sz;
            if (hi != nullptr) {
                const int c = (2ULL * hi[g] >= sz) ? 0 : 1;  // McStage.cpp:145
                v[2 + 3 * c] = 1;
                v[3 + 3 * c] = sz;
                v[4 + 3 * c] = sz;
            }
        }
    }
    block_fold(v, (1u << 1) | (1u << 4) | (1u << 7), out + (blockIdx.x % kMcRows) * kMcSlots);
}

template <typename S>
__global__ void __launch_bounds__(kBlock) mc_download_kernel(uint64_t n, const S* spin, int32_t* out) {
    for (uint64_t i = blockIdx.x * static_cast<uint64_t>(kBlock) + threadIdx.x; i < n;
         i += static_cast<uint64_t>(gridDim.x) * kBlock)
        out[i] = static_cast<int32_t>(spin[i]);
}

template <typename S>
__global__ void __launch_bounds__(kBlock) mc_labels_kernel(uint64_t n, const S* spin, const uint32_t* parent,
                                                           uint32_t* out) {
    for (uint64_t i = blockIdx.x * static_cast<uint64_t>(kBlock) + threadIdx.x; i < n;
         i += static_cast<uint64_t>(gridDim.x) * kBlock)
        out[i] = spin[i] != 0 ? parent[i] : 0xFFFFFFFFu;
}

// ------------------------------------------------------------------------------------------------
// launchers
// ------------------------------------------------------------------------------------------------

namespace {
inline unsigned grid_for(uint64_t n, unsigned cap = 65535u * 4u) {
    const uint64_t g = (n + kBlock - 1) / kBlock;
    return static_cast<unsigned>(g < 1 ? 1 : (g > cap ? cap : g));
}
// One block per item (row, segment), at most cap blocks.
inline unsigned blocks_for(uint64_t items, unsigned cap) {
    return static_cast<unsigned>(items < 1 ? 1 : (items > cap ? cap : items));
}
}  // namespace

void ca_launch_terrain(const CaParams& P, const CaDev& D, const int32_t* height, hipStream_t s) {
    const uint64_t n = static_cast<uint64_t>(P.nx) * P.ny * P.nz;
    hipLaunchKernelGGL(ca_terrain_kernel, dim3(grid_for(n, 16384)), dim3(kBlock), 0, s, P, D, height);
}

void ca_launch_vent_tiles(const CaParams& P, const CaDev& D, hipStream_t s) {
    hipLaunchKernelGGL(ca_vent_tiles_kernel, dim3(grid_for(static_cast<uint64_t>(P.nvent_slots))), dim3(kBlock), 0, s, P, D);
}

void ca_launch_seeds(const CaParams&, const CaDev& D, const uint32_t* cells, const int32_t* flows, const float* temps,
                     int n, hipStream_t s) {
    hipLaunchKernelGGL(ca_seeds_kernel, dim3(1), dim3(64), 0, s, D, cells, flows, temps, n);
}

// The host's hold switch (CaCtrl::hold); switching it off releases a held simulation (it steps on).
__global__ void ca_hold_kernel(CaCtrl* c, int hold) {
    if (threadIdx.x != 0) return;
    c->hold = hold;
    if (!hold && c->stop == kStopHeld) c->stop = 0;
}
void ca_launch_hold(CaCtrl* c, int hold, hipStream_t s) {
    hipLaunchKernelGGL(ca_hold_kernel, dim3(1), dim3(64), 0, s, c, hold);
}

void ca_launch_init_stop(const CaParams& P, const CaDev& D, hipStream_t s) {
    hipLaunchKernelGGL(ca_init_stop_kernel, dim3(1), dim3(64), 0, s, P, D);
}

void ca_launch_init_flags(const CaParams& P, const CaDev& D, hipStream_t s) {
    const unsigned g = P.ntiles < 2048 ? P.ntiles : 2048;
    hipLaunchKernelGGL(ca_init_flags_kernel, dim3(g ? g : 1), dim3(kBlock), 0, s, P, D);
}

// A launch whose own dispatch records `start` / `stop` when they are given (kernel time without the launch gaps).
template <typename Kernel, typename... Args>
void launch(Kernel kernel, dim3 grid, hipStream_t s, hipEvent_t start, hipEvent_t stop, Args... args) {
    if (!start && !stop)
        hipLaunchKernelGGL(kernel, grid, dim3(kBlock), 0, s, args...);
    else
        hipExtLaunchKernelGGL(kernel, grid, dim3(kBlock), 0, s, start, stop, 0, args...);
}

// Events of kernel k of a step: ev[2k] recorded at the start of its first launch, ev[2k + 1] at the end of its last
// (a kernel launched once per tile shape); rec[k] tells that the pair was recorded.
struct StepEvents {
    hipEvent_t* ev;
    bool* rec;
    hipEvent_t start(int k, bool first) const {
        if (!ev || !first) return nullptr;
        rec[k] = true;
        return ev[2 * k];
    }
    hipEvent_t stop(int k, bool last) const { return ev && last ? ev[2 * k + 1] : nullptr; }
};

template <int TY, int TZ>
void launch_descent(int shape, dim3 g, hipStream_t s, hipEvent_t a, hipEvent_t b, const BatchDev& B, int par,
                    uint32_t fin_blocks, int desc) {
    launch(ca_descent_kernel<TY, TZ>, g, s, a, b, B, shape, par, fin_blocks, desc);
}
template <int TY, int TZ, class F>
void launch_rules(int shape, dim3 g, hipStream_t s, hipEvent_t a, hipEvent_t b, const BatchDev& B, int par) {
    launch(ca_rules_kernel<TY, TZ, F>, g, s, a, b, B, shape, par);
}
void descent_shape(int k, dim3 g, hipStream_t s, hipEvent_t a, hipEvent_t b, const BatchDev& B, int par,
                   uint32_t fin_blocks, int desc) {
    switch (k) {
        case 0: launch_descent<4, 4>(0, g, s, a, b, B, par, fin_blocks, desc); break;
        case 1: launch_descent<2, 8>(1, g, s, a, b, B, par, fin_blocks, desc); break;
        case 2: launch_descent<8, 2>(2, g, s, a, b, B, par, fin_blocks, desc); break;
        case 3: launch_descent<1, 16>(3, g, s, a, b, B, par, fin_blocks, desc); break;
        default: launch_descent<16, 1>(4, g, s, a, b, B, par, fin_blocks, desc); break;
    }
}
template <class F>
void rules_shape(int k, dim3 g, hipStream_t s, hipEvent_t a, hipEvent_t b, const BatchDev& B, int par) {
    switch (k) {
        case 0: launch_rules<4, 4, F>(0, g, s, a, b, B, par); break;
        case 1: launch_rules<2, 8, F>(1, g, s, a, b, B, par); break;
        case 2: launch_rules<8, 2, F>(2, g, s, a, b, B, par); break;
        case 3: launch_rules<1, 16, F>(3, g, s, a, b, B, par); break;
        default: launch_rules<16, 1, F>(4, g, s, a, b, B, par); break;
    }
}

// One step.  Kernel order: [ca_front] ca_descent (+ the previous step's finish when L.fin_prev) ca_rules [ca_finish];
// ca_front and ca_finish only when a member needs the scan.  Every kernel of a step reads the
// list copy L.par.  Without lava mode there is no D(n) to compute, but ca_descent still runs (it clears the previous
// copy's counts and does the pending finish).
void ca_launch_step(const BatchDev& B, const StepLaunch& L, hipStream_t s, hipEvent_t* ev, bool* rec) {
    const StepEvents E{ev, rec};
    if (L.scan) launch(ca_front_kernel, dim3(grid_for(B.ntiles)), s, E.start(0, true), E.stop(0, true), B, L.shapes, L.par);
    int shapes[kShapes], ns = 0;
    for (int k = 0; k < kShapes; ++k)
        if ((L.shapes >> k) & 1u) shapes[ns++] = k;
    // the first descent launch clears the counts and, with nslots blocks before the tile blocks, does the finish
    for (int i = 0; i < ns; ++i) {
        const uint32_t fin_blocks = i == 0 && L.fin_prev ? B.nslots : 0u;
        descent_shape(shapes[i], dim3(static_cast<unsigned>(L.grid[shapes[i]]) + fin_blocks), s, E.start(1, i == 0),
                      E.stop(1, i == ns - 1), B, L.par, fin_blocks, L.desc_list ? 1 : 0);
    }
    for (int i = 0; i < ns; ++i) {
        const dim3 gt(static_cast<unsigned>(L.grid[shapes[i]]));
        if (L.flow16)
            rules_shape<int16_t>(shapes[i], gt, s, E.start(2, i == 0), E.stop(2, i == ns - 1), B, L.par);
        else
            rules_shape<int32_t>(shapes[i], gt, s, E.start(2, i == 0), E.stop(2, i == ns - 1), B, L.par);
    }
    if (L.scan) launch(ca_finish_kernel, dim3(B.nslots), s, E.start(3, true), E.stop(3, true), B);
}

void ca_launch_finish(const BatchDev& B, hipStream_t s, hipEvent_t* ev) {
    launch(ca_finish_kernel, dim3(B.nslots), s, ev ? ev[0] : nullptr, ev ? ev[1] : nullptr, B);
}

void ca_launch_seed(const BatchDev& B, uint32_t slot, uint32_t ntiles, int par, hipStream_t s) {
    hipLaunchKernelGGL(ca_seed_kernel, dim3(grid_for(ntiles)), dim3(kBlock), 0, s, B, slot, par);
}

void ca_launch_purge(const BatchDev& B, const SlotMask& slots, int par, hipStream_t s) {
    hipLaunchKernelGGL(ca_purge_kernel, dim3(1), dim3(1024), 0, s, B, slots, par);
}

void ca_launch_status(const BatchDev& B, int32_t* out, hipStream_t s) {
    hipLaunchKernelGGL(ca_status_kernel, dim3(grid_for(B.nslots, 64)), dim3(kBlock), 0, s, B, out);
}

void ca_write_sim(uint32_t index, const CaSim* sim, hipStream_t s) {
    (void)hipMemcpyToSymbolAsync(HIP_SYMBOL(g_sims), sim, sizeof(CaSim), static_cast<size_t>(index) * sizeof(CaSim),
                                 hipMemcpyHostToDevice, s);
}

__global__ void __launch_bounds__(kBlock) fill_u32_kernel(uint32_t* p, uint32_t v, uint64_t n) {
    for (uint64_t j = blockIdx.x * static_cast<uint64_t>(kBlock) + threadIdx.x; j < n;
         j += static_cast<uint64_t>(gridDim.x) * kBlock)
        p[j] = v;
}

void fill_u32(uint32_t* p, uint32_t v, uint64_t n, hipStream_t s) {
    if (n == 0) return;
    hipLaunchKernelGGL(fill_u32_kernel, dim3(grid_for(n, 4096)), dim3(kBlock), 0, s, p, v, n);
}

void ca_launch_owner(uint32_t* owner, uint32_t base, uint32_t n, uint32_t slot, hipStream_t s) {
    if (n == 0) return;
    hipLaunchKernelGGL(ca_owner_kernel, dim3(grid_for(n, 4096)), dim3(kBlock), 0, s, owner, base, n, slot);
}

void ca_launch_ff_estimate(const CaParams& P, const CaDev& D, const BatchDev& B, uint32_t slot, int shape, int par,
                           int32_t t0, hipStream_t s) {
    hipLaunchKernelGGL(ca_ff_init_kernel, dim3(1), dim3(1), 0, s, D);
    const dim3 g(static_cast<unsigned>(std::min<uint64_t>((static_cast<uint64_t>(P.ntiles) + kSub - 1) / kSub, 256u) * kSub));
    hipLaunchKernelGGL(ca_ff_estimate_kernel, g, dim3(kBlock), 0, s, B, slot, shape, par, P, D, t0);
}

void ca_launch_fast_forward(const CaParams& P, const CaDev& D, const BatchDev& B, uint32_t slot, int shape, int par,
                            int32_t t0, int32_t h, bool until_solid, hipStream_t s, hipEvent_t* ev) {
    if (h <= 0) return;
    if (ev) (void)hipEventRecord(ev[0], s);
    hipLaunchKernelGGL(ca_ff_init_kernel, dim3(1), dim3(1), 0, s, D);
    fill_u32(reinterpret_cast<uint32_t*>(D.hist_active + t0), 0u, 2ull * static_cast<uint64_t>(h), s);
    fill_u32(reinterpret_cast<uint32_t*>(D.hist_changed + t0), 0u, 2ull * static_cast<uint64_t>(h), s);
    // the step kernels' grid for the simulation's tiles: a multiple of kSub blocks, at most 32 per CU-sized share
    const dim3 g(static_cast<unsigned>(std::min<uint64_t>((static_cast<uint64_t>(P.ntiles) + kSub - 1) / kSub, 256u) * kSub));
    hipLaunchKernelGGL(ca_ff_tiles_kernel<false>, g, dim3(kBlock), 0, s, B, slot, shape, par, P, D, t0, h);
    hipLaunchKernelGGL(ca_ff_scan_kernel, dim3(1), dim3(kBlock), 0, s, P, D, t0, h, until_solid ? 1 : 0);
    hipLaunchKernelGGL(ca_ff_tiles_kernel<true>, g, dim3(kBlock), 0, s, B, slot, shape, par, P, D, t0, h);
    hipLaunchKernelGGL(ca_ff_relabel_kernel, dim3(grid_for(B.subcap, 64), kSub), dim3(kBlock), 0, s, B, D, slot, shape,
                       par);
    if (ev) (void)hipEventRecord(ev[1], s);
}

#define VCA_SPIN_DISPATCH(bytes, CALL)                 \
    do {                                               \
        if ((bytes) == 1) {                            \
            using S = uint8_t;                         \
            CALL;                                      \
        } else if ((bytes) == 2) {                     \
            using S = uint16_t;                        \
            CALL;                                      \
        } else {                                       \
            using S = int32_t;                         \
            CALL;                                      \
        }                                              \
    } while (0)

void mc_launch_color(const McParams& P, void* spin, int bytes, const uint32_t* list, const uint64_t* mob_thr,
                     uint32_t count, uint64_t prefix, const uint64_t* accept, unsigned long long* counters,
                     hipStream_t s) {
    if (count == 0) return;
    VCA_SPIN_DISPATCH(bytes, hipLaunchKernelGGL(mc_color_kernel<S>, dim3(grid_for(count, 4096u)), dim3(kBlock), 0, s,
                                                P, static_cast<S*>(spin), list, mob_thr, count, prefix, accept,
                                                counters));
}

void mc_launch_sweeps(const McParams& P, void* spin, int bytes, const uint32_t* list, const uint64_t* mob_thr,
                      const McColours& C, uint64_t sweep0, int32_t nsweeps, const uint64_t* accept,
                      unsigned long long* counters, McGridBar* bar, uint64_t limit, unsigned grid, hipStream_t s) {
    if (nsweeps <= 0 || grid == 0) return;
    VCA_SPIN_DISPATCH(bytes, hipLaunchKernelGGL(mc_sweeps_kernel<S>, dim3(grid), dim3(kBlock), 0, s, P,
                                                static_cast<S*>(spin), list, mob_thr, C, sweep0, nsweeps, accept,
                                                counters, bar, limit));
}

void mc_launch_random(const McParams& P, void* spin, int bytes, const uint32_t* list, const uint64_t* mob_thr,
                      uint32_t count, uint64_t prefix, const uint64_t* accept, unsigned long long* counters,
                      hipStream_t s) {
    if (count == 0) return;
    VCA_SPIN_DISPATCH(bytes, hipLaunchKernelGGL(mc_random_kernel<S>, dim3(1), dim3(64), 0, s, P, static_cast<S*>(spin),
                                                list, mob_thr, count, prefix, accept, counters));
}

void mc_launch_init_spins(const McParams& P, void* spin, int bytes, const uint32_t* list, uint32_t count,
                          hipStream_t s) {
    if (count == 0) return;
    VCA_SPIN_DISPATCH(bytes, hipLaunchKernelGGL(mc_init_kernel<S>, dim3(grid_for(count, 1u << 30)), dim3(kBlock), 0, s,
                                                P, static_cast<S*>(spin), list, count));
}

void mc_launch_energy(const McParams& P, const void* spin, int bytes, const uint32_t* list, uint32_t count,
                      uint64_t prefix, unsigned long long* unlike, hipStream_t s) {
    if (count == 0) return;
    VCA_SPIN_DISPATCH(bytes, hipLaunchKernelGGL(mc_energy_kernel<S>, dim3(grid_for(count, 1u << 30)), dim3(kBlock), 0,
                                                s, P, static_cast<const S*>(spin), list, count, prefix, unlike));
}

void mc_launch_grains(const McParams& P, const void* spin, int bytes, const uint32_t* list, uint32_t count,
                      uint32_t* parent, uint32_t* size, uint32_t* hi, const float* mobility, unsigned long long* out,
                      hipStream_t s) {
    if (count == 0) return;
    const unsigned g = grid_for(count, 1u << 30);
    hipLaunchKernelGGL(mc_grain_init_kernel, dim3(g), dim3(kBlock), 0, s, list, count, parent, size, hi);
    VCA_SPIN_DISPATCH(bytes, hipLaunchKernelGGL(mc_grain_union_kernel<S>, dim3(g), dim3(kBlock), 0, s, P,
                                                static_cast<const S*>(spin), list, count, parent));
    hipLaunchKernelGGL(mc_grain_flatten_kernel, dim3(g), dim3(kBlock), 0, s, list, count, parent, size, hi, mobility,
                       out);
    hipLaunchKernelGGL(mc_grain_roots_kernel, dim3(g), dim3(kBlock), 0, s, list, count, parent, size, hi, out);
}

void mc_launch_download(const McParams& P, const void* spin, int bytes, int32_t* out, hipStream_t s) {
    const uint64_t n = static_cast<uint64_t>(P.nx) * P.ny * P.nz;
    VCA_SPIN_DISPATCH(bytes, hipLaunchKernelGGL(mc_download_kernel<S>, dim3(grid_for(n, 16384)), dim3(kBlock), 0, s, n,
                                                static_cast<const S*>(spin), out));
}

void mc_launch_labels(const McParams& P, const void* spin, int bytes, const uint32_t* parent, uint32_t* out,
                      hipStream_t s) {
    const uint64_t n = static_cast<uint64_t>(P.nx) * P.ny * P.nz;
    VCA_SPIN_DISPATCH(bytes, hipLaunchKernelGGL(mc_labels_kernel<S>, dim3(grid_for(n, 16384)), dim3(kBlock), 0, s, n,
                                                static_cast<const S*>(spin), parent, out));
}

// ------------------------------------------------------------------------------------------------
// After the CA: flow sums of the final state and the MC domain lists, on the device
// ------------------------------------------------------------------------------------------------

// Calls fn(i, t) for every LAVA / BASALT cell i (type t) of the n cells of `type` (16-byte aligned): each lane
// reads 16 cells with one load and skips a word of four without a domain byte in a few ALU operations, so a
// pass over the grid keeps many loads in flight instead of waiting on one byte per lane.
template <typename Fn>
__device__ inline void for_domain_cells(uint64_t n, const uint8_t* type, Fn&& fn) {
    const uint64_t chunks = (n + 15) / 16;
    for (uint64_t q = blockIdx.x * static_cast<uint64_t>(kBlock) + threadIdx.x; q < chunks;
         q += static_cast<uint64_t>(gridDim.x) * kBlock) {
        const uint64_t i0 = q * 16;
        uint32_t w[4];
        if (i0 + 16 <= n) {
            const uint4 v = reinterpret_cast<const uint4*>(type)[q];
            w[0] = v.x;
            w[1] = v.y;
            w[2] = v.z;
            w[3] = v.w;
        } else {  // the last, partial chunk: missing cells read as AIR
            uint64_t lo = 0, hi = 0;
            for (uint64_t i = i0; i < n; ++i) {
                const uint32_t sft = static_cast<uint32_t>(i - i0);
                const uint64_t t = type[i];
                if (sft < 8)
                    lo |= t << (8 * sft);
                else
                    hi |= t << (8 * (sft - 8));
            }
            w[0] = static_cast<uint32_t>(lo);
            w[1] = static_cast<uint32_t>(lo >> 32);
            w[2] = static_cast<uint32_t>(hi);
            w[3] = static_cast<uint32_t>(hi >> 32);
        }
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            // a byte of y is zero iff that cell is LAVA (2) or BASALT (3); the zero-byte test is exact
            const uint32_t y = (w[k] ^ 0x02020202u) & 0xFEFEFEFEu;
            if (((y - 0x01010101u) & ~y & 0x80808080u) == 0) continue;
#pragma unroll
            for (int b = 0; b < 4; ++b) {
                const uint32_t t = (w[k] >> (8 * b)) & 0xFFu;
                if ((t & 0xFEu) == 2u) fn(i0 + 4 * k + b, t);
            }
        }
    }
}

// flowSums (Simulation.cpp:408-466): per flowId f in 1..nflows the LAVA and BASALT cells and the largest squared
// horizontal distance to the flow's vent.  With at most kFlowLds flows a block gathers its sums in LDS and adds
// them to the totals once (per-wave atomics on the few total addresses serialised: 2.1 ms on a 512x512x256 grid);
// with more flows each cell goes to the totals directly, spread over many addresses.
constexpr int kFlowLds = 256;

__global__ void __launch_bounds__(kBlock) ca_flow_sums_kernel(CaParams P, CaDev D, int par, const int32_t* vent_xy,
                                                               int32_t nflows, unsigned long long* lava,
                                                               unsigned long long* basalt,
                                                               unsigned long long* maxd2) {
    __shared__ uint32_t s_cnt[2][kFlowLds + 1];  // LAVA, BASALT
    __shared__ unsigned long long s_d2[kFlowLds + 1];
    const bool in_lds = nflows <= kFlowLds;
    if (in_lds) {
        for (int f = static_cast<int>(threadIdx.x); f <= nflows; f += kBlock) {
            s_cnt[0][f] = 0;
            s_cnt[1][f] = 0;
            s_d2[f] = 0;
        }
        __syncthreads();
    }
    const uint64_t nx = static_cast<uint64_t>(P.nx), ny = static_cast<uint64_t>(P.ny);
    for_domain_cells(nx * ny * static_cast<uint64_t>(P.nz), D.type[par], [&](uint64_t i, uint32_t t) {
        const int32_t f = D.flow[i];
        if (f < 1 || f > nflows) return;
        const uint64_t r = i / nx;
        const long long dx = static_cast<long long>(i - r * nx) - vent_xy[2 * (f - 1)];
        const long long dy = static_cast<long long>(r % ny) - vent_xy[2 * (f - 1) + 1];
        const unsigned long long d2 = static_cast<unsigned long long>(dx * dx + dy * dy);
        if (in_lds) {
            atomicAdd(&s_cnt[t == kLava ? 0 : 1][f], 1u);
            atomicMax(&s_d2[f], d2);
        } else {
            atomicAdd(t == kLava ? &lava[f] : &basalt[f], 1ULL);
            atomicMax(&maxd2[f], d2);
        }
    });
    if (!in_lds) return;
    __syncthreads();
    for (int f = 1 + static_cast<int>(threadIdx.x); f <= nflows; f += kBlock) {
        const uint32_t nl = s_cnt[0][f], nb = s_cnt[1][f];
        if (nl) atomicAdd(&lava[f], static_cast<unsigned long long>(nl));
        if (nb) atomicAdd(&basalt[f], static_cast<unsigned long long>(nb));
        if (nl | nb) atomicMax(&maxd2[f], s_d2[f]);
    }
}

// Before a read of the temperatures after `steps` steps (copy par = steps & 1): every BASALT cell's last temperature
// into that copy (the step that solidified the cell at step stamp s wrote it into copy (s + 1) & 1 only).
__global__ void __launch_bounds__(kBlock) ca_final_temps_kernel(uint64_t n, const uint8_t* type, const uint16_t* solid,
                                                                float* t0, float* t1, int par) {
    float* dst = par ? t1 : t0;
    const float* src = par ? t0 : t1;
    for (uint64_t i = blockIdx.x * static_cast<uint64_t>(kBlock) + threadIdx.x; i < n;
         i += static_cast<uint64_t>(gridDim.x) * kBlock)
        if (type[i] == kBasalt && ((solid[i] + 1) & 1) != par) dst[i] = src[i];
}

// Marks seen[|dx|] for the LAVA / BASALT cells of flow `flow` whose squared distance to (vx, vy) is d2: which
// lattice points of that circle the flow reaches (the reach is the largest hypot among them).
__global__ void __launch_bounds__(kBlock) ca_flow_points_kernel(CaParams P, CaDev D, int par, int32_t flow, int32_t vx,
                                                                 int32_t vy, unsigned long long d2, uint32_t* seen,
                                                                 uint32_t cap) {
    const uint32_t rows = static_cast<uint32_t>(P.ny) * static_cast<uint32_t>(P.nz);
    for (uint32_t r = blockIdx.x; r < rows; r += gridDim.x) {
        const int y = static_cast<int>(r % static_cast<uint32_t>(P.ny));
        const uint32_t row = r * static_cast<uint32_t>(P.nx);
        for (int x = static_cast<int>(threadIdx.x); x < P.nx; x += kBlock) {
            const uint8_t t = D.type[par][row + static_cast<uint32_t>(x)];
            if (t != kLava && t != kBasalt) continue;
            if (D.flow[row + static_cast<uint32_t>(x)] != flow) continue;
            const long long dx = x - vx, dy = y - vy;
            if (static_cast<unsigned long long>(dx * dx + dy * dy) != d2) continue;
            const unsigned long long a = static_cast<unsigned long long>(dx < 0 ? -dx : dx);
            if (a < cap) seen[a] = 1u;
        }
    }
}

// Cooling mobility (CoolingMobility.cpp:8-129), three passes over the CA result: red[0] = end step (largest
// solidStep / arrivalStep of the domain), red[1] / red[2] = smallest / largest time in the liquid state, then
// the mobility of every cell (0 outside the domain).  Same double arithmetic as the host formula.
__device__ inline bool mc_domain_cell(uint8_t t) { return t == kBasalt || t == kLava; }

// Block-wide min and max of one value per thread (every thread of the block calls; the results are valid in
// thread 0).
__device__ inline void block_min_max(uint32_t& lo, uint32_t& hi) {
    __shared__ uint32_t s_lo[kBlock / 64], s_hi[kBlock / 64];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        const uint32_t a = __shfl_xor(lo, o), b = __shfl_xor(hi, o);
        lo = a < lo ? a : lo;
        hi = b > hi ? b : hi;
    }
    if ((threadIdx.x & 63u) == 0) {
        s_lo[threadIdx.x >> 6] = lo;
        s_hi[threadIdx.x >> 6] = hi;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int k = 1; k < kBlock / 64; ++k) {
            lo = s_lo[k] < lo ? s_lo[k] : lo;
            hi = s_hi[k] > hi ? s_hi[k] : hi;
        }
    }
}

// One atomic per block on each of the shared results (per-wave atomics on one address serialised).
__global__ void __launch_bounds__(kBlock) mc_cool_end_kernel(uint64_t n, const uint8_t* type, const uint16_t* solid,
                                                              const uint16_t* arrival, uint32_t* red) {
    uint32_t lo = 0, e = 0;
    for_domain_cells(n, type, [&](uint64_t i, uint32_t) {
        const uint32_t a = solid[i], b = arrival[i];
        e = a > e ? a : e;
        e = b > e ? b : e;
    });
    block_min_max(lo, e);
    if (threadIdx.x == 0 && e) atomicMax(&red[0], e);
}

__device__ inline uint32_t mc_cool_time(uint8_t t, uint32_t solid, uint32_t arrival, uint32_t end) {
    const uint32_t sd = t == kBasalt ? solid : end;
    return sd > arrival ? sd - arrival : 0u;
}

__global__ void __launch_bounds__(kBlock) mc_cool_range_kernel(uint64_t n, const uint8_t* type, const uint16_t* solid,
                                                                const uint16_t* arrival, uint32_t* red) {
    const uint32_t end = red[0];
    uint32_t lo = 0xFFFFFFFFu, hi = 0;
    for_domain_cells(n, type, [&](uint64_t i, uint32_t t) {
        const uint32_t tm = mc_cool_time(static_cast<uint8_t>(t), solid[i], arrival[i], end);
        lo = tm < lo ? tm : lo;
        hi = tm > hi ? tm : hi;
    });
    block_min_max(lo, hi);
    if (threadIdx.x == 0) {
        if (lo != 0xFFFFFFFFu) atomicMin(&red[1], lo);
        if (hi) atomicMax(&red[2], hi);
    }
}

__global__ void __launch_bounds__(kBlock) mc_cool_mobility_kernel(uint64_t n, const uint8_t* type,
                                                                   const uint16_t* solid, const uint16_t* arrival,
                                                                   const uint32_t* red, double mobility_min,
                                                                   float* mob) {
    const uint32_t end = red[0], tmin = red[1], tmax = red[2];
    for (uint64_t i = blockIdx.x * static_cast<uint64_t>(kBlock) + threadIdx.x; i < n;
         i += static_cast<uint64_t>(gridDim.x) * kBlock) {
        const uint8_t t = type[i];
        float m = 0.0f;
        if (mc_domain_cell(t)) {
            if (tmax == tmin) {
                m = 1.0f;
            } else {
                const uint32_t tm = mc_cool_time(t, solid[i], arrival[i], end);
                const double norm = static_cast<double>(tm - tmin) / static_cast<double>(tmax - tmin);
                m = static_cast<float>(mobility_min + (1.0 - mobility_min) * norm);
            }
        }
        mob[i] = m;
    }
}

// Exact integer form of `toUniform(r) >= m` (the mobility skip) for a float mobility m: skip iff (r >> 11) >= thr.
__device__ inline uint64_t mobility_threshold_dev(float m) {
    const double two53 = 9007199254740992.0;
    if (m != m) return 1ULL << 53;  // u >= NaN is false: never skipped
    if (!(m > 0.0f)) return 0;      // u >= m for every u in [0, 1)
    const double v = ceil(static_cast<double>(m) * two53);
    return v >= two53 ? (1ULL << 53) : static_cast<uint64_t>(v);
}

// Domain lists of the MC, colour-major with ascending global index inside each colour (buildColors,
// PottsMC.cpp:145-224), built by an ordered compaction: per-segment colour counts, a scan per colour, then each
// segment writes its cells at their ranks.  Segments are kMcSeg consecutive cells, processed 256 at a time.
constexpr int kMcSeg = 4096;

__device__ inline bool mc_list_member(const McListSrc& L, uint64_t i) {
    if (L.mode == 0) return true;
    const uint8_t v = L.dom[i];
    return L.mode == 1 ? v != 0 : mc_domain_cell(v);
}
__device__ inline int mc_list_colour(const McListSrc& L, uint64_t i) {
    if (L.ncol <= 1) return 0;
    // i < n <= 2e9: 32-bit division
    const uint32_t nx = static_cast<uint32_t>(L.dims[0]), ny = static_cast<uint32_t>(L.dims[1]);
    const uint32_t ii = static_cast<uint32_t>(i);
    const uint32_t x = ii % nx, q = ii / nx;
    const uint32_t y = q % ny, z = q / ny;
    if (L.parity) return static_cast<int>((x + y + z) & 1u);
    const uint32_t c[3] = {x, y, z};
    int k = 0, mul = 1;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const int cls = (L.per[a] && (L.dims[a] & 1) && c[a] == static_cast<uint32_t>(L.dims[a] - 1)) ? 2
                                                                                                     : static_cast<int>(c[a] & 1u);
        k += cls * mul;
        mul *= 3;
    }
    return L.index[k];
}

__global__ void __launch_bounds__(kBlock) mc_list_count_kernel(McListSrc L, uint32_t* cnt) {
    __shared__ uint32_t s_cnt[27];
    const uint64_t nseg = (L.n + kMcSeg - 1) / kMcSeg;
    const int lane = static_cast<int>(threadIdx.x & 63u);
    for (uint64_t g = blockIdx.x; g < nseg; g += gridDim.x) {
        __syncthreads();
        if (threadIdx.x < 27) s_cnt[threadIdx.x] = 0;
        __syncthreads();
        for (int r = 0; r < kMcSeg; r += kBlock) {
            const uint64_t i = g * kMcSeg + static_cast<uint64_t>(r) + threadIdx.x;
            const bool in = i < L.n && mc_list_member(L, i);
            const int c = in ? mc_list_colour(L, i) : 0;
            unsigned long long pend = __ballot(in);
            while (pend) {
                const int leader = __ffsll(static_cast<long long>(pend)) - 1;
                const int cl = __shfl(c, leader);
                const unsigned long long m = __ballot(in && c == cl);
                if (lane == leader) atomicAdd(&s_cnt[cl], static_cast<uint32_t>(__popcll(m)));
                pend &= ~m;
            }
        }
        __syncthreads();
        if (threadIdx.x < static_cast<unsigned>(L.ncol)) cnt[static_cast<uint64_t>(threadIdx.x) * nseg + g] = s_cnt[threadIdx.x];
    }
}

// Exclusive scan of each colour's segment counts (one block per colour); totals[c] = the colour's cells.
__global__ void __launch_bounds__(1024) mc_list_scan_kernel(uint32_t* cnt, uint64_t nseg, uint32_t* totals) {
    __shared__ uint32_t s_w[16];
    __shared__ uint32_t s_carry;
    uint32_t* v = cnt + static_cast<uint64_t>(blockIdx.x) * nseg;
    const int lane = static_cast<int>(threadIdx.x & 63u), w = static_cast<int>(threadIdx.x >> 6);
    if (threadIdx.x == 0) s_carry = 0;
    __syncthreads();
    for (uint64_t b = 0; b < nseg; b += 1024) {
        const uint64_t j = b + threadIdx.x;
        const uint32_t x = j < nseg ? v[j] : 0u;
        uint32_t inc = x;  // inclusive scan inside the wave
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            const uint32_t y = __shfl_up(inc, o);
            if (lane >= o) inc += y;
        }
        if (lane == 63) s_w[w] = inc;
        __syncthreads();
        uint32_t before = s_carry;
        for (int k = 0; k < w; ++k) before += s_w[k];
        if (j < nseg) v[j] = before + inc - x;
        __syncthreads();
        if (threadIdx.x == 1023) s_carry = before + inc;
        __syncthreads();
    }
    if (threadIdx.x == 0) totals[blockIdx.x] = s_carry;
}

__global__ void __launch_bounds__(kBlock) mc_list_scatter_kernel(McListSrc L, const uint32_t* off,
                                                                  const uint32_t* colour_base, uint32_t* list,
                                                                  const float* mob, uint64_t* thr) {
    __shared__ uint32_t s_off[27];
    __shared__ uint32_t s_wt[27][kBlock / 64];
    const uint64_t nseg = (L.n + kMcSeg - 1) / kMcSeg;
    const int lane = static_cast<int>(threadIdx.x & 63u), w = static_cast<int>(threadIdx.x >> 6);
    const unsigned long long below = (1ULL << lane) - 1ULL;
    for (uint64_t g = blockIdx.x; g < nseg; g += gridDim.x) {
        __syncthreads();
        if (threadIdx.x < static_cast<unsigned>(L.ncol))
            s_off[threadIdx.x] = colour_base[threadIdx.x] + off[static_cast<uint64_t>(threadIdx.x) * nseg + g];
        for (int r = 0; r < kMcSeg; r += kBlock) {
            const uint64_t i = g * kMcSeg + static_cast<uint64_t>(r) + threadIdx.x;
            const bool in = i < L.n && mc_list_member(L, i);
            const int c = in ? mc_list_colour(L, i) : 0;
            if (lane < L.ncol) s_wt[lane][w] = 0;  // this wave's column (in program order before its writes)
            uint32_t rank = 0;
            unsigned long long pend = __ballot(in);
            while (pend) {
                const int leader = __ffsll(static_cast<long long>(pend)) - 1;
                const int cl = __shfl(c, leader);
                const unsigned long long m = __ballot(in && c == cl);
                if (in && c == cl) rank = static_cast<uint32_t>(__popcll(m & below));
                if (lane == leader) s_wt[cl][w] = static_cast<uint32_t>(__popcll(m));
                pend &= ~m;
            }
            __syncthreads();
            if (in) {
                uint32_t pos = s_off[c] + rank;
                for (int k = 0; k < w; ++k) pos += s_wt[c][k];
                list[pos] = static_cast<uint32_t>(i);
                if (thr) thr[pos] = mobility_threshold_dev(mob[i]);
            }
            __syncthreads();
            if (threadIdx.x < static_cast<unsigned>(L.ncol)) {
                uint32_t add = 0;
                for (int k = 0; k < kBlock / 64; ++k) add += s_wt[threadIdx.x][k];
                s_off[threadIdx.x] += add;
            }
            __syncthreads();
        }
    }
}

void ca_launch_flow_sums(const CaParams& P, const CaDev& D, int par, const int32_t* vent_xy, int32_t nflows,
                         unsigned long long* lava, unsigned long long* basalt, unsigned long long* maxd2,
                         hipStream_t s) {
    const uint64_t n = static_cast<uint64_t>(P.nx) * P.ny * P.nz;
    hipLaunchKernelGGL(ca_flow_sums_kernel, dim3(grid_for((n + 15) / 16, 2048)), dim3(kBlock), 0, s, P, D, par, vent_xy,
                       nflows,
                       lava, basalt, maxd2);
}

void ca_launch_final_temps(const CaDev& D, uint64_t n, int par, hipStream_t s) {
    hipLaunchKernelGGL(ca_final_temps_kernel, dim3(grid_for(n, 8192)), dim3(kBlock), 0, s, n, D.type[par], D.solid,
                       D.temp[0], D.temp[1], par);
}

void ca_launch_flow_points(const CaParams& P, const CaDev& D, int par, int32_t flow, int32_t vx, int32_t vy,
                           unsigned long long d2, uint32_t* seen, uint32_t cap, hipStream_t s) {
    const uint64_t rows = static_cast<uint64_t>(P.ny) * P.nz;
    hipLaunchKernelGGL(ca_flow_points_kernel, dim3(blocks_for(rows, 16384)), dim3(kBlock), 0, s, P, D, par, flow, vx, vy,
                       d2, seen, cap);
}

void mc_launch_cooling(uint64_t n, const uint8_t* type, const uint16_t* solid, const uint16_t* arrival,
                       uint32_t* red, double mobility_min, float* mob, hipStream_t s) {
    const unsigned g = grid_for((n + 15) / 16, 2048);
    hipLaunchKernelGGL(mc_cool_end_kernel, dim3(g), dim3(kBlock), 0, s, n, type, solid, arrival, red);
    hipLaunchKernelGGL(mc_cool_range_kernel, dim3(g), dim3(kBlock), 0, s, n, type, solid, arrival, red);
    hipLaunchKernelGGL(mc_cool_mobility_kernel, dim3(grid_for(n, 65536)), dim3(kBlock), 0, s, n, type, solid, arrival,
                       red, mobility_min, mob);
}

uint64_t mc_list_segments(uint64_t n) { return (n + kMcSeg - 1) / kMcSeg; }

void mc_launch_list_count(const McListSrc& L, uint32_t* cnt, uint32_t* totals, hipStream_t s) {
    const uint64_t nseg = mc_list_segments(L.n);
    hipLaunchKernelGGL(mc_list_count_kernel, dim3(blocks_for(nseg, 16384)), dim3(kBlock), 0, s, L, cnt);
    hipLaunchKernelGGL(mc_list_scan_kernel, dim3(static_cast<unsigned>(L.ncol)), dim3(1024), 0, s, cnt, nseg, totals);
}

void mc_launch_list_scatter(const McListSrc& L, const uint32_t* off, const uint32_t* colour_base, uint32_t* list,
                            const float* mob, uint64_t* thr, hipStream_t s) {
    const uint64_t nseg = mc_list_segments(L.n);
    hipLaunchKernelGGL(mc_list_scatter_kernel, dim3(blocks_for(nseg, 16384)), dim3(kBlock), 0, s, L, off, colour_base,
                       list, mob, thr);
}

}  // namespace vca

// HIP kernels of the lava CA step and the Potts/Metropolis sweep, written for gfx950 (CDNA4).
//
// CA step (one synchronous step of engine/src/ca/LavaCA.hpp) = five stream-ordered launches:
//   ca_front   : active tile list — a 64x4x4 (3D) / 64x1x16 (x-z) tile is visited when it holds lava,
//                or holds AIR next to a tile with lava (the reference's "may change" set,
//                engine/src/ca/CaRules.hpp:69-79, at tile granularity).  Wave ballot + one atomic per
//                wave appends the tile ids.
//   ca_descent : D(n) of every lava cell of the visited tiles (LavaCA.hpp:421-429, 877-888).
//   ca_rules   : the rules of every visited cell read the start-of-step state only and append the cell's
//                new state to a change list (wave-aggregated append), count lava/born/changed, and set the
//                tile's lava/air flags from the new state (LavaCA.hpp:963-1048, 1109-1202).
//   ca_apply   : scatters the change list into the state (LavaCA.hpp:1052-1066).
//   ca_finish  : one lane: histories, stop rule, counter reset (LavaCA.hpp:1206-1229).
// Every kernel first reads the device stop flag and returns when a stop rule already fired, so the host
// can enqueue many steps without waiting.
//
// MC checkerboard sweep (engine/src/mc/PottsMC.cpp:393-514): one launch per colour over the colour's list
// of domain cells; Metropolis and mobility decisions are exact integer compares against host-built 53-bit
// thresholds (no exp on the device).
//
// Temperatures use only IEEE float subtract / compare (built with -ffp-contract=off).
#include <hip/hip_runtime.h>

#include "vca_internal.h"

namespace vca {

namespace {

__device__ inline bool wrap_axis(int& v, int n, int per) {
    if (static_cast<unsigned>(v) < static_cast<unsigned>(n)) return true;
    if (!per) return false;
    v = v < 0 ? v + n : v - n;
    return true;
}

__device__ inline unsigned long long wave_sum(unsigned long long v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    return v;
}

// Wave-aggregated slot reservation: every lane of the wave must call this (pred may differ per lane).
__device__ inline uint32_t wave_append(bool pred, uint32_t* counter) {
    const unsigned long long mask = __ballot(pred);
    const int lane = static_cast<int>(threadIdx.x & 63u);
    uint32_t base = 0;
    if (mask != 0) {
        const int leader = __ffsll(static_cast<long long>(mask)) - 1;
        if (lane == leader) base = atomicAdd(counter, static_cast<uint32_t>(__popcll(mask)));
        base = __shfl(base, leader);
    }
    return base + static_cast<uint32_t>(__popcll(mask & ((1ULL << lane) - 1ULL)));
}

// ------------------------------------------------------------------------------------------------
// CA helpers
// ------------------------------------------------------------------------------------------------

struct TileCoord {
    int x0, y0, z0;
};
__device__ inline TileCoord tile_origin(const CaParams& P, uint32_t t) {
    const int tx = static_cast<int>(t % static_cast<uint32_t>(P.ntx));
    const uint32_t r = t / static_cast<uint32_t>(P.ntx);
    const int ty = static_cast<int>(r % static_cast<uint32_t>(P.nty));
    const int tz = static_cast<int>(r / static_cast<uint32_t>(P.nty));
    return TileCoord{tx * kTileX, ty * P.ty, tz * P.tz};
}
// Cell k (0..3) of thread threadIdx.x inside the tile.
__device__ inline void tile_cell(const CaParams& P, const TileCoord& o, int k, int& x, int& y, int& z) {
    const int lx = static_cast<int>(threadIdx.x & 63u);
    const int row = static_cast<int>(threadIdx.x >> 6) + 4 * k;
    x = o.x0 + lx;
    y = o.y0 + row % P.ty;
    z = o.z0 + row / P.ty;
}

// hasSupport (LavaCA.hpp:390-399); col = x + nx*y of an interior column.
__device__ inline bool has_support(const CaDev& D, uint32_t col, uint32_t plane, int z) {
    if (z == 0) return true;
    const uint8_t b = D.type[col + plane * static_cast<uint32_t>(z - 1)];
    if (b == kRock || b == kBasalt) return true;
    return b == kLava && D.spill != nullptr && z - 1 < D.spill[col];
}

// dropDepth (LavaCA.hpp:404-414)
__device__ int drop_depth(const CaDev& D, uint32_t col, uint32_t plane, int z) {
    for (int d = 0; d <= kMaxDrop && z - d >= 0; ++d) {
        if (d > 0 && D.type[col + plane * static_cast<uint32_t>(z - d)] != kAir) return -1;
        if (has_support(D, col, plane, z - d)) return d;
    }
    return kMaxDrop;
}

// Candidate vote with 4 register slots (engine/src/ca/CaRules.hpp:127-183).  More than 4 distinct flowIds
// among the candidates sets `overflow`; the caller then re-runs the gather with VoteWide.
struct Vote4 {
    int n = 0;
    bool overflow = false;
    int32_t id[4];
    int cnt[4];
    float mt[4];
    __device__ void add(int32_t f, float t) {
        bool found = false;
#pragma unroll
        for (int s = 0; s < 4; ++s) {
            if (s < n && id[s] == f) {
                ++cnt[s];
                if (mt[s] < t) mt[s] = t;  // std::max(a, b) == (a < b ? b : a)
                found = true;
            }
        }
        if (found) return;
        if (n == 4) {
            overflow = true;
            return;
        }
#pragma unroll
        for (int s = 0; s < 4; ++s) {
            if (s == n) {
                id[s] = f;
                cnt[s] = 1;
                mt[s] = t;
            }
        }
        ++n;
    }
    // CandidateVote::winner: most votes; ties -> ascending flowId, pick hash % k.
    __device__ void winner(uint64_t h_index, int32_t& flow, float& temp) const {
        int best = 0;
#pragma unroll
        for (int s = 0; s < 4; ++s)
            if (s < n && cnt[s] > best) best = cnt[s];
        int k = 0;
#pragma unroll
        for (int s = 0; s < 4; ++s) k += (s < n && cnt[s] == best) ? 1 : 0;
        int pick = 0;
        if (n > 1 && k > 1) pick = static_cast<int>(splitmix64(h_index ^ kSaltTieBreak) % static_cast<uint64_t>(k));
#pragma unroll
        for (int s = 0; s < 4; ++s) {
            if (s < n && cnt[s] == best) {
                int rank = 0;
#pragma unroll
                for (int q = 0; q < 4; ++q) rank += (q < n && cnt[q] == best && id[q] < id[s]) ? 1 : 0;
                if (rank == pick) {
                    flow = id[s];
                    temp = mt[s];
                }
            }
        }
    }
};

// Vote over any number of distinct flowIds (private memory; only used when Vote4 overflowed).
struct VoteWide {
    int n = 0;
    bool overflow = false;
    int32_t id[32];
    int cnt[32];
    float mt[32];
    __device__ void add(int32_t f, float t) {
        for (int s = 0; s < n; ++s) {
            if (id[s] == f) {
                ++cnt[s];
                if (mt[s] < t) mt[s] = t;
                return;
            }
        }
        id[n] = f;
        cnt[n] = 1;
        mt[n] = t;
        ++n;
    }
    __device__ void winner(uint64_t h_index, int32_t& flow, float& temp) const {
        int best = 0;
        for (int s = 0; s < n; ++s) best = cnt[s] > best ? cnt[s] : best;
        int k = 0;
        for (int s = 0; s < n; ++s) k += cnt[s] == best ? 1 : 0;
        int pick = 0;
        if (n > 1 && k > 1) pick = static_cast<int>(splitmix64(h_index ^ kSaltTieBreak) % static_cast<uint64_t>(k));
        for (int s = 0; s < n; ++s) {
            if (cnt[s] != best) continue;
            int rank = 0;
            for (int q = 0; q < n; ++q) rank += (cnt[q] == best && id[q] < id[s]) ? 1 : 0;
            if (rank == pick) {
                flow = id[s];
                temp = mt[s];
            }
        }
    }
};

// tryPonding (LavaCA.hpp:443-461)
template <class V>
__device__ void try_ponding(const CaParams& P, const CaDev& D, int x, int y, int z, uint32_t col, uint32_t plane,
                            V& v) {
    if (z > D.spill[col] || z == 0) return;
    const uint32_t b = col + plane * static_cast<uint32_t>(z - 1);
    if (D.type[b] != kLava) return;
    if (!has_support(D, col, plane, z - 1)) return;
    const int hx[4] = {-1, 1, 0, 0};
    const int hy[4] = {0, 0, -1, 1};
    const int nh = P.ny > 1 ? 4 : 2;
    for (int k = 0; k < nh; ++k) {
        int qx = x + hx[k], qy = y + hy[k];
        if (!wrap_axis(qx, P.nx, P.per[0]) || !wrap_axis(qy, P.ny, P.per[1])) return;  // VOID
        const uint8_t t = D.type[static_cast<uint32_t>(qx) + static_cast<uint32_t>(P.nx) * static_cast<uint32_t>(qy) +
                                 plane * static_cast<uint32_t>(z - 1)];
        if (t == kAir) return;
    }
    v.add(D.flow[b], D.temp[b] - P.cool_lat);
}

// gatherLavaCandidates (LavaCA.hpp:465-519) for the AIR cell (x, y, z).
template <class V>
__device__ void gather_lava(const CaParams& P, const CaDev& D, int x, int y, int z, uint32_t mask, V& v) {
    const uint32_t plane = static_cast<uint32_t>(P.nx) * static_cast<uint32_t>(P.ny);
    const uint32_t col = static_cast<uint32_t>(x) + static_cast<uint32_t>(P.nx) * static_cast<uint32_t>(y);
    int cdrop = -100;
#pragma unroll
    for (int k = 0; k < 27; ++k) {
        if (k == 13 || !((mask >> k) & 1u)) continue;
        const int dx = k % 3 - 1, dy = (k / 3) % 3 - 1, dz = k / 9 - 1;
        if (dz < 0) {
            if (dx == 0 && dy == 0 && P.ponding) try_ponding(P, D, x, y, z, col, plane, v);
            continue;
        }
        int qx = x + dx, qy = y + dy, qz = z + dz;
        if (!wrap_axis(qx, P.nx, P.per[0]) || !wrap_axis(qy, P.ny, P.per[1]) || !wrap_axis(qz, P.nz, P.per[2]))
            continue;
        const uint32_t ni = static_cast<uint32_t>(qx) + static_cast<uint32_t>(P.nx) * static_cast<uint32_t>(qy) +
                            plane * static_cast<uint32_t>(qz);
        if (D.type[ni] != kLava) continue;
        float cooling;
        if (dz > 0 && dx == 0 && dy == 0) {
            cooling = P.cool_fall;
        } else {
            const int d = D.descent[ni];
            if (d < 0) continue;
            if (cdrop == -100) cdrop = drop_depth(D, col, plane, z);
            if (dz > 0) {
                if (cdrop != 0 || d != 1) continue;
                cooling = P.cool_fall;
            } else {
                if (cdrop != d || (d == 1 && P.nb.diagonal_falls)) continue;
                cooling = P.cool_lat;
            }
        }
        v.add(D.flow[ni], D.temp[ni] - cooling);
    }
}

// Classic grain growth: every lava (grain) neighbour votes (LavaCA.hpp:1165-1202).
template <class V>
__device__ void gather_classic(const CaParams& P, const CaDev& D, int x, int y, int z, uint32_t mask, V& v) {
    const uint32_t plane = static_cast<uint32_t>(P.nx) * static_cast<uint32_t>(P.ny);
#pragma unroll
    for (int k = 0; k < 27; ++k) {
        if (k == 13 || !((mask >> k) & 1u)) continue;
        int qx = x + (k % 3 - 1), qy = y + ((k / 3) % 3 - 1), qz = z + (k / 9 - 1);
        if (!wrap_axis(qx, P.nx, P.per[0]) || !wrap_axis(qy, P.ny, P.per[1]) || !wrap_axis(qz, P.nz, P.per[2]))
            continue;
        const uint32_t ni = static_cast<uint32_t>(qx) + static_cast<uint32_t>(P.nx) * static_cast<uint32_t>(qy) +
                            plane * static_cast<uint32_t>(qz);
        if (D.type[ni] == kLava) v.add(D.flow[ni], 0.0f);
    }
}

__device__ __noinline__ bool wide_vote(const CaParams& P, const CaDev& D, int x, int y, int z, uint32_t mask,
                                        uint64_t h_index, int32_t& flow, float& temp) {
    VoteWide v;
    if (P.classic)
        gather_classic(P, D, x, y, z, mask, v);
    else
        gather_lava(P, D, x, y, z, mask, v);
    if (v.n == 0) return false;
    v.winner(h_index, flow, temp);
    return true;
}

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
        D.type[i] = z < h ? kRock : kAir;
    }
}

__global__ void ca_seeds_kernel(CaDev D, const uint32_t* cells, const int32_t* flows, const float* temps, int n) {
    // Sequential in list order, like placeSeeds (LavaCA.hpp:215-227): a later seed in the same cell wins.
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    for (int k = 0; k < n; ++k) {
        const uint32_t i = cells[k];
        D.type[i] = kLava;
        D.flow[i] = flows[k];
        D.temp[i] = temps[k];
        D.solid[i] = 0;
        D.arrival[i] = 0;
    }
}

// Tile flags and the initial LAVA / AIR counts (countInitialState, LavaCA.hpp:1252-1280).
__global__ void __launch_bounds__(kBlock) ca_init_flags_kernel(CaParams P, CaDev D) {
    for (uint32_t t = blockIdx.x; t < P.ntiles; t += gridDim.x) {
        const TileCoord o = tile_origin(P, t);
        int lava = 0, air = 0;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            int x, y, z;
            tile_cell(P, o, k, x, y, z);
            if (x < P.nx && y < P.ny && z < P.nz) {
                const uint8_t ty = D.type[static_cast<uint32_t>(x) +
                                          static_cast<uint32_t>(P.nx) * (static_cast<uint32_t>(y) +
                                                                         static_cast<uint32_t>(P.ny) * z)];
                lava += ty == kLava;
                air += ty == kAir;
            }
        }
        const unsigned long long wl = wave_sum(static_cast<unsigned long long>(lava));
        const unsigned long long wa = wave_sum(static_cast<unsigned long long>(air));
        if ((threadIdx.x & 63u) == 0) {
            if (wl) atomicAdd(&D.ctrl->lava_total, wl);
            if (wa) atomicAdd(&D.ctrl->air_total, wa);
        }
        const int any_lava = __syncthreads_or(lava);
        const int any_air = __syncthreads_or(air);
        if (threadIdx.x == 0) D.tile_flags[t] = static_cast<uint8_t>((any_lava ? kFlagLava : 0) | (any_air ? kFlagAir : 0));
    }
}

// ------------------------------------------------------------------------------------------------
// CA step kernels
// ------------------------------------------------------------------------------------------------

__global__ void __launch_bounds__(kBlock) ca_front_kernel(CaParams P, CaDev D) {
    if (D.ctrl->stop) return;
    const uint32_t t = blockIdx.x * kBlock + threadIdx.x;
    bool act = false;
    if (t < P.ntiles) {
        const uint8_t f = D.tile_flags[t];
        if (P.dense) {
            act = true;
        } else if (!P.classic && (f & kFlagLava)) {
            act = true;  // lava cools (rule 2)
        } else if (f & kFlagAir) {
            if (P.classic && (f & kFlagLava)) {
                act = true;
            } else {
                const int tx = static_cast<int>(t % static_cast<uint32_t>(P.ntx));
                const uint32_t r = t / static_cast<uint32_t>(P.ntx);
                const int ty = static_cast<int>(r % static_cast<uint32_t>(P.nty));
                const int tz = static_cast<int>(r / static_cast<uint32_t>(P.nty));
                for (int dz = -1; dz <= 1 && !act; ++dz) {
                    int qz = tz + dz;
                    if (!wrap_axis(qz, P.ntz, P.per[2])) continue;
                    for (int dy = -1; dy <= 1 && !act; ++dy) {
                        int qy = ty + dy;
                        if (!wrap_axis(qy, P.nty, P.per[1])) continue;
                        for (int dx = -1; dx <= 1; ++dx) {
                            int qx = tx + dx;
                            if (!wrap_axis(qx, P.ntx, P.per[0])) continue;
                            const uint32_t q = static_cast<uint32_t>(qx) +
                                               static_cast<uint32_t>(P.ntx) *
                                                   (static_cast<uint32_t>(qy) + static_cast<uint32_t>(P.nty) * qz);
                            if (D.tile_flags[q] & kFlagLava) {
                                act = true;
                                break;
                            }
                        }
                    }
                }
            }
        }
    }
    const uint32_t slot = wave_append(act, &D.ctrl->nactive);
    if (act) D.active[slot] = t;
}

__global__ void __launch_bounds__(kBlock) ca_descent_kernel(CaParams P, CaDev D) {
    if (D.ctrl->stop || P.classic) return;
    const uint32_t nact = D.ctrl->nactive;
    const uint32_t plane = static_cast<uint32_t>(P.nx) * static_cast<uint32_t>(P.ny);
    for (uint32_t s = blockIdx.x; s < nact; s += gridDim.x) {
        const TileCoord o = tile_origin(P, D.active[s]);
#pragma unroll 1
        for (int k = 0; k < 4; ++k) {
            int x, y, z;
            tile_cell(P, o, k, x, y, z);
            if (x >= P.nx || y >= P.ny || z >= P.nz) continue;
            const uint32_t col = static_cast<uint32_t>(x) + static_cast<uint32_t>(P.nx) * static_cast<uint32_t>(y);
            const uint32_t i = col + plane * static_cast<uint32_t>(z);
            if (D.type[i] != kLava) continue;
            int d = -1;
            if (has_support(D, col, plane, z)) {
                d = 0;
                const uint32_t side = P.nb.side_mask;
#pragma unroll
                for (int q = 9; q < 18; ++q) {  // dz == 0 positions
                    if (q == 13 || !((side >> q) & 1u)) continue;
                    int qx = x + (q % 3 - 1), qy = y + ((q / 3) % 3 - 1);
                    if (!wrap_axis(qx, P.nx, P.per[0]) || !wrap_axis(qy, P.ny, P.per[1])) continue;
                    const uint32_t qc = static_cast<uint32_t>(qx) + static_cast<uint32_t>(P.nx) * static_cast<uint32_t>(qy);
                    if (D.type[qc + plane * static_cast<uint32_t>(z)] == kAir) {
                        const int dd = drop_depth(D, qc, plane, z);
                        d = dd > d ? dd : d;
                    }
                }
            }
            D.descent[i] = static_cast<int8_t>(d);
        }
    }
}

__global__ void __launch_bounds__(kBlock) ca_rules_kernel(CaParams P, CaDev D) {
    if (D.ctrl->stop) return;
    const uint32_t nact = D.ctrl->nactive;
    const int32_t t = D.ctrl->steps;
    const uint64_t prefix = hash_prefix(P.seed, static_cast<uint64_t>(t));
    const bool erupting = t < P.eruption_steps;
    const uint16_t stamp = static_cast<uint16_t>(t < 65535 ? t : 65535);
    const uint32_t plane = static_cast<uint32_t>(P.nx) * static_cast<uint32_t>(P.ny);
    unsigned long long lava_after = 0, born = 0, changed = 0;
    for (uint32_t s = blockIdx.x; s < nact; s += gridDim.x) {
        const uint32_t tile = D.active[s];
        const TileCoord o = tile_origin(P, tile);
        int tile_lava = 0, tile_air = 0;
#pragma unroll 1
        for (int k = 0; k < 4; ++k) {
            int x, y, z;
            tile_cell(P, o, k, x, y, z);
            bool want = false;
            Change c;
            if (x < P.nx && y < P.ny && z < P.nz) {
                const uint32_t i = static_cast<uint32_t>(x) + static_cast<uint32_t>(P.nx) * static_cast<uint32_t>(y) +
                                   plane * static_cast<uint32_t>(z);
                const uint8_t ty = D.type[i];
                if (ty == kLava) {
                    if (P.classic) {
                        tile_lava = 1;
                    } else {
                        // rule 2 (LavaCA.hpp:1129-1147, 973-994)
                        const int32_t f = D.flow[i];
                        float tc = D.temp[i];
                        const float before = tc;
                        if (erupting && f > 0 && f < P.nvent_slots && D.vent_cell[f] == i)
                            tc = P.erupt_temp;
                        else
                            tc = tc - P.cool_amb;
                        if (tc < P.solidus) {
                            want = true;
                            c = Change{i, f, tc, stamp, kBasalt, kKindSolidify};
                            ++changed;
                        } else {
                            ++lava_after;
                            tile_lava = 1;
                            if (tc != before) {
                                want = true;
                                c = Change{i, f, tc, 0, kLava, kKindTemp};
                                ++changed;
                            }
                        }
                    }
                } else if (ty == kAir) {
                    // rule 1 (LavaCA.hpp:1115-1128) / classic (LavaCA.hpp:1175-1190)
                    uint32_t mask = P.nb.mask[0];
                    uint64_t h_index = 0;
                    const bool random_nb = P.nb.variants > 1;
                    if (random_nb) {
                        h_index = splitmix64(prefix ^ i);
                        mask = P.nb.mask[splitmix64(h_index ^ kSaltNeighborhood) % static_cast<uint64_t>(P.nb.variants)];
                    }
                    Vote4 v;
                    if (P.classic)
                        gather_classic(P, D, x, y, z, mask, v);
                    else
                        gather_lava(P, D, x, y, z, mask, v);
                    bool has = v.n > 0;
                    int32_t wf = 0;
                    float wt = 0.0f;
                    if (has) {
                        if (!random_nb) h_index = splitmix64(prefix ^ i);
                        if (v.overflow)
                            has = wide_vote(P, D, x, y, z, mask, h_index, wf, wt);
                        else
                            v.winner(h_index, wf, wt);
                    }
                    if (has && (P.classic || wt >= P.solidus)) {
                        want = true;
                        c = Change{i, wf, P.classic ? 0.0f : wt, stamp, kLava, kKindBorn};
                        ++lava_after;
                        ++born;
                        ++changed;
                        tile_lava = 1;
                    } else {
                        tile_air = 1;
                    }
                }
            }
            const uint32_t slot = wave_append(want, &D.ctrl->nchanges);
            if (want) D.changes[slot] = c;
        }
        const int any_lava = __syncthreads_or(tile_lava);
        const int any_air = __syncthreads_or(tile_air);
        if (threadIdx.x == 0)
            D.tile_flags[tile] = static_cast<uint8_t>((any_lava ? kFlagLava : 0) | (any_air ? kFlagAir : 0));
    }
    lava_after = wave_sum(lava_after);
    born = wave_sum(born);
    changed = wave_sum(changed);
    if ((threadIdx.x & 63u) == 0) {
        if (lava_after) atomicAdd(&D.ctrl->lava_after, lava_after);
        if (born) atomicAdd(&D.ctrl->born, born);
        if (changed) atomicAdd(&D.ctrl->changed, changed);
    }
}

__global__ void __launch_bounds__(kBlock) ca_apply_kernel(CaDev D) {
    if (D.ctrl->stop) return;
    const uint32_t n = D.ctrl->nchanges;
    for (uint32_t j = blockIdx.x * kBlock + threadIdx.x; j < n; j += gridDim.x * kBlock) {
        const Change c = D.changes[j];
        D.type[c.idx] = c.type;
        D.temp[c.idx] = c.temp;
        if (c.kind == kKindBorn) {
            D.flow[c.idx] = c.flow;
            D.arrival[c.idx] = c.stamp;
        } else if (c.kind == kKindSolidify) {
            D.solid[c.idx] = c.stamp;
        }
    }
}

// finishStep (LavaCA.hpp:1206-1229) + frontCounts (LavaCA.hpp:865-871)
__global__ void ca_finish_kernel(CaParams P, CaDev D) {
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    CaCtrl* c = D.ctrl;
    if (c->stop) return;
    const int32_t t = c->steps;
    const unsigned long long active = P.classic ? c->lava_total + c->born : c->lava_after;
    const unsigned long long air = c->air_total - c->born;
    D.hist_active[t] = active;
    D.hist_changed[t] = c->changed;
    c->steps = t + 1;
    c->lava_total = active;
    c->air_total = air;
    int32_t stop = 0;
    if (!P.classic && active == 0)
        stop = 1;  // NoLava
    else if (P.classic && air == 0)
        stop = 2;  // NoAir
    else if (c->changed == 0 && P.nb.variants == 1)
        stop = 3;  // NoChange
    else if (t + 1 >= P.max_steps)
        stop = 4;  // MaxSteps
    c->stop = stop;
    c->tiles_visited += c->nactive;
    c->changes_total += c->nchanges;
    c->lava_after = 0;
    c->born = 0;
    c->changed = 0;
    c->nactive = 0;
    c->nchanges = 0;
}

// ------------------------------------------------------------------------------------------------
// MC kernels
// ------------------------------------------------------------------------------------------------

namespace {

template <typename S>
__device__ inline int32_t nb_spin(const McParams& P, const S* spin, int x, int y, int z, int k) {
    int qx = x + (k % 3 - 1), qy = y + ((k / 3) % 3 - 1), qz = z + (k / 9 - 1);
    if (!wrap_axis(qx, P.nx, P.per[0]) || !wrap_axis(qy, P.ny, P.per[1]) || !wrap_axis(qz, P.nz, P.per[2])) return 0;
    return static_cast<int32_t>(
        spin[static_cast<uint32_t>(qx) + static_cast<uint32_t>(P.nx) *
                                             (static_cast<uint32_t>(qy) + static_cast<uint32_t>(P.ny) * qz)]);
}

__device__ inline void cell_xyz(const McParams& P, uint32_t g, int& x, int& y, int& z) {
    x = static_cast<int>(g % static_cast<uint32_t>(P.nx));
    const uint32_t r = g / static_cast<uint32_t>(P.nx);
    y = static_cast<int>(r % static_cast<uint32_t>(P.ny));
    z = static_cast<int>(r / static_cast<uint32_t>(P.ny));
}

// One Metropolis trial (PottsMC.cpp:403-443): h = hashIndex(prefix, g) already computed, mob_thr is the
// exact 53-bit threshold ceil(m * 2^53) of the cell (skip iff (r >> 11) >= thr), accept is the table
// T[before][after] (kDownhill: dE <= 0, else accept iff (r >> 11) < T).
template <typename S>
__device__ inline void mc_trial(const McParams& P, S* spin, uint32_t g, uint64_t h, int var, const uint64_t* accept,
                                unsigned long long& att, unsigned long long& acc, unsigned long long& up) {
    ++att;
    int x, y, z;
    cell_xyz(P, g, x, y, z);
    const int32_t cur = static_cast<int32_t>(spin[g]);
    int32_t prop = 1 + static_cast<int32_t>(splitmix64(h ^ kSaltMcOrient) % static_cast<uint64_t>(P.Q - 1));
    if (prop >= cur) ++prop;
    const uint32_t mask = P.nb.mask[var];
    int in_dom = 0, same_a = 0, same_b = 0;
#pragma unroll
    for (int k = 0; k < 27; ++k) {
        if (k == 13 || !((mask >> k) & 1u)) continue;
        const int32_t sj = nb_spin(P, spin, x, y, z, k);
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

}  // namespace

template <typename S>
__global__ void __launch_bounds__(kBlock) mc_color_kernel(McParams P, S* spin, const uint32_t* list,
                                                          const uint64_t* mob_thr, uint32_t count, uint64_t prefix,
                                                          const uint64_t* accept, unsigned long long* counters) {
    unsigned long long att = 0, acc = 0, up = 0;
    const uint32_t j = blockIdx.x * kBlock + threadIdx.x;
    if (j < count) {
        const uint32_t g = list[j];
        const uint64_t h = splitmix64(prefix ^ g);
        if (mob_thr == nullptr || (splitmix64(h ^ kSaltMcMobility) >> 11) < mob_thr[j]) {
            const int var = P.nb.variants > 1
                                ? static_cast<int>(splitmix64(h ^ kSaltNeighborhood) % static_cast<uint64_t>(P.nb.variants))
                                : 0;
            mc_trial(P, spin, g, h, var, accept, att, acc, up);
        }
    }
    att = wave_sum(att);
    acc = wave_sum(acc);
    up = wave_sum(up);
    if ((threadIdx.x & 63u) == 0) {
        if (att) atomicAdd(&counters[0], att);
        if (acc) atomicAdd(&counters[1], acc);
        if (up) atomicAdd(&counters[2], up);
    }
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
                            ? static_cast<int>(splitmix64(splitmix64(prefix ^ g) ^ kSaltNeighborhood) %
                                               static_cast<uint64_t>(P.nb.variants))
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
    unsigned long long u = 0;
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
            mask = P.nb.mask[splitmix64(splitmix64(prefix ^ g) ^ kSaltNeighborhood) % static_cast<uint64_t>(P.nb.variants)];
        }
        int in_dom = 0, same = 0;
#pragma unroll
        for (int k = 0; k < 27; ++k) {
            if (k == 13 || !((mask >> k) & 1u)) continue;
            const int32_t sj = nb_spin(P, spin, x, y, z, k);
            in_dom += sj != 0;
            same += sj == s;
        }
        u = static_cast<unsigned long long>(in_dom - same);
    }
    u = wave_sum(u);
    if ((threadIdx.x & 63u) == 0 && u) atomicAdd(out, u);
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
#pragma unroll
    for (int k = 14; k < 27; ++k) {
        if (!((mask >> k) & 1u)) continue;
        int qx = x + (k % 3 - 1), qy = y + ((k / 3) % 3 - 1), qz = z + (k / 9 - 1);
        if (!wrap_axis(qx, P.nx, P.per[0]) || !wrap_axis(qy, P.ny, P.per[1]) || !wrap_axis(qz, P.nz, P.per[2]))
            continue;
        const uint32_t q = static_cast<uint32_t>(qx) +
                           static_cast<uint32_t>(P.nx) * (static_cast<uint32_t>(qy) + static_cast<uint32_t>(P.ny) * qz);
        if (spin[q] == s) uf_unite(parent, g, q);
    }
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
    high = wave_sum(high);
    if ((threadIdx.x & 63u) == 0 && high) atomicAdd(&out[8], high);
}

__global__ void __launch_bounds__(kBlock) mc_grain_roots_kernel(const uint32_t* list, uint32_t count,
                                                                 uint32_t* parent, const uint32_t* size,
                                                                 const uint32_t* hi, unsigned long long* out) {
    const uint32_t j = blockIdx.x * kBlock + threadIdx.x;
    if (j >= count) return;
    const uint32_t g = list[j];
    const uint32_t r = uf_find(parent, g);
    parent[g] = r;  // flatten for label download (same value any concurrent reader could see)
    if (r != g) return;
    const unsigned long long sz = size[g];
    atomicAdd(&out[0], 1ULL);
    atomicMax(&out[1], sz);
    if (hi != nullptr) {
        const int c = (2ULL * hi[g] >= sz) ? 0 : 1;  // McStage.cpp:145
        atomicAdd(&out[2 + 3 * c], 1ULL);
        atomicAdd(&out[3 + 3 * c], sz);
        atomicMax(&out[4 + 3 * c], sz);
    }
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
}  // namespace

void ca_launch_terrain(const CaParams& P, const CaDev& D, const int32_t* height, hipStream_t s) {
    const uint64_t n = static_cast<uint64_t>(P.nx) * P.ny * P.nz;
    hipLaunchKernelGGL(ca_terrain_kernel, dim3(grid_for(n, 16384)), dim3(kBlock), 0, s, P, D, height);
}

void ca_launch_seeds(const CaParams&, const CaDev& D, const uint32_t* cells, const int32_t* flows, const float* temps,
                     int n, hipStream_t s) {
    hipLaunchKernelGGL(ca_seeds_kernel, dim3(1), dim3(64), 0, s, D, cells, flows, temps, n);
}

void ca_launch_init_flags(const CaParams& P, const CaDev& D, hipStream_t s) {
    const unsigned g = P.ntiles < 16384 ? P.ntiles : 16384;
    hipLaunchKernelGGL(ca_init_flags_kernel, dim3(g ? g : 1), dim3(kBlock), 0, s, P, D);
}

void ca_launch_step(const CaParams& P, const CaDev& D, int grid_tiles, hipStream_t s, hipEvent_t* ev) {
    const unsigned gt = static_cast<unsigned>(grid_tiles);
    if (ev) hipEventRecord(ev[0], s);
    hipLaunchKernelGGL(ca_front_kernel, dim3(grid_for(P.ntiles)), dim3(kBlock), 0, s, P, D);
    if (ev) hipEventRecord(ev[1], s);
    if (!P.classic) hipLaunchKernelGGL(ca_descent_kernel, dim3(gt), dim3(kBlock), 0, s, P, D);
    if (ev) hipEventRecord(ev[2], s);
    hipLaunchKernelGGL(ca_rules_kernel, dim3(gt), dim3(kBlock), 0, s, P, D);
    if (ev) hipEventRecord(ev[3], s);
    hipLaunchKernelGGL(ca_apply_kernel, dim3(gt), dim3(kBlock), 0, s, D);
    if (ev) hipEventRecord(ev[4], s);
    hipLaunchKernelGGL(ca_finish_kernel, dim3(1), dim3(64), 0, s, P, D);
    if (ev) hipEventRecord(ev[5], s);
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
    VCA_SPIN_DISPATCH(bytes, hipLaunchKernelGGL(mc_color_kernel<S>, dim3(grid_for(count, 1u << 30)), dim3(kBlock), 0, s,
                                                P, static_cast<S*>(spin), list, mob_thr, count, prefix, accept,
                                                counters));
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

}  // namespace vca

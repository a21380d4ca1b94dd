/* CPU restatement of the reference engine's lava CA step and Potts/Metropolis sweep.
 *
 * TEST INFRASTRUCTURE ONLY (the checker, never the product).  Every function cites the
 * reference file:line it restates (paths relative to the reference root).  The code is
 * written for clarity, not speed: a single thread, unframed arrays, coordinates wrapped
 * per axis.  Floating-point work mirrors the reference exactly: float subtraction / max /
 * compare for temperatures, double for the Metropolis and mobility decisions; build with
 * -ffp-contract=off.
 *
 * Pinning (what fixes each part without running the reference): see oracle/README.md and
 * tests/test_oracle.py.
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

/* ======================================================================================
 * Hash — engine/src/core/Rng.hpp:22-74
 * ==================================================================================== */
static const uint64_t kSalts[9] = {
    0x4E45494748424F52ULL, /* neighbourhood variant   Rng.hpp:22 */
    0x5449455F42524B31ULL, /* flowId tie-break        Rng.hpp:23 */
    0x5445525241494E31ULL, /* terrain noise           Rng.hpp:24 */
    0x56454E5453454544ULL, /* vent placement          Rng.hpp:25 */
    0x4D43494E49544F52ULL, /* MC initial orientation  Rng.hpp:27 */
    0x4D435049434B4345ULL, /* MC random-mode pick     Rng.hpp:28 */
    0x4D434D4F42494C49ULL, /* MC mobility skip        Rng.hpp:29 */
    0x4D434F5249454E54ULL, /* MC proposed orientation Rng.hpp:30 */
    0x4D43414343455054ULL, /* MC Metropolis accept    Rng.hpp:31 */
};
enum { S_NB = 0, S_TIE, S_TERRAIN, S_VENTS, S_INIT, S_PICK, S_MOB, S_ORIENT, S_ACCEPT };

uint64_t orc_salt(int which) { return (which >= 0 && which < 9) ? kSalts[which] : 0; }

/* Rng.hpp:35-40 */
uint64_t orc_splitmix64(uint64_t x) {
    x += 0x9E3779B97F4A7C15ULL;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ULL;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBULL;
    return x ^ (x >> 31);
}
/* Rng.hpp:66-72 (prefix/index/salt split; composed == hashRand, Rng.hpp:46-52) */
uint64_t orc_hash_prefix(uint64_t seed, uint64_t step) {
    return orc_splitmix64(orc_splitmix64(seed ^ 0xA0761D6478BD642FULL) ^ step);
}
uint64_t orc_hash_index(uint64_t prefix, uint64_t index) { return orc_splitmix64(prefix ^ index); }
uint64_t orc_hash_salt(uint64_t h, uint64_t salt) { return orc_splitmix64(h ^ salt); }
uint64_t orc_hash_rand(uint64_t seed, uint64_t step, uint64_t index, uint64_t salt) {
    uint64_t h = orc_splitmix64(seed ^ 0xA0761D6478BD642FULL);
    h = orc_splitmix64(h ^ step);
    h = orc_splitmix64(h ^ index);
    return orc_splitmix64(h ^ salt);
}
/* Rng.hpp:74 — 53 top bits scaled to [0, 1) */
static double to_uniform(uint64_t r) { return (double)(r >> 11) * (1.0 / 9007199254740992.0); }

/* ======================================================================================
 * Neighbourhoods — engine/src/core/Neighborhood.cpp:18-119
 * ==================================================================================== */
static int iabs(int v) { return v < 0 ? -v : v; }

/* Filter of variant `v` of kind `kind` on plane offsets (u, v2) (2D) — Neighborhood.cpp:73-99 */
static int keep2d(int kind, int variant, int u, int v) {
    switch (kind) {
        case ORC_VON_NEUMANN: return iabs(u) + iabs(v) == 1;
        case ORC_MOORE: return 1;
        case ORC_PENTAGONAL:
            return variant == 0 ? u != -1 : variant == 1 ? u != 1 : variant == 2 ? v != -1 : v != 1;
        case ORC_HEXAGONAL:
            return variant == 0 ? !((u == -1 && v == -1) || (u == 1 && v == 1))
                                : !((u == -1 && v == 1) || (u == 1 && v == -1));
    }
    return 0;
}
/* Neighborhood.cpp:100-116 */
static int keep3d(int kind, int variant, int dx, int dy, int dz) {
    switch (kind) {
        case ORC_VON_NEUMANN: return iabs(dx) + iabs(dy) + iabs(dz) == 1;
        case ORC_MOORE: return 1;
        case ORC_PENTAGONAL: {
            const int a[6] = {dx != -1, dx != 1, dy != -1, dy != 1, dz != -1, dz != 1};
            return a[variant];
        }
    }
    return 0;
}

/* Neighborhood.cpp:55-119: variant lists in the reference order (2D: v outer, u inner over the
 * plane of the two axes of size > 1; 3D: dz, dy, dx). */
int orc_nb_build(int kind, int nx, int ny, int nz, orc_nb* out) {
    memset(out, 0, sizeof(*out));
    const int sizes[3] = {nx, ny, nz};
    const int axes = (nx > 1) + (ny > 1) + (nz > 1);
    if (axes < 2) return -1;
    if (axes == 2) {
        int plane[2], k = 0;
        for (int a = 0; a < 3; ++a)
            if (sizes[a] > 1) plane[k++] = a;
        out->count = kind == ORC_PENTAGONAL ? 4 : kind == ORC_HEXAGONAL ? 2 : 1;
        for (int var = 0; var < out->count; ++var) {
            int n = 0;
            for (int v = -1; v <= 1; ++v)
                for (int u = -1; u <= 1; ++u) {
                    if ((u == 0 && v == 0) || !keep2d(kind, var, u, v)) continue;
                    int d[3] = {0, 0, 0};
                    d[plane[0]] = u;
                    d[plane[1]] = v;
                    out->d[var][n][0] = (int8_t)d[0];
                    out->d[var][n][1] = (int8_t)d[1];
                    out->d[var][n][2] = (int8_t)d[2];
                    ++n;
                }
            out->size[var] = n;
        }
        return 0;
    }
    if (kind == ORC_HEXAGONAL) return -1; /* Neighborhood.cpp:111-116 */
    out->count = kind == ORC_PENTAGONAL ? 6 : 1;
    for (int var = 0; var < out->count; ++var) {
        int n = 0;
        for (int dz = -1; dz <= 1; ++dz)
            for (int dy = -1; dy <= 1; ++dy)
                for (int dx = -1; dx <= 1; ++dx) {
                    if ((dx == 0 && dy == 0 && dz == 0) || !keep3d(kind, var, dx, dy, dz)) continue;
                    out->d[var][n][0] = (int8_t)dx;
                    out->d[var][n][1] = (int8_t)dy;
                    out->d[var][n][2] = (int8_t)dz;
                    ++n;
                }
        out->size[var] = n;
    }
    return 0;
}

/* ======================================================================================
 * mt19937-64 (Matsumoto & Nishimura, "Tables of 64-bit Mersenne Twisters", 2000) — the
 * generator std::mt19937_64 specifies; the reference seeds it with splitmix64(seed ^ salt)
 * (engine/src/terrain/VentPlacer.cpp:37,75).
 * ==================================================================================== */
void orc_mt64_seed(orc_mt64* g, uint64_t seed) {
    g->mt[0] = seed;
    for (int i = 1; i < 312; ++i)
        g->mt[i] = 6364136223846793005ULL * (g->mt[i - 1] ^ (g->mt[i - 1] >> 62)) + (uint64_t)i;
    g->mti = 312;
}
uint64_t orc_mt64_next(orc_mt64* g) {
    static const uint64_t mag[2] = {0ULL, 0xB5026F5AA96619E9ULL};
    const uint64_t UM = 0xFFFFFFFF80000000ULL, LM = 0x7FFFFFFFULL;
    if (g->mti >= 312) {
        int i;
        uint64_t x;
        for (i = 0; i < 312 - 156; ++i) {
            x = (g->mt[i] & UM) | (g->mt[i + 1] & LM);
            g->mt[i] = g->mt[i + 156] ^ (x >> 1) ^ mag[x & 1ULL];
        }
        for (; i < 311; ++i) {
            x = (g->mt[i] & UM) | (g->mt[i + 1] & LM);
            g->mt[i] = g->mt[i + (156 - 312)] ^ (x >> 1) ^ mag[x & 1ULL];
        }
        x = (g->mt[311] & UM) | (g->mt[0] & LM);
        g->mt[311] = g->mt[155] ^ (x >> 1) ^ mag[x & 1ULL];
        g->mti = 0;
    }
    uint64_t x = g->mt[g->mti++];
    x ^= (x >> 29) & 0x5555555555555555ULL;
    x ^= (x << 17) & 0x71D67FFFEDA60000ULL;
    x ^= (x << 37) & 0xFFF7EEE000000000ULL;
    x ^= (x >> 43);
    return x;
}

/* ======================================================================================
 * Spill levels by priority-flood — engine/src/terrain/TerrainGenerator.cpp:86-160
 * (Barnes, Lehman & Mulla 2014).  Binary min-heap on (level, column index).
 * ==================================================================================== */
typedef struct { int32_t level; int64_t idx; } heap_item;
static int item_less(heap_item a, heap_item b) { return a.level < b.level || (a.level == b.level && a.idx < b.idx); }
static void heap_push(heap_item* h, int64_t* n, heap_item it) {
    int64_t i = (*n)++;
    h[i] = it;
    while (i > 0) {
        int64_t p = (i - 1) / 2;
        if (!item_less(h[i], h[p])) break;
        heap_item t = h[i]; h[i] = h[p]; h[p] = t;
        i = p;
    }
}
static heap_item heap_pop(heap_item* h, int64_t* n) {
    heap_item top = h[0];
    h[0] = h[--(*n)];
    int64_t i = 0;
    for (;;) {
        int64_t l = 2 * i + 1, r = l + 1, m = i;
        if (l < *n && item_less(h[l], h[m])) m = l;
        if (r < *n && item_less(h[r], h[m])) m = r;
        if (m == i) break;
        heap_item t = h[i]; h[i] = h[m]; h[m] = t;
        i = m;
    }
    return top;
}

void orc_fill_depressions(const int32_t* hm, int nx, int ny, int px, int py, int32_t* out) {
    const int64_t n = (int64_t)nx * ny;
    uint8_t* closed = (uint8_t*)calloc((size_t)n, 1);
    heap_item* heap = (heap_item*)malloc(sizeof(heap_item) * (size_t)(n + 1));
    int64_t hn = 0;
    memcpy(out, hm, sizeof(int32_t) * (size_t)n);
    /* outlets: columns on a non-periodic edge (TerrainGenerator.cpp:104-113) */
    for (int y = 0; y < ny; ++y)
        for (int x = 0; x < nx; ++x) {
            const int ex = !px && nx > 1 && (x == 0 || x == nx - 1);
            const int ey = !py && ny > 1 && (y == 0 || y == ny - 1);
            if (ex || ey) {
                int64_t i = x + (int64_t)nx * y;
                if (!closed[i]) { closed[i] = 1; heap_push(heap, &hn, (heap_item){hm[i], i}); }
            }
        }
    if (hn == 0 && n > 0) { /* fully periodic: the lowest columns (TerrainGenerator.cpp:114-124) */
        int32_t lo = hm[0];
        for (int64_t i = 1; i < n; ++i) if (hm[i] < lo) lo = hm[i];
        for (int64_t i = 0; i < n; ++i)
            if (hm[i] == lo && !closed[i]) { closed[i] = 1; heap_push(heap, &hn, (heap_item){hm[i], i}); }
    }
    static const int dx[4] = {-1, 1, 0, 0}, dy[4] = {0, 0, -1, 1};
    while (hn > 0) { /* TerrainGenerator.cpp:127-157 */
        heap_item it = heap_pop(heap, &hn);
        const int x = (int)(it.idx % nx), y = (int)(it.idx / nx);
        for (int k = 0; k < 4; ++k) {
            int sx = x + dx[k], sy = y + dy[k];
            if (sx < 0 || sx >= nx) { if (!px || nx == 1) continue; sx = (sx + nx) % nx; }
            if (sy < 0 || sy >= ny) { if (!py || ny == 1) continue; sy = (sy + ny) % ny; }
            const int64_t j = sx + (int64_t)nx * sy;
            if (closed[j]) continue;
            closed[j] = 1;
            out[j] = hm[j] > it.level ? hm[j] : it.level;
            heap_push(heap, &hn, (heap_item){out[j], j});
        }
    }
    free(heap);
    free(closed);
}

/* ======================================================================================
 * Lava / classic CA — engine/src/ca/LavaCA.hpp, engine/src/ca/CaRules.hpp
 * ==================================================================================== */
#define MAX_DROP 16 /* LavaCA.hpp:1293 */

struct orc_ca {
    orc_ca_params p;
    int periodic[3];
    orc_nb nb;
    int nside;
    int8_t side[26][3];   /* dz == 0 offsets, union over variants (LavaCA.hpp:102-112) */
    int diagonal_falls;   /* any variant has dz == -1 with dx or dy != 0 */
    int64_t n;
    uint8_t* type;
    int32_t* flow;
    float* temp;
    uint16_t* solid;
    uint16_t* arrival;
    int8_t* descent;
    int32_t* spill;       /* NULL in classic mode */
    int64_t* vent_cell;   /* indexed by flowId; -1 = none (LavaCA.hpp:215-227, 335-341) */
    int nvent_slots;
    int steps;
    int stop;
    uint64_t* hist_active;
    uint64_t* hist_changed;
    /* change list of one step */
    int64_t* ch_idx;
    uint8_t* ch_type;
    int32_t* ch_flow;
    float* ch_temp;
    uint16_t* ch_solid;
    uint16_t* ch_arrival;
};

static int64_t gidx(const orc_ca* s, int x, int y, int z) {
    return x + (int64_t)s->p.nx * (y + (int64_t)s->p.ny * z);
}

/* Neighbour coordinate along one axis: wrap if periodic, -1 (VOID) if outside (Boundary.hpp:66-131). */
static int axis_coord(int v, int n, int periodic) {
    if (v >= 0 && v < n) return v;
    if (!periodic) return -1;
    return v < 0 ? v + n : v - n;
}
/* Type of (x, y, z) with boundary handling; fills the wrapped index. */
static uint8_t type_at(const orc_ca* s, int x, int y, int z, int64_t* idx) {
    const int qx = axis_coord(x, s->p.nx, s->periodic[0]);
    const int qy = axis_coord(y, s->p.ny, s->periodic[1]);
    const int qz = axis_coord(z, s->p.nz, s->periodic[2]);
    if (qx < 0 || qy < 0 || qz < 0) return ORC_VOID;
    const int64_t i = gidx(s, qx, qy, qz);
    if (idx) *idx = i;
    return s->type[i];
}

static int is_solid_ground(uint8_t t) { return t == ORC_ROCK || t == ORC_BASALT; } /* CaRules.hpp:122 */

/* LavaCA.hpp:320-323: spill level with wrapped column coordinates */
static int spill_at(const orc_ca* s, int x, int y) {
    x = axis_coord(x, s->p.nx, 1);
    y = axis_coord(y, s->p.ny, 1);
    return s->spill[x + s->p.nx * y];
}

/* LavaCA.hpp:390-399.  (x, y) are interior (already wrapped) column coordinates. */
static int has_support(const orc_ca* s, int x, int y, int z) {
    if (z == 0) return 1;
    const uint8_t below = s->type[gidx(s, x, y, z - 1)];
    if (is_solid_ground(below)) return 1;
    return below == ORC_LAVA && s->spill != NULL && z - 1 < spill_at(s, x, y);
}

/* LavaCA.hpp:404-414 */
static int drop_depth(const orc_ca* s, int x, int y, int z) {
    for (int d = 0; d <= MAX_DROP && z - d >= 0; ++d) {
        if (d > 0 && s->type[gidx(s, x, y, z - d)] != ORC_AIR) return -1;
        if (has_support(s, x, y, z - d)) return d;
    }
    return MAX_DROP;
}

/* LavaCA.hpp:421-429: D(n) of a supported lava cell; -1 if unsupported (LavaCA.hpp:551-566). */
static int descent_of(const orc_ca* s, int x, int y, int z) {
    if (!has_support(s, x, y, z)) return -1;
    int best = 0;
    for (int k = 0; k < s->nside; ++k) {
        const int tx = axis_coord(x + s->side[k][0], s->p.nx, s->periodic[0]);
        const int ty = axis_coord(y + s->side[k][1], s->p.ny, s->periodic[1]);
        if (tx < 0 || ty < 0) continue; /* VOID frame */
        if (s->type[gidx(s, tx, ty, z)] == ORC_AIR) {
            const int d = drop_depth(s, tx, ty, z);
            if (d > best) best = d;
        }
    }
    return best;
}

/* CandidateVote — CaRules.hpp:127-183 */
typedef struct { int n; int32_t ids[32]; int counts[32]; float maxt[32]; } vote_t;
static void vote_add(vote_t* v, int32_t id, float t) {
    for (int i = 0; i < v->n; ++i)
        if (v->ids[i] == id) {
            ++v->counts[i];
            if (v->maxt[i] < t) v->maxt[i] = t; /* std::max(a, b) = a < b ? b : a */
            return;
        }
    v->ids[v->n] = id;
    v->counts[v->n] = 1;
    v->maxt[v->n] = t;
    ++v->n;
}
static int vote_winner(const vote_t* v, uint64_t seed, uint64_t step, uint64_t cell) {
    if (v->n == 1) return 0;
    int best = 0;
    for (int i = 0; i < v->n; ++i) if (v->counts[i] > best) best = v->counts[i];
    int tied[32], k = 0;
    for (int i = 0; i < v->n; ++i) if (v->counts[i] == best) tied[k++] = i;
    if (k == 1) return tied[0];
    for (int a = 1; a < k; ++a) { /* ascending flowId (insertion sort; ids are distinct) */
        int t = tied[a], b = a - 1;
        while (b >= 0 && v->ids[tied[b]] > v->ids[t]) { tied[b + 1] = tied[b]; --b; }
        tied[b + 1] = t;
    }
    const uint64_t r = orc_hash_rand(seed, step, cell, kSalts[S_TIE]);
    return tied[(int)(r % (uint64_t)k)];
}

/* LavaCA.hpp:443-461 */
static void try_ponding(const orc_ca* s, int x, int y, int z, vote_t* vote) {
    if (s->spill == NULL || z > s->spill[x + s->p.nx * y]) return;
    const int bz = z - 1;
    if (bz < 0) return; /* frame below z = 0 is VOID */
    const int64_t b = gidx(s, x, y, bz);
    if (s->type[b] != ORC_LAVA) return;
    if (!has_support(s, x, y, bz)) return;
    const int hx[4] = {-1, 1, 0, 0}, hy[4] = {0, 0, -1, 1};
    const int nh = s->p.ny > 1 ? 4 : 2; /* LavaCA.hpp:93-98 */
    for (int k = 0; k < nh; ++k) {
        const uint8_t t = type_at(s, x + hx[k], y + hy[k], bz, NULL);
        if (t == ORC_AIR || t == ORC_VOID) return;
    }
    vote_add(vote, s->flow[b], s->temp[b] - s->p.cooling_lateral);
}

/* LavaCA.hpp:465-519: candidates for the AIR cell (x, y, z), global index idx. */
static void gather_lava(const orc_ca* s, int x, int y, int z, uint64_t prefix, int64_t idx, vote_t* vote) {
    vote->n = 0;
    int var = 0;
    if (s->nb.count > 1) var = (int)(orc_hash_salt(orc_hash_index(prefix, (uint64_t)idx), kSalts[S_NB]) % (uint64_t)s->nb.count);
    int cdrop = -100; /* lazily computed (LavaCA.hpp:469) */
    for (int k = 0; k < s->nb.size[var]; ++k) {
        const int dx = s->nb.d[var][k][0], dy = s->nb.d[var][k][1], dz = s->nb.d[var][k][2];
        if (dz < 0) {
            if (s->p.ponding && dx == 0 && dy == 0) try_ponding(s, x, y, z, vote);
            continue;
        }
        int64_t ni;
        if (type_at(s, x + dx, y + dy, z + dz, &ni) != ORC_LAVA) continue;
        float cooling;
        if (dz > 0 && dx == 0 && dy == 0) {
            cooling = s->p.cooling_fall;
        } else {
            const int d = s->descent[ni];
            if (d < 0) continue;
            if (cdrop == -100) cdrop = drop_depth(s, x, y, z);
            if (dz > 0) {
                if (cdrop != 0 || d != 1) continue;
                cooling = s->p.cooling_fall;
            } else {
                if (cdrop != d || (d == 1 && s->diagonal_falls)) continue;
                cooling = s->p.cooling_lateral;
            }
        }
        vote_add(vote, s->flow[ni], s->temp[ni] - cooling);
    }
}

static void add_unique(int8_t (*list)[3], int* n, const int8_t* o) {
    for (int i = 0; i < *n; ++i)
        if (list[i][0] == o[0] && list[i][1] == o[1] && list[i][2] == o[2]) return;
    memcpy(list[*n], o, 3);
    ++*n;
}

orc_ca* orc_ca_create(const orc_ca_params* p, const int32_t* height, const int32_t* spill,
                      const int32_t* vents, int nvents) {
    orc_ca* s = (orc_ca*)calloc(1, sizeof(orc_ca));
    s->p = *p;
    if (orc_nb_build(p->neighborhood, p->nx, p->ny, p->nz, &s->nb) != 0) { free(s); return NULL; }
    /* Boundary.cpp:5-11: z periodic only in classic mode; axes of size 1 never wrap */
    s->periodic[0] = p->periodic && p->nx > 1;
    s->periodic[1] = p->periodic && p->ny > 1;
    s->periodic[2] = p->periodic && p->classic && p->nz > 1;
    for (int v = 0; v < s->nb.count; ++v)
        for (int k = 0; k < s->nb.size[v]; ++k) {
            const int8_t* o = s->nb.d[v][k];
            if (o[2] == 0) add_unique(s->side, &s->nside, o);
            else if (o[2] == -1 && (o[0] != 0 || o[1] != 0)) s->diagonal_falls = 1;
        }
    s->n = (int64_t)p->nx * p->ny * p->nz;
    const size_t n = (size_t)s->n;
    s->type = (uint8_t*)calloc(n, 1);
    s->flow = (int32_t*)calloc(n, 4);
    s->temp = (float*)calloc(n, 4);
    s->solid = (uint16_t*)calloc(n, 2);
    s->arrival = (uint16_t*)calloc(n, 2);
    s->descent = (int8_t*)malloc(n);
    memset(s->descent, -1, n);
    s->ch_idx = (int64_t*)malloc(n * 8);
    s->ch_type = (uint8_t*)malloc(n);
    s->ch_flow = (int32_t*)malloc(n * 4);
    s->ch_temp = (float*)malloc(n * 4);
    s->ch_solid = (uint16_t*)malloc(n * 2);
    s->ch_arrival = (uint16_t*)malloc(n * 2);
    const int hcap = p->max_steps > 0 ? p->max_steps : 1;
    s->hist_active = (uint64_t*)calloc((size_t)hcap, 8);
    s->hist_changed = (uint64_t*)calloc((size_t)hcap, 8);
    /* terrain: ROCK for z < min(h, nz) (LavaCA.hpp:185-197) */
    if (!p->classic && height != NULL) {
        for (int y = 0; y < p->ny; ++y)
            for (int x = 0; x < p->nx; ++x) {
                int h = height[x + p->nx * y];
                if (h > p->nz) h = p->nz;
                for (int z = 0; z < h; ++z) s->type[gidx(s, x, y, z)] = ORC_ROCK;
            }
        if (spill != NULL) {
            s->spill = (int32_t*)malloc(sizeof(int32_t) * (size_t)p->nx * p->ny);
            memcpy(s->spill, spill, sizeof(int32_t) * (size_t)p->nx * p->ny);
        }
    }
    /* seeds (LavaCA.hpp:215-227) */
    int maxf = 0;
    for (int k = 0; k < nvents; ++k) if (vents[4 * k + 3] > maxf) maxf = vents[4 * k + 3];
    s->nvent_slots = maxf + 1;
    s->vent_cell = (int64_t*)malloc(sizeof(int64_t) * (size_t)s->nvent_slots);
    for (int f = 0; f < s->nvent_slots; ++f) s->vent_cell[f] = -1;
    const float t0 = p->classic ? 0.0f : p->erupt_temp;
    for (int k = 0; k < nvents; ++k) {
        const int64_t i = gidx(s, vents[4 * k], vents[4 * k + 1], vents[4 * k + 2]);
        s->type[i] = ORC_LAVA;
        s->flow[i] = vents[4 * k + 3];
        s->temp[i] = t0;
        s->solid[i] = 0;
        s->arrival[i] = 0;
        if (vents[4 * k + 3] >= 0) s->vent_cell[vents[4 * k + 3]] = i;
    }
    /* countInitialState (LavaCA.hpp:1252-1280) */
    uint64_t lava = 0, air = 0;
    for (int64_t i = 0; i < s->n; ++i) { lava += s->type[i] == ORC_LAVA; air += s->type[i] == ORC_AIR; }
    s->stop = ORC_RUNNING;
    if (!p->classic && lava == 0) s->stop = ORC_NO_LAVA;
    else if (p->classic && air == 0) s->stop = ORC_NO_AIR;
    else if (p->max_steps <= 0) s->stop = ORC_MAX_STEPS;
    return s;
}

void orc_ca_destroy(orc_ca* s) {
    if (!s) return;
    free(s->type); free(s->flow); free(s->temp); free(s->solid); free(s->arrival); free(s->descent);
    free(s->spill); free(s->vent_cell); free(s->hist_active); free(s->hist_changed);
    free(s->ch_idx); free(s->ch_type); free(s->ch_flow); free(s->ch_temp); free(s->ch_solid); free(s->ch_arrival);
    free(s);
}

/* LavaCA.hpp:335-341 */
static int is_vent(const orc_ca* s, int32_t flow, int64_t i) {
    return flow > 0 && flow < s->nvent_slots && s->vent_cell[flow] == i;
}

/* One synchronous step: the reference's full sweep (LavaCA.hpp:577-588 computeDescent +
 * updateLava :1109-1162, or updateClassic :1165-1202), then finishStep (:1206-1229). The new
 * states are collected in a change list and applied after all cells were evaluated. */
int orc_ca_step(orc_ca* s) {
    if (s->stop != ORC_RUNNING) return 0;
    const orc_ca_params* p = &s->p;
    const uint64_t t = (uint64_t)s->steps;
    const int erupting = s->steps < p->eruption_steps;
    const uint16_t stamp = (uint16_t)(s->steps < 65535 ? s->steps : 65535);
    const uint64_t prefix = orc_hash_prefix(p->seed, t);
    uint64_t lava = 0, air = 0, changed = 0;
    int64_t nch = 0;
    vote_t vote;
    if (!p->classic) {
        for (int z = 0; z < p->nz; ++z)
            for (int y = 0; y < p->ny; ++y)
                for (int x = 0; x < p->nx; ++x) {
                    const int64_t i = gidx(s, x, y, z);
                    s->descent[i] = (int8_t)(s->type[i] == ORC_LAVA ? descent_of(s, x, y, z) : -1);
                }
    }
    for (int z = 0; z < p->nz; ++z)
        for (int y = 0; y < p->ny; ++y)
            for (int x = 0; x < p->nx; ++x) {
                const int64_t i = gidx(s, x, y, z);
                const uint8_t ty = s->type[i];
                if (ty == ORC_AIR) {
                    if (p->classic) {
                        vote.n = 0;
                        int var = 0;
                        if (s->nb.count > 1) var = (int)(orc_hash_salt(orc_hash_index(prefix, (uint64_t)i), kSalts[S_NB]) % (uint64_t)s->nb.count);
                        for (int k = 0; k < s->nb.size[var]; ++k) {
                            int64_t ni;
                            if (type_at(s, x + s->nb.d[var][k][0], y + s->nb.d[var][k][1], z + s->nb.d[var][k][2], &ni) == ORC_LAVA)
                                vote_add(&vote, s->flow[ni], 0.0f);
                        }
                        if (vote.n > 0) {
                            const int w = vote_winner(&vote, p->seed, t, (uint64_t)i);
                            s->ch_idx[nch] = i; s->ch_type[nch] = ORC_LAVA; s->ch_flow[nch] = vote.ids[w];
                            s->ch_temp[nch] = 0.0f; s->ch_solid[nch] = 0; s->ch_arrival[nch] = stamp; ++nch;
                            ++lava; ++changed;
                        } else {
                            ++air;
                        }
                        continue;
                    }
                    gather_lava(s, x, y, z, prefix, i, &vote);
                    if (vote.n > 0) {
                        const int w = vote_winner(&vote, p->seed, t, (uint64_t)i);
                        const float tw = vote.maxt[w];
                        if (tw >= p->solidus_temp) {
                            s->ch_idx[nch] = i; s->ch_type[nch] = ORC_LAVA; s->ch_flow[nch] = vote.ids[w];
                            s->ch_temp[nch] = tw; s->ch_solid[nch] = 0; s->ch_arrival[nch] = stamp; ++nch;
                            ++lava; ++changed;
                            continue;
                        }
                    }
                    ++air;
                } else if (ty == ORC_LAVA) {
                    if (p->classic) { ++lava; continue; }
                    float tc = s->temp[i];
                    const float before = tc;
                    if (erupting && is_vent(s, s->flow[i], i)) tc = p->erupt_temp;
                    else tc -= p->cooling_ambient;
                    if (tc < p->solidus_temp) {
                        s->ch_idx[nch] = i; s->ch_type[nch] = ORC_BASALT; s->ch_flow[nch] = s->flow[i];
                        s->ch_temp[nch] = tc; s->ch_solid[nch] = stamp; s->ch_arrival[nch] = s->arrival[i]; ++nch;
                        ++changed;
                    } else {
                        ++lava;
                        if (tc != before) {
                            s->ch_idx[nch] = i; s->ch_type[nch] = ORC_LAVA; s->ch_flow[nch] = s->flow[i];
                            s->ch_temp[nch] = tc; s->ch_solid[nch] = s->solid[i]; s->ch_arrival[nch] = s->arrival[i]; ++nch;
                            ++changed;
                        }
                    }
                }
            }
    for (int64_t k = 0; k < nch; ++k) {
        const int64_t i = s->ch_idx[k];
        s->type[i] = s->ch_type[k]; s->flow[i] = s->ch_flow[k]; s->temp[i] = s->ch_temp[k];
        s->solid[i] = s->ch_solid[k]; s->arrival[i] = s->ch_arrival[k];
    }
    /* finishStep (LavaCA.hpp:1206-1229) */
    s->hist_active[s->steps] = lava;
    s->hist_changed[s->steps] = changed;
    ++s->steps;
    if (!p->classic && lava == 0) s->stop = ORC_NO_LAVA;
    else if (p->classic && air == 0) s->stop = ORC_NO_AIR;
    else if (changed == 0 && s->nb.count == 1) s->stop = ORC_NO_CHANGE;
    else if (s->steps >= p->max_steps) s->stop = ORC_MAX_STEPS;
    return s->stop == ORC_RUNNING;
}

int orc_ca_run(orc_ca* s) {
    while (orc_ca_step(s)) {
    }
    return s->steps;
}
int orc_ca_steps(const orc_ca* s) { return s->steps; }
int orc_ca_stop_reason(const orc_ca* s) { return s->stop; }

void orc_ca_get(const orc_ca* s, uint8_t* type, int32_t* flow, float* temp, uint16_t* solid, uint16_t* arrival) {
    const size_t n = (size_t)s->n;
    if (type) memcpy(type, s->type, n);
    if (flow) memcpy(flow, s->flow, n * 4);
    if (temp) memcpy(temp, s->temp, n * 4);
    if (solid) memcpy(solid, s->solid, n * 2);
    if (arrival) memcpy(arrival, s->arrival, n * 2);
}
int orc_ca_histories(const orc_ca* s, uint64_t* active, uint64_t* changed, int cap) {
    const int m = s->steps < cap ? s->steps : cap;
    if (active) memcpy(active, s->hist_active, sizeof(uint64_t) * (size_t)m);
    if (changed) memcpy(changed, s->hist_changed, sizeof(uint64_t) * (size_t)m);
    return s->steps;
}

/* ======================================================================================
 * Potts MC — engine/src/mc/PottsMC.cpp
 * ==================================================================================== */
struct orc_mc {
    orc_mc_params p;
    orc_nb nb;
    int64_t n;
    int32_t* spin;
    uint8_t* dom;
    float* mob;            /* NULL = mobility off */
    int64_t* list;         /* domain cells: index order (random mode) or colour-major (checkerboard) */
    int64_t nlist;
    int colors;
    int64_t color_start[28];
    int sweeps;
    uint64_t attempts, accepted, uphill;
};

static int64_t mc_gidx(const orc_mc* m, int x, int y, int z) {
    return x + (int64_t)m->p.nx * (y + (int64_t)m->p.ny * z);
}
/* Spin of the neighbour (x+dx, y+dy, z+dz): 0 outside the domain or beyond an absorbing edge
 * (PottsMC.cpp:226-277 frame; unlike-count PottsMC.hpp:174-204). */
static int32_t mc_nb_spin(const orc_mc* m, int x, int y, int z, const int8_t* o) {
    const int qx = axis_coord(x + o[0], m->p.nx, m->p.periodic[0]);
    const int qy = axis_coord(y + o[1], m->p.ny, m->p.periodic[1]);
    const int qz = axis_coord(z + o[2], m->p.nz, m->p.periodic[2]);
    if (qx < 0 || qy < 0 || qz < 0) return 0;
    return m->spin[mc_gidx(m, qx, qy, qz)];
}

/* PottsMC.cpp:145-151 */
static int axis_class(const orc_mc* m, int c, int a) {
    const int n = a == 0 ? m->p.nx : a == 1 ? m->p.ny : m->p.nz;
    if (m->p.periodic[a] && (n % 2) == 1 && c == n - 1) return 2;
    return c & 1;
}

orc_mc* orc_mc_create(const orc_mc_params* p, const uint8_t* domain, const float* mobility) {
    if (p->Q < 2) return NULL;
    orc_mc* m = (orc_mc*)calloc(1, sizeof(orc_mc));
    m->p = *p;
    if (orc_nb_build(p->neighborhood, p->nx, p->ny, p->nz, &m->nb) != 0) { free(m); return NULL; }
    m->n = (int64_t)p->nx * p->ny * p->nz;
    const size_t n = (size_t)m->n;
    m->spin = (int32_t*)calloc(n, 4);
    m->dom = (uint8_t*)malloc(n);
    for (size_t i = 0; i < n; ++i) m->dom[i] = domain ? (domain[i] != 0) : 1;
    if (mobility) { m->mob = (float*)malloc(n * 4); memcpy(m->mob, mobility, n * 4); }
    m->list = (int64_t*)malloc(sizeof(int64_t) * (n ? n : 1));
    /* initialize (PottsMC.cpp:104-143) */
    for (int64_t g = 0; g < m->n; ++g)
        if (m->dom[g]) {
            m->spin[g] = 1 + (int32_t)(orc_hash_rand(p->seed, 0, (uint64_t)g, kSalts[S_INIT]) % (uint64_t)p->Q);
            m->list[m->nlist++] = g;
        }
    if (p->checkerboard) { /* buildColors (PottsMC.cpp:160-224) */
        const int sz[3] = {p->nx, p->ny, p->nz};
        int odd_periodic = 0;
        for (int a = 0; a < 3; ++a) odd_periodic |= p->periodic[a] && sz[a] % 2 == 1;
        const int parity = p->neighborhood == ORC_VON_NEUMANN && !odd_periodic;
        int color_index[27];
        if (parity) {
            m->colors = 2;
        } else {
            int classes[3];
            for (int a = 0; a < 3; ++a) classes[a] = sz[a] == 1 ? 1 : (p->periodic[a] && sz[a] % 2 == 1 ? 3 : 2);
            m->colors = 0;
            for (int cz = 0; cz < classes[2]; ++cz)
                for (int cy = 0; cy < classes[1]; ++cy)
                    for (int cx = 0; cx < classes[0]; ++cx) color_index[cx + 3 * cy + 9 * cz] = m->colors++;
        }
        int64_t k = 0;
        for (int c = 0; c < m->colors; ++c) {
            m->color_start[c] = k;
            for (int z = 0; z < p->nz; ++z)
                for (int y = 0; y < p->ny; ++y)
                    for (int x = 0; x < p->nx; ++x) {
                        const int64_t g = mc_gidx(m, x, y, z);
                        if (!m->dom[g]) continue;
                        const int col = parity ? ((x + y + z) & 1)
                                               : color_index[axis_class(m, x, 0) + 3 * axis_class(m, y, 1) + 9 * axis_class(m, z, 2)];
                        if (col == c) m->list[k++] = g;
                    }
        }
        m->color_start[m->colors] = k;
    }
    return m;
}

void orc_mc_destroy(orc_mc* m) {
    if (!m) return;
    free(m->spin); free(m->dom); free(m->mob); free(m->list);
    free(m);
}

/* One trial on cell g with index hash h (PottsMC.cpp:403-443 checkerboard, :330-387 random). */
static void mc_trial(orc_mc* m, int64_t g, uint64_t h, int variant) {
    if (m->mob && to_uniform(orc_hash_salt(h, kSalts[S_MOB])) >= (double)m->mob[g]) return;
    ++m->attempts;
    const int x = (int)(g % m->p.nx);
    const int y = (int)((g / m->p.nx) % m->p.ny);
    const int z = (int)(g / ((int64_t)m->p.nx * m->p.ny));
    const int32_t current = m->spin[g];
    int32_t proposed = 1 + (int32_t)(orc_hash_salt(h, kSalts[S_ORIENT]) % (uint64_t)(m->p.Q - 1));
    if (proposed >= current) ++proposed;
    int in_domain = 0, same_a = 0, same_b = 0;
    for (int k = 0; k < m->nb.size[variant]; ++k) {
        const int32_t sj = mc_nb_spin(m, x, y, z, m->nb.d[variant][k]);
        in_domain += sj != 0;
        same_a += sj == current;
        same_b += sj == proposed;
    }
    const int before = in_domain - same_a, after = in_domain - same_b;
    const double delta = m->p.J * after - m->p.J * before;
    int accept = delta <= 0.0;
    if (!accept && m->p.kT > 0.0) {
        accept = to_uniform(orc_hash_salt(h, kSalts[S_ACCEPT])) < exp(-delta / m->p.kT);
        m->uphill += accept;
    }
    if (accept) {
        m->spin[g] = proposed;
        ++m->accepted;
    }
}

uint64_t orc_mc_sweep(orc_mc* m) {
    const uint64_t before = m->accepted;
    const uint64_t prefix = orc_hash_prefix(m->p.seed, (uint64_t)m->sweeps);
    const int nv = m->nb.count;
    if (m->p.checkerboard) { /* sweepColors (PottsMC.cpp:450-514) */
        for (int c = 0; c < m->colors; ++c)
            for (int64_t j = m->color_start[c]; j < m->color_start[c + 1]; ++j) {
                const int64_t g = m->list[j];
                const uint64_t h = orc_hash_index(prefix, (uint64_t)g);
                const int var = nv > 1 ? (int)(orc_hash_salt(h, kSalts[S_NB]) % (uint64_t)nv) : 0;
                mc_trial(m, g, h, var);
            }
    } else { /* sweepTrials (PottsMC.cpp:312-391): n picks with replacement */
        const uint64_t n = (uint64_t)m->nlist;
        for (uint64_t a = 0; a < n; ++a) {
            const uint64_t h = orc_hash_index(prefix, a);
            const int64_t g = m->list[orc_hash_salt(h, kSalts[S_PICK]) % n];
            const int var = nv > 1 ? (int)(orc_hash_salt(orc_hash_index(prefix, (uint64_t)g), kSalts[S_NB]) % (uint64_t)nv) : 0;
            mc_trial(m, g, h, var);
        }
    }
    ++m->sweeps;
    return m->accepted - before;
}

void orc_mc_counters(const orc_mc* m, uint64_t* attempts, uint64_t* accepted, uint64_t* uphill) {
    if (attempts) *attempts = m->attempts;
    if (accepted) *accepted = m->accepted;
    if (uphill) *uphill = m->uphill;
}
int orc_mc_colors(const orc_mc* m) { return m->colors; }
uint64_t orc_mc_domain_cells(const orc_mc* m) { return (uint64_t)m->nlist; }
void orc_mc_spins(const orc_mc* m, int32_t* out) { memcpy(out, m->spin, sizeof(int32_t) * (size_t)m->n); }

/* totalEnergy (PottsMC.cpp:590-664) */
double orc_mc_total_energy(const orc_mc* m) {
    uint64_t unlike = 0;
    if (m->nb.count == 1) { /* forward pairs (positive linear offset) once each */
        for (int64_t j = 0; j < m->nlist; ++j) {
            const int64_t g = m->list[j];
            const int x = (int)(g % m->p.nx), y = (int)((g / m->p.nx) % m->p.ny), z = (int)(g / ((int64_t)m->p.nx * m->p.ny));
            const int32_t s = m->spin[g];
            for (int k = 0; k < m->nb.size[0]; ++k) {
                const int8_t* o = m->nb.d[0][k];
                const int fwd = o[2] > 0 || (o[2] == 0 && (o[1] > 0 || (o[1] == 0 && o[0] > 0)));
                if (!fwd) continue;
                const int32_t sj = mc_nb_spin(m, x, y, z, o);
                unlike += (uint64_t)(sj != 0) - (uint64_t)(sj == s);
            }
        }
        return m->p.J * (double)unlike;
    }
    const uint64_t prefix = orc_hash_prefix(m->p.seed, (uint64_t)m->sweeps);
    for (int64_t j = 0; j < m->nlist; ++j) {
        const int64_t g = m->list[j];
        const int x = (int)(g % m->p.nx), y = (int)((g / m->p.nx) % m->p.ny), z = (int)(g / ((int64_t)m->p.nx * m->p.ny));
        const int var = (int)(orc_hash_salt(orc_hash_index(prefix, (uint64_t)g), kSalts[S_NB]) % (uint64_t)m->nb.count);
        const int32_t s = m->spin[g];
        int ind = 0, same = 0;
        for (int k = 0; k < m->nb.size[var]; ++k) {
            const int32_t sj = mc_nb_spin(m, x, y, z, m->nb.d[var][k]);
            ind += sj != 0;
            same += sj == s;
        }
        unlike += (uint64_t)(ind - same);
    }
    return 0.5 * m->p.J * (double)unlike;
}

/* Connected components of equal positive orientation (Grains.cpp:33-150): union-find with
 * the smaller index kept as representative; Moore or von Neumann connectivity; periodic axes. */
static uint32_t uf_find(uint32_t* parent, uint32_t i) {
    while (parent[i] != i) { parent[i] = parent[parent[i]]; i = parent[i]; }
    return i;
}
uint64_t orc_grains(const int32_t* crystal, int nx, int ny, int nz, int moore, const int periodic[3],
                    uint32_t* root_opt, uint64_t* cells_out, uint64_t* largest_out) {
    const int64_t n = (int64_t)nx * ny * nz;
    orc_nb conn;
    orc_nb_build(moore ? ORC_MOORE : ORC_VON_NEUMANN, nx, ny, nz, &conn);
    uint32_t* parent = (uint32_t*)malloc(sizeof(uint32_t) * (size_t)(n ? n : 1));
    uint32_t* size = (uint32_t*)calloc((size_t)(n ? n : 1), sizeof(uint32_t));
    for (int64_t i = 0; i < n; ++i) parent[i] = (uint32_t)i;
    for (int z = 0; z < nz; ++z)
        for (int y = 0; y < ny; ++y)
            for (int x = 0; x < nx; ++x) {
                const int64_t i = x + (int64_t)nx * (y + (int64_t)ny * z);
                const int32_t s = crystal[i];
                if (s <= 0) continue;
                for (int k = 0; k < conn.size[0]; ++k) {
                    const int8_t* o = conn.d[0][k];
                    const int qx = axis_coord(x + o[0], nx, periodic[0]);
                    const int qy = axis_coord(y + o[1], ny, periodic[1]);
                    const int qz = axis_coord(z + o[2], nz, periodic[2]);
                    if (qx < 0 || qy < 0 || qz < 0) continue;
                    const int64_t j = qx + (int64_t)nx * (qy + (int64_t)ny * qz);
                    if (crystal[j] != s) continue;
                    const uint32_t a = uf_find(parent, (uint32_t)i), b = uf_find(parent, (uint32_t)j);
                    if (a < b) parent[b] = a; else if (b < a) parent[a] = b;
                }
            }
    uint64_t grains = 0, cells = 0, largest = 0;
    for (int64_t i = 0; i < n; ++i) {
        if (crystal[i] <= 0) { if (root_opt) root_opt[i] = 0xFFFFFFFFu; continue; }
        const uint32_t r = uf_find(parent, (uint32_t)i);
        if (root_opt) root_opt[i] = r;
        ++size[r];
        ++cells;
        grains += (r == (uint32_t)i);
    }
    for (int64_t i = 0; i < n; ++i) if (size[i] > largest) largest = size[i];
    free(parent);
    free(size);
    if (cells_out) *cells_out = cells;
    if (largest_out) *largest_out = largest;
    return grains;
}

/* measure (PottsMC.cpp:535-553) */
void orc_mc_measure(const orc_mc* m, orc_mc_stats* out, uint32_t* root_opt) {
    out->mcs = m->sweeps;
    out->grains = orc_grains(m->spin, m->p.nx, m->p.ny, m->p.nz, m->p.neighborhood != ORC_VON_NEUMANN,
                             m->p.periodic, root_opt, &out->cells, &out->largest);
    out->mean_grain_size = out->grains > 0 ? (double)out->cells / (double)out->grains : 0.0;
    out->energy = orc_mc_total_energy(m);
}

/* CoolingMobility.cpp:8-129 */
void orc_mobility_from_cooling(const uint8_t* type, const uint16_t* solid, const uint16_t* arrival, size_t n,
                               double mobility_min, uint8_t* domain_out, float* mobility_out) {
    uint32_t end = 0;
    for (size_t i = 0; i < n; ++i) {
        domain_out[i] = (type[i] == ORC_BASALT || type[i] == ORC_LAVA) ? 1 : 0;
        if (domain_out[i]) {
            const uint32_t a = solid[i] > arrival[i] ? solid[i] : arrival[i];
            if (a > end) end = a;
        }
    }
    uint32_t tmin = 0xFFFFFFFFu, tmax = 0;
    uint32_t* time = (uint32_t*)calloc(n ? n : 1, 4);
    for (size_t i = 0; i < n; ++i) {
        if (!domain_out[i]) continue;
        const uint32_t sd = type[i] == ORC_BASALT ? solid[i] : end;
        time[i] = sd > arrival[i] ? sd - arrival[i] : 0;
        if (time[i] < tmin) tmin = time[i];
        if (time[i] > tmax) tmax = time[i];
    }
    for (size_t i = 0; i < n; ++i) {
        if (!domain_out[i]) { mobility_out[i] = 0.0f; continue; }
        if (tmax == tmin) { mobility_out[i] = 1.0f; continue; }
        const double norm = (double)(time[i] - tmin) / (double)(tmax - tmin);
        mobility_out[i] = (float)(mobility_min + (1.0 - mobility_min) * norm);
    }
    free(time);
}

/* ======================================================================================
 * Terrain and vents (host-side inputs of the hot path), so whole configs can be replayed
 * independently of the product's host code.
 * ==================================================================================== */
/* TerrainGenerator.cpp:8-12 */
static double lattice_value(uint64_t seed, int octave, int i, int j) {
    const uint64_t key = (uint64_t)(uint32_t)i | ((uint64_t)(uint32_t)j << 32);
    return 2.0 * to_uniform(orc_hash_rand(seed, (uint64_t)octave, key, kSalts[S_TERRAIN])) - 1.0;
}
static double smooth_step(double t) { return t * t * (3.0 - 2.0 * t); } /* TerrainGenerator.cpp:15 */
/* TerrainGenerator.cpp:18-30 */
static double value_noise(uint64_t seed, int octave, int spacing, int x, int y) {
    const int i = x / spacing, j = y / spacing;
    const double fx = smooth_step((double)(x % spacing) / spacing);
    const double fy = smooth_step((double)(y % spacing) / spacing);
    const double v00 = lattice_value(seed, octave, i, j);
    const double v10 = lattice_value(seed, octave, i + 1, j);
    const double v01 = lattice_value(seed, octave, i, j + 1);
    const double v11 = lattice_value(seed, octave, i + 1, j + 1);
    const double a = v00 + (v10 - v00) * fx;
    const double b = v01 + (v11 - v01) * fx;
    return a + (b - a) * fy;
}
/* TerrainGenerator.cpp:45-51 */
static double smooth_noise(uint64_t seed, int x, int y) {
    const double n = 1.0 * value_noise(seed, 0, 16, x, y) + 0.5 * value_noise(seed, 1, 8, x, y) +
                     0.25 * value_noise(seed, 2, 4, x, y);
    return n / 1.75;
}
static double cone_height(double peak, double base, double r) {
    const double f = 1.0 - r / base;
    return peak * (f > 0.0 ? f : 0.0);
}
/* TerrainGenerator.cpp:39-43 */
static double center_distance(int nx, int ny, int x, int y) {
    const double cx = (nx - 1) / 2.0, cy = (ny - 1) / 2.0;
    return hypot(x - cx, y - cy);
}
/* TerrainGenerator.cpp:53-84 */
void orc_cone(int nx, int ny, int nz, double peak, double base, double crater_r, double crater_d, double noise,
              uint64_t seed, int32_t* out) {
    const double amplitude = noise * peak;
    const double floor_h = cone_height(peak, base, crater_r) - crater_d;
    for (int y = 0; y < ny; ++y)
        for (int x = 0; x < nx; ++x) {
            const double r = center_distance(nx, ny, x, y);
            double h = (r < crater_r) ? floor_h : cone_height(peak, base, r);
            h += amplitude * smooth_noise(seed, x, y);
            int cells = (int)ceil(h);
            if (cells < 0) cells = 0;
            if (cells > nz) cells = nz;
            out[x + nx * y] = cells;
        }
}

/* VentPlacer.cpp:11-58 (placement crater = 0, random = 1).  Returns the number of vents
 * written (x, y, z, flowId per vent), or -1 when there are fewer candidate columns than count. */
int orc_place_vents(int count, int placement, int is_cone, double crater_r, double base_r, int nx, int ny, int nz,
                    const int32_t* hm, uint64_t seed, int32_t* out) {
    int64_t ncand = 0;
    int32_t* cand = (int32_t*)malloc(sizeof(int32_t) * 2 * (size_t)nx * (size_t)ny);
    for (int y = 0; y < ny; ++y)
        for (int x = 0; x < nx; ++x) {
            const int h = hm[x + nx * y];
            if (h >= nz) continue;
            const double r = center_distance(nx, ny, x, y);
            const int ok = placement == 0 ? r <= crater_r : (r > crater_r && h >= 1 && (!is_cone || r < base_r));
            if (ok) { cand[2 * ncand] = x; cand[2 * ncand + 1] = y; ++ncand; }
        }
    if (ncand < count) { free(cand); return -1; }
    orc_mt64 g;
    orc_mt64_seed(&g, orc_splitmix64(seed ^ kSalts[S_VENTS]));
    for (int i = 0; i < count; ++i) {
        const uint64_t remaining = (uint64_t)(ncand - i);
        const int64_t j = i + (int64_t)(orc_mt64_next(&g) % remaining);
        const int32_t tx = cand[2 * i], ty = cand[2 * i + 1];
        cand[2 * i] = cand[2 * j]; cand[2 * i + 1] = cand[2 * j + 1];
        cand[2 * j] = tx; cand[2 * j + 1] = ty;
        out[4 * i] = cand[2 * i];
        out[4 * i + 1] = cand[2 * i + 1];
        out[4 * i + 2] = hm[cand[2 * i] + nx * cand[2 * i + 1]];
        out[4 * i + 3] = i + 1;
    }
    free(cand);
    return count;
}

/* VentPlacer.cpp:83-104: classic seeds in distinct random cells. */
int orc_place_seeds(int count, int nx, int ny, int nz, uint64_t seed, int32_t* out) {
    const uint64_t n = (uint64_t)nx * ny * nz;
    uint8_t* taken = (uint8_t*)calloc((size_t)n, 1);
    orc_mt64 g;
    orc_mt64_seed(&g, orc_splitmix64(seed ^ kSalts[S_VENTS]));
    int k = 0;
    while (k < count) {
        const uint64_t idx = orc_mt64_next(&g) % n;
        if (taken[idx]) continue;
        taken[idx] = 1;
        out[4 * k] = (int32_t)(idx % (uint64_t)nx);
        out[4 * k + 1] = (int32_t)((idx / (uint64_t)nx) % (uint64_t)ny);
        out[4 * k + 2] = (int32_t)(idx / ((uint64_t)nx * (uint64_t)ny));
        out[4 * k + 3] = k + 1;
        ++k;
    }
    free(taken);
    return count;
}

/* Test hooks: set one orientation (PottsMC::setOrientation, PottsMC.cpp:566-576) and read the
 * colour-major domain list (PottsMC.hpp:150-151). */
void orc_mc_set_spin(orc_mc* m, int64_t g, int32_t s) { m->spin[g] = s; }
int orc_mc_color_list(const orc_mc* m, int64_t* starts, int64_t* list) {
    for (int c = 0; c <= m->colors; ++c) starts[c] = m->color_start[c];
    memcpy(list, m->list, sizeof(int64_t) * (size_t)m->nlist);
    return m->colors;
}

// Internal types shared by the HIP kernels (vca_device.hip) and the C-ABI layer (vca_api.hip).
//
// Grid layout in HBM: structure of arrays, unframed, global-index order i = x + nx*(y + ny*z).
// Boundaries are evaluated by index wrap (periodic axis) or as VOID (absorbing axis) instead of a
// copied frame (the reference's frame: engine/src/core/Boundary.hpp:66-131).
//
// Neighbourhoods are bitmasks over the 27 positions of a 3x3x3 block, position
//   k = (dx+1) + 3*(dy+1) + 9*(dz+1)      (k = 13 is the cell itself),
// one mask per variant, variants in the reference order (engine/src/core/Neighborhood.cpp:55-119).
// The order of offsets inside a variant does not matter on the device: the flowId vote, the maximum
// temperature and the Potts unlike-counts are all order independent.  k > 13 is exactly the
// reference's "forward" half (positive linear offset, engine/src/mc/PottsMC.cpp:599-604).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace vca {

constexpr uint8_t kAir = 0, kRock = 1, kLava = 2, kBasalt = 3;
constexpr int kMaxDrop = 16;  // engine/src/ca/LavaCA.hpp:1293

// Salts (engine/src/core/Rng.hpp:22-31)
constexpr uint64_t kSaltNeighborhood = 0x4E45494748424F52ULL;
constexpr uint64_t kSaltTieBreak = 0x5449455F42524B31ULL;
constexpr uint64_t kSaltMcInit = 0x4D43494E49544F52ULL;
constexpr uint64_t kSaltMcPick = 0x4D435049434B4345ULL;
constexpr uint64_t kSaltMcMobility = 0x4D434D4F42494C49ULL;
constexpr uint64_t kSaltMcOrient = 0x4D434F5249454E54ULL;
constexpr uint64_t kSaltMcAccept = 0x4D43414343455054ULL;

__host__ __device__ inline uint64_t splitmix64(uint64_t x) {
    x += 0x9E3779B97F4A7C15ULL;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ULL;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBULL;
    return x ^ (x >> 31);
}
// hashPrefix / hashIndex / hashSalt (engine/src/core/Rng.hpp:66-72)
__host__ __device__ inline uint64_t hash_prefix(uint64_t seed, uint64_t step) {
    return splitmix64(splitmix64(seed ^ 0xA0761D6478BD642FULL) ^ step);
}

// Plain data (no member initialisers): it lives in the constant-memory simulation table.
struct NbTable {
    int32_t variants;
    uint32_t mask[6];
    int32_t max_size;
    uint32_t any_mask;            // union over variants
    uint32_t side_mask;           // union over variants of dz == 0 positions (D(n), LavaCA.hpp:102-112)
    int32_t diagonal_falls;       // some variant has dz == -1 with dx|dy != 0
    uint32_t conn_mask;           // grain connectivity: Moore, or von Neumann when the kind is von Neumann
    uint32_t var_r32;             // 2^32 mod variants and 1 / variants: the CA rules draw a cell's variant without
    double var_inv;               // integer division (mod_small_inv)
};

// Tile = 64 x TY x TZ cells (TY*TZ = 16): one 256-thread workgroup, 4 cells per thread.
constexpr int kTileX = 64;
constexpr int kTileCells = 1024;
constexpr int kBlock = 256;

struct CaParams {
    int32_t nx, ny, nz;
    int32_t classic;
    int32_t per[3];
    uint64_t seed;
    int32_t max_steps, eruption_steps;
    float erupt_temp, solidus, cool_amb, cool_lat, cool_fall;
    int32_t ponding;
    int32_t dense;
    NbTable nb;
    int32_t ty, tz;               // tile extent in y and z
    int32_t ntx, nty, ntz;        // tiles per axis
    uint32_t ntiles;
    // t / ntx and r / nty for t, r < 2^23 as (x * m) >> p (exact; see magic_div)
    uint32_t div_ntx_m, div_nty_m;
    int32_t div_ntx_p, div_nty_p;
    int32_t nvent_slots;
};

// Division by an invariant d >= 1 for dividends x < 2^23: p = 23 + ceil(log2 d), m = ceil(2^p / d) (< 2^24);
// then floor(x / d) == (x * m) >> p, because m * d - 2^p < d <= 2^(p - 23) and x < 2^23.
inline void magic_div(uint32_t d, uint32_t& m, int32_t& p) {
    int l = 0;
    while ((1ULL << l) < d) ++l;
    p = 23 + l;
    m = static_cast<uint32_t>(((1ULL << p) + d - 1) / d);
}
__host__ __device__ inline uint32_t div_magic(uint32_t x, uint32_t m, int32_t p) {
    return static_cast<uint32_t>((static_cast<uint64_t>(x) * m) >> p);
}

struct CaCtrl {
    int32_t steps;
    int32_t stop;
    unsigned long long lava_total, air_total;
    unsigned long long tiles_visited, changes_total;
    unsigned long long tiles_staged;
    // first step from which no AIR cell has LAVA at a position the rules read (none can ever be born again: every
    // later step only cools the lava cells, with the vents' eruption), 0 while there is one; set by the finish of
    // the step before it
    int32_t quiet;
    // = steps when the last step changed no cell's type and none of its AIR candidates had a source the rules accept
    // (under any neighbourhood variant) with a candidate temperature at the solidus or above: until a cell's type
    // changes, no cell can be born; 0 otherwise (set by every finish)
    int32_t qdead;
    int32_t born_last;            // cells born in the last step were counted (the rules kernel then skips the check)
    // the fast-forward: the step it last started from (-1: never) and the steps it completed in all
    int32_t ff_from, ff_steps;
    int32_t ff_sol;               // fast-forward: steps from its start to the first step in which a lava cell solidifies
    uint64_t prefix;              // hashPrefix(seed, steps): the hash prefix of the next step (kept with `steps` by
                                  // every kernel that advances it, so that the rules kernel loads it)
    // set by the host while it steps a lava-front simulation with the fast-forward on: the finish of a step after
    // which no AIR cell has a candidate (quiet) holds the simulation (stop = kStopHeld) instead of letting the
    // remaining launched steps run; a held simulation's steps are no-ops that keep its tiles listed, until the host's
    // fast-forward (which the quiet condition admits) takes over from the held step
    int32_t hold;
};
constexpr int32_t kStopHeld = 8;  // CaCtrl::stop of a held simulation (running for every reader outside the kernels)

// Counters of one step spread over kPartRows rows (a tile's counts go to row tile % kPartRows: visited tiles
// from ca_front, the rest from the rules kernel; ca_finish sums and clears them), so that no counter is one
// address every block updates.
struct CaPart {
    unsigned long long lava_after, born, changed, tiles, staged;
    unsigned long long air_cand;  // AIR cells with LAVA at a position the rules read (candidates of rule 1)
    unsigned long long air_alive; // of those, the ones not born that a later step could bear (see CaCtrl::qdead)
    unsigned long long tchanged;  // cells whose type changed (born or solidified)
};
constexpr int kPartRows = 256;

// MC counters spread the same way: kMcRows rows; a sweep's counter row holds attempts, accepted, uphill;
// a measurement row holds kMcSlots values: 0 grains, 1 largest (max), 2..4 / 5..7 count, cells, largest (max)
// of the high / low mobility class, 8 high-mobility cells, 9 unlike neighbour pairs.
constexpr int kMcRows = 64;
constexpr int kMcCounters = 4;
constexpr int kMcSlots = 16;

struct CaDev {
    // cell types at the start of a step, by the step's parity: step t reads type[t & 1] and writes the tile's cells
    // that differ into type[(t + 1) & 1] -- its own changes, and the changes of step t - 1 that copy does not hold yet
    // (every tile that changed a cell in step t - 1 is visited in step t); type[steps & 1] is the state after the
    // last step
    uint8_t* type[2];
    // flowId, solidification and arrival steps: written in place by the step that changes the cell (a cell becomes
    // LAVA only from AIR, and no rule reads the flowId of a cell that is not LAVA at the step's start)
    int32_t* flow;
    // temperatures at the start of a step, by the step's parity: step t reads temp[t & 1] and writes every lava cell's
    // new temperature (also the last one of a cell that solidifies) to temp[(t + 1) & 1].  temp[steps & 1] is the
    // state after the last step except for BASALT cells, whose last temperature is in temp[(solid + 1) & 1] (no rule
    // reads the temperature of a cell that is not LAVA); ca_final_temps copies those before a read of the state
    float* temp[2];
    uint16_t* solid;
    uint16_t* arrival;
    int8_t* descent;
    int32_t* spill;               // nx*ny, nullptr in classic mode
    uint32_t* vent_cell;          // by flowId, 0xFFFFFFFF = none
    uint8_t* tile_flags;          // bit0: has lava, bit1: has air
    uint8_t* tile_near;           // 1: one of the 26 neighbour tiles held lava when its last step ended (set by
                                  // that tile's rules, read and cleared by ca_front in lava mode; kept in classic)
    uint8_t* tile_halo;           // kHaloBytes per tile: byte q = 1 when the neighbour tile in direction q holds lava
    uint32_t* tile_tchg;          // per tile: 1 + the last step in which a cell of the tile changed type (0: never)
    uint32_t* listed;             // per tile: the step whose tile list holds the tile (lava front: the rules kernel
                                  // pushes the next step's tiles; the stamp lists each tile once)
    uint32_t* dirtied;            // per tile: the step whose descent list holds the tile (stamp, as `listed`)
    uint8_t* tile_vent;           // per tile: 1 when one of its cells is the vent cell of some flowId (the only tiles
                                  // whose lava cells the eruption rule has to test)
    CaPart* part;                 // kPartRows rows
    CaCtrl* ctrl;
    unsigned long long* hist_active;
    unsigned long long* hist_changed;
};
constexpr uint8_t kFlagLava = 1, kFlagAir = 2;
constexpr int kHaloBytes = 32;   // 27 direction bytes per tile, padded

// Tile shapes (64 x TY x TZ): 0 = 4x4, 1 = 2x8, 2 = 8x2 (3D), 3 = 1x16 (x-z plane, ny = 1), 4 = 16x1 (nz = 1).
constexpr int kShapes = 5;
constexpr int kShapeTY[kShapes] = {4, 2, 8, 1, 16};
constexpr int kShapeTZ[kShapes] = {4, 8, 2, 16, 1};

// One simulation of a batch, as the step kernels see it.
struct CaSim {
    CaParams P;
    CaDev D;
    int32_t running;              // 1 while the simulation takes part in the batch's steps
    int32_t shape;
    int32_t push;                 // lava front (lava mode, ca.sweep = front): the rules kernel lists the next step's
                                  // tiles; otherwise ca_front scans the tiles every step
};

// A batch: simulations in slots, their tiles in one global tile range.  Work-list entries are
// slot << kTileBits | tile (tile local to the simulation); the visited-tile lists carry the simulation's step
// count in the high word as well, so the rules kernel reads both with one load.
constexpr int kTileBits = 22;
constexpr uint32_t kTileMask = (1u << kTileBits) - 1u;
constexpr int kMaxSlots = 1024;
constexpr uint32_t kNoSlot = 0xFFFFFFFFu;

// Tile lists come in two copies: a step reads the copy of its parity and the rules kernel pushes the next step's
// tiles into the other one (lava front); the first kernel of a step clears the counts of the copy the previous step
// read.  Each list (parity, shape) can be split into kSub sub-lists: a tile goes to the sub-list its index within its
// simulation selects (sub_of in vca_device.hip), each sub-list's count on a 64-byte line of its own -- device atomics
// on one 64-byte line serialise (about 12 ns each on MI355X, measured), and thousands of tiles append per step.
// Measured with 32 sub-lists: no gain on the 1024 x 1024 x 384 caldera, 3 % slower on the 512 x 512 x 256 cone (the
// producers' grouping by sub-list costs more scalar issue than the serialised appends), so one list is used.
constexpr int kSub = 1;
struct BatchCtrl {
    uint32_t nactive[2][kShapes][kSub][16];  // staged tiles per step parity, shape and sub-list (count at [0])
};

// Simulations of every batch live in one constant-memory table on each device (read by the step kernels through
// scalar loads); a batch owns entries sim_base .. sim_base + its slot capacity.
constexpr int kSimTable = 4096;

struct BatchDev {
    uint32_t sim_base;            // first table entry of this batch (slot s -> entry sim_base + s)
    uint32_t* owner;              // global tile -> slot (kNoSlot for a free tile)
    uint32_t* tile_base;          // slot -> first global tile
    uint32_t ntiles;              // global tiles in use (high-water mark)
    uint32_t nslots;
    uint32_t subcap;              // capacity of each sub-list: a simulation of n tiles holds at most ceil(n / kSub) of
                                  // a sub-list's entries, so ceil(tiles / kSub) + slots covers every batch
    uint64_t* active;             // 2 x kShapes x kSub sub-lists of subcap entries (parity-major): step << 32 | entry
    // 2 x kShapes descent lists of subcap entries (parity-major, entries as in `active`): the tiles whose D(n) may have
    // changed -- every tile whose 3 x 3 columns down to the layers its drops reach hold a tile that changed a cell's
    // type in the previous step, pushed by that step's rules kernel; nullptr when the batch keeps none (the descent
    // then checks every listed tile)
    uint64_t* desc_list;
    uint32_t* desc_count;         // [2][kShapes][16]: entries of descent list (parity, shape) at [par][shape][0]
    BatchCtrl* ctrl;
};

struct McParams {
    int32_t nx, ny, nz;
    int32_t per[3];
    uint64_t seed;
    int32_t Q;
    NbTable nb;
};

constexpr uint64_t kDownhill = ~0ULL;  // accept-table entry for dE <= 0

// Colour passes of a checkerboard sweep: colour c is list[start[c] .. start[c + 1]) (at most 3 x 3 x 3 colours).
struct McColours {
    uint32_t start[28];
    int32_t count;
};

// Barrier state of the persistent sweep kernel (one per MC simulation, zero at creation): arrival counters of the
// eight block groups, of the groups, the generation and the give-up flag, each on a 256-byte line of its own.
struct McGridBar {
    uint32_t group[8][64];
    uint32_t top[64];
    uint32_t gen[64];
    uint32_t abort[64];
};

// Source of the MC domain lists built on the device: mode 0 every cell, 1 dom[i] != 0, 2 dom = CA cell types
// (LAVA or BASALT).  ncol colours (1 = one list in index order); parity colouring or index[] over the axis
// classes as in buildColors (PottsMC.cpp:145-224).
struct McListSrc {
    uint64_t n;
    const uint8_t* dom;
    int32_t mode;
    int32_t ncol;
    int32_t parity;
    int32_t dims[3];
    int32_t per[3];
    int32_t index[27];
};

// ---- launchers (vca_device.hip) ----
void ca_launch_terrain(const CaParams& P, const CaDev& D, const int32_t* height, hipStream_t s);
void ca_launch_vent_tiles(const CaParams& P, const CaDev& D, hipStream_t s);
void ca_launch_seeds(const CaParams& P, const CaDev& D, const uint32_t* cells, const int32_t* flows,
                     const float* temps, int n, hipStream_t s);
void ca_launch_init_flags(const CaParams& P, const CaDev& D, hipStream_t s);
void ca_launch_hold(CaCtrl* c, int hold, hipStream_t s);
void ca_launch_init_stop(const CaParams& P, const CaDev& D, hipStream_t s);
// One step of every running simulation of the batch.
struct StepLaunch {
    uint32_t shapes;              // bit mask of the tile shapes present
    bool lava_mode;               // some member runs in lava mode (D(n) is needed)
    bool flow16;                  // every flowId of the batch fits int16_t (the rules kernel holds them in 16 bits)
    bool scan;                    // some member needs ca_front (classic mode or ca.sweep = full): then ca_finish
                                  // runs after every step; otherwise the previous step's finish runs inside the
                                  // first kernel of the next step
    int par;                      // parity of the step (which list copy it reads)
    bool fin_prev;                // the previous step's finish is still to be done (by this step's first kernel)
    bool desc_list;               // the descent walks the descent lists (every member's previous step ran its rules
                                  // kernel in this batch), else every listed tile with its own check
    int grid[kShapes];            // blocks of the tile kernels per shape: at most the batch's grid, at most the tiles
                                  // of the members with that shape (a list holds each tile at most once)
};
// Kernel timing: kCaKernels kernels in launch order; ev[2k], ev[2k + 1] (when non-null) are recorded by kernel k's
// own dispatch (start and end); a kernel not launched in this step leaves its pair unrecorded (flag false).
constexpr int kCaKernels = 4;     // front, descent, rules, finish
void ca_launch_step(const BatchDev& B, const StepLaunch& L, hipStream_t s, hipEvent_t* ev, bool* recorded);
// The finish of the last step enqueued (after a run of steps whose finish is done by the next step's first kernel);
// ev (nullable): the start / end events of the launch.
void ca_launch_finish(const BatchDev& B, hipStream_t s, hipEvent_t* ev);
// Lists the tiles of slot's simulation that the next step stages (lava front: tiles with lava or with lava in their
// 1-cell halo), into the copy of parity par.
void ca_launch_seed(const BatchDev& B, uint32_t slot, uint32_t ntiles, int par, hipStream_t s);
// Removes the entries of the slots in `slots` from the copy of parity par (their simulations left the batch; entries
// they pushed for the next step would otherwise stay until a step reads and skips them).
struct SlotMask {
    uint32_t bits[kMaxSlots / 32];
};
void ca_launch_purge(const BatchDev& B, const SlotMask& slots, int par, hipStream_t s);
// p[0 .. n) = v (a kernel, not a memset: usable where the caller may capture the stream)
void fill_u32(uint32_t* p, uint32_t v, uint64_t n, hipStream_t s);
// owner[base .. base + n) = slot
void ca_launch_owner(uint32_t* owner, uint32_t base, uint32_t n, uint32_t slot, hipStream_t s);
// out[4*s] = steps, out[4*s+1] = stop (-1 when slot s is not running), out[4*s+2] = CaCtrl::quiet, out[4*s+3] =
// CaCtrl::qdead, for s < nslots.
void ca_launch_status(const BatchDev& B, int32_t* out, hipStream_t s);
// Fast-forward of one simulation from step t0 over at most h steps, while no cell can be born: each lava cell's
// cooling (and its vent's eruption) replayed on its own, the histories of those steps, the state after them in both
// type and both temperature copies, and CaCtrl as stepping on would leave it -- stopped when a stop rule fires within
// them.  until_solid (CaCtrl::qdead): only up to and including the first step in which a lava cell solidifies (a
// type change may let a cell be born afterwards); otherwise (CaCtrl::quiet) over all h.  When the simulation goes
// on, its entries in list copy par of batch B (slot) are moved to the new step.  ev (nullable): start / end events.
void ca_launch_fast_forward(const CaParams& P, const CaDev& D, const BatchDev& B, uint32_t slot, int shape, int par,
                            int32_t t0, int32_t h, bool until_solid, hipStream_t s, hipEvent_t* ev);
// Estimate of the steps from t0 to the first solidification into CaCtrl::ff_sol (no replay): whether a fast-forward
// up to it is worth its launches.
void ca_launch_ff_estimate(const CaParams& P, const CaDev& D, const BatchDev& B, uint32_t slot, int shape, int par,
                           int32_t t0, hipStream_t s);
// Writes entry `index` of the device's simulation table (stream ordered; `sim` must stay valid until the
// stream reaches the copy).
void ca_write_sim(uint32_t index, const CaSim* sim, hipStream_t s);

// After the CA: per-flow LAVA / BASALT counts and largest squared vent distance (arrays of nflows + 1, zeroed by
// the caller; vent_xy = x, y of flow f at 2*(f-1)); seen[|dx|] = 1 for the cells of one flow at squared distance d2.
// The last temperature of every BASALT cell into temperature copy par (= steps & 1) before that copy is read.
void ca_launch_final_temps(const CaDev& D, uint64_t n, int par, hipStream_t s);
// par: the type copy holding the state (the step count's parity)
void ca_launch_flow_sums(const CaParams& P, const CaDev& D, int par, const int32_t* vent_xy, int32_t nflows,
                         unsigned long long* lava, unsigned long long* basalt, unsigned long long* maxd2,
                         hipStream_t s);
void ca_launch_flow_points(const CaParams& P, const CaDev& D, int par, int32_t flow, int32_t vx, int32_t vy,
                           unsigned long long d2, uint32_t* seen, uint32_t cap, hipStream_t s);
// Cooling mobility of every cell from the CA result; red[3] = {0, 0xFFFFFFFF, 0} on entry (end, tmin, tmax).
void mc_launch_cooling(uint64_t n, const uint8_t* type, const uint16_t* solid, const uint16_t* arrival,
                       uint32_t* red, double mobility_min, float* mob, hipStream_t s);
// Ordered domain lists: count + scan (cnt: ncol * segments entries; totals: ncol), then scatter (colour_base =
// first list slot of each colour; thr != nullptr also writes the mobility threshold of each entry from mob).
uint64_t mc_list_segments(uint64_t n);
void mc_launch_list_count(const McListSrc& L, uint32_t* cnt, uint32_t* totals, hipStream_t s);
void mc_launch_list_scatter(const McListSrc& L, const uint32_t* off, const uint32_t* colour_base, uint32_t* list,
                            const float* mob, uint64_t* thr, hipStream_t s);

// Colour pass of the checkerboard sweep.  spin_bytes: 1, 2 or 4.
void mc_launch_color(const McParams& P, void* spin, int spin_bytes, const uint32_t* list, const uint64_t* mob_thr,
                     uint32_t count, uint64_t prefix, const uint64_t* accept, unsigned long long* counters,
                     hipStream_t s);
// nsweeps checkerboard sweeps in one launch of `grid` blocks (at most one per CU), a device-wide barrier between
// colour passes; `limit` bounds each barrier wait in ticks of the device's constant-rate clock.
void mc_launch_sweeps(const McParams& P, void* spin, int spin_bytes, const uint32_t* list, const uint64_t* mob_thr,
                      const McColours& C, uint64_t sweep0, int32_t nsweeps, const uint64_t* accept,
                      unsigned long long* counters, McGridBar* bar, uint64_t limit, unsigned grid, hipStream_t s);
void mc_launch_random(const McParams& P, void* spin, int spin_bytes, const uint32_t* list, const uint64_t* mob_thr,
                      uint32_t count, uint64_t prefix, const uint64_t* accept, unsigned long long* counters,
                      hipStream_t s);
void mc_launch_init_spins(const McParams& P, void* spin, int spin_bytes, const uint32_t* list, uint32_t count,
                          hipStream_t s);
void mc_launch_energy(const McParams& P, const void* spin, int spin_bytes, const uint32_t* list, uint32_t count,
                      uint64_t prefix, unsigned long long* unlike, hipStream_t s);
// Grain labels: parent (N), size (N), hi (N, nullable); out[0..] = grains, largest, class counters.
void mc_launch_grains(const McParams& P, const void* spin, int spin_bytes, const uint32_t* list, uint32_t count,
                      uint32_t* parent, uint32_t* size, uint32_t* hi, const float* mobility,
                      unsigned long long* out, hipStream_t s);
void mc_launch_download(const McParams& P, const void* spin, int spin_bytes, int32_t* out, hipStream_t s);
void mc_launch_labels(const McParams& P, const void* spin, int spin_bytes, const uint32_t* parent, uint32_t* out,
                      hipStream_t s);

}  // namespace vca

/* vca.h — C ABI of the MI355X lava-CA / Potts-MC engine (libvca.so).
 *
 * The reference engine has no FFI layer: its hot paths sit behind two C++ class templates,
 * LavaCA<Storage> (engine/src/ca/LavaCA.hpp:80-296) and PottsMC (engine/src/mc/PottsMC.hpp:52-268),
 * called by Simulation::runWith (engine/src/app/Simulation.cpp:158-292) and McStage::run
 * (engine/src/app/McStage.cpp:93-301).  This header replaces those two classes with opaque handles
 * whose state lives in GPU memory; each entry point names the member function it stands in for.
 *
 * Conventions
 *   - Every function returns VCA_OK (0) or a negative VCA_ERR_* code; vca_last_error() gives the
 *     message of the last failure on the calling thread.  No C++ exception crosses this boundary.
 *   - Host arrays are owned by the caller; device memory is owned by the handle.
 *   - Arrays are unframed and in global-index order i = x + nx*(y + ny*z) (engine/src/core/Types.hpp:95-100),
 *     the order of ca_result.bin / mc_result.bin.
 *   - One handle is driven by one host thread at a time.  Calls on one context are stream ordered.
 */
#ifndef VCA_H
#define VCA_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define VCA_OK 0
#define VCA_ERR_ARG (-1)      /* invalid argument (reference: std::invalid_argument / ConfigError) */
#define VCA_ERR_HIP (-2)      /* HIP runtime failure */
#define VCA_ERR_NOMEM (-3)    /* device allocation failed */
#define VCA_ERR_STATE (-4)    /* call not valid in the handle's state */
#define VCA_ERR_DEVICE (-5)   /* no usable gfx950 device / library built without it */

/* Cell types (engine/src/core/Types.hpp:13-19); values are part of the file format. */
#define VCA_AIR 0
#define VCA_ROCK 1
#define VCA_LAVA 2
#define VCA_BASALT 3

/* Neighbourhoods (engine/src/core/Types.hpp:22) */
#define VCA_VON_NEUMANN 0
#define VCA_MOORE 1
#define VCA_PENTAGONAL 2
#define VCA_HEXAGONAL 3

/* Stop reasons (engine/src/ca/CaRules.hpp:108) */
#define VCA_RUNNING 0
#define VCA_NO_LAVA 1
#define VCA_NO_AIR 2
#define VCA_NO_CHANGE 3
#define VCA_MAX_STEPS 4

typedef struct vca_ctx vca_ctx;
typedef struct vca_ca vca_ca;
typedef struct vca_mc vca_mc;

/* ---- library / context ------------------------------------------------------------------ */
int vca_version(void);
/* Copies the last error message of this thread (NUL-terminated, truncated to cap). */
int vca_last_error(char* buf, size_t cap);
/* Context = one device + one HIP stream.  stream == NULL creates a private default-priority stream;
 * otherwise the caller's hipStream_t is used (not destroyed by vca_ctx_destroy). */
int vca_ctx_create(int device, void* stream, vca_ctx** out);
int vca_ctx_destroy(vca_ctx* ctx);
int vca_ctx_synchronize(vca_ctx* ctx);
/* Free / total device memory in bytes (admission control of the serving loop). */
int vca_device_memory(vca_ctx* ctx, uint64_t* free_bytes, uint64_t* total_bytes);

/* ---- lava / classic CA: replaces LavaCA<Storage> ---------------------------------------- */
typedef struct {
    int32_t nx, ny, nz;          /* Dims (Types.hpp:38-72); 2D lava is the x-z plane (ny = 1) */
    int32_t classic;             /* CaParams::mode: 0 lava, 1 classic grain growth (CaRules.hpp:93-105) */
    int32_t neighborhood;        /* VCA_VON_NEUMANN .. VCA_HEXAGONAL */
    int32_t periodic;            /* BoundaryKind: 0 absorbing, 1 periodic (Boundary.cpp:5-11) */
    uint64_t seed;               /* CaParams::seed */
    int32_t max_steps;           /* CaParams::maxSteps (<= 65535) */
    int32_t eruption_steps;      /* CaParams::eruptionSteps */
    float erupt_temp, solidus_temp, cooling_ambient, cooling_lateral, cooling_fall;
    int32_t ponding;             /* CaParams::ponding */
    int32_t dense;               /* 0: visit only tiles near lava (ca.sweep = front); 1: every tile (full) */
    int32_t tile_shape;          /* 0: default for the dims; 1..5: 64x4x4, 64x2x8, 64x8x2, 64x1x16, 64x16x1
                                    (results do not depend on it; vca_ca_autotune picks one) */
} vca_ca_desc;

typedef struct {
    int32_t x, y, z, flow_id;    /* Vent (engine/src/terrain/VentPlacer.hpp:18-23) */
} vca_vent;

/* LavaCA ctor + applyTerrain + placeSeeds (LavaCA.hpp:84-227).  height / spill: nx*ny column
 * arrays (height = rock cells per column, spill = priority-flood spill level; both NULL in classic
 * mode).  Counts the initial state (countInitialState, LavaCA.hpp:1252-1280). */
int vca_ca_create(vca_ctx* ctx, const vca_ca_desc* desc, const int32_t* height, const int32_t* spill,
                  const vca_vent* vents, int32_t nvents, vca_ca** out);
int vca_ca_destroy(vca_ca* ca);
/* LavaCA::step x n (LavaCA.hpp:230-256): enqueues up to max_new_steps synchronous steps; steps after a
 * stop rule fired are no-ops on the device.  Blocks until done, returns the step count and stop reason.
 * When the run would reach the simulation's last step and the simulation has become quiescent (lava mode, the
 * eruption over, and in the last step no AIR cell had LAVA at a position the rules read: from then on every step
 * only cools the lava cells, LavaCA.hpp:1129-1147), the remaining steps are completed by the fast-forward: each lava
 * cell's cooling replayed on its own, with the same float arithmetic, histories and stop rule (finishStep,
 * LavaCA.hpp:1206-1229) -- the state, histories and step count are those of stepping on. */
int vca_ca_run(vca_ca* ca, int32_t max_new_steps, int32_t* steps_done, int32_t* stop_reason);
/* The fast-forward of vca_ca_run, vca_ca_run_noalloc and vca_batch_status / vca_batch_run (no counterpart in the
 * reference, which steps on): on (1, the default) or off (0, every step is launched).  Results do not depend on it. */
int vca_ca_set_fast_forward(vca_ca* ca, int32_t on);
/* The step the fast-forward started from (-1 when it did not run) and the steps it completed. */
int vca_ca_fast_forwarded(vca_ca* ca, int32_t* from_step, int32_t* steps);
/* vca_ca_run without any device allocation.  No counterpart in the reference: for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that forbid it).  The simulation (created
 * beforehand by vca_ca_create) is advanced exactly as vca_ca_run does; the batch-of-one state vca_ca_run allocates on
 * first use (list counts, tile owners, the two copies of the tile lists) lives instead in a static device array of
 * 16 MiB per device, set up by vca_ctx_create: grids of at most 199,713 tiles of 1024 cells, about 204 M cells
 * (VCA_ERR_ARG beyond, with the limit in the message).  Calls on one device are serialised by a lock; the array keeps the last call's
 * contents and is re-initialised by every call.  VCA_ERR_STATE when the simulation is in a batch or was advanced by
 * vca_ca_run / vca_ca_enqueue before (their private batch holds its lists). */
int vca_ca_run_noalloc(vca_ca* ca, int32_t max_new_steps, int32_t* steps_done, int32_t* stop_reason);
/* Enqueue only (no host wait); pair with vca_ca_status. */
int vca_ca_enqueue(vca_ca* ca, int32_t new_steps);
/* LavaCA::steps / stopReason (LavaCA.hpp:262-264); synchronises the context stream. */
int vca_ca_status(vca_ca* ca, int32_t* steps_done, int32_t* stop_reason);
/* activeHistory / changedHistory (LavaCA.hpp:266-270); returns the number of entries in *n. */
int vca_ca_histories(vca_ca* ca, uint64_t* active, uint64_t* changed, int32_t cap, int32_t* n);
/* LavaCA::result (LavaCA.hpp:276): any pointer may be NULL. */
int vca_ca_download(vca_ca* ca, uint8_t* type, int32_t* flow, float* temp, uint16_t* solid, uint16_t* arrival);
/* Work statistics: tiles visited and changes applied over all steps, tile size in cells. */
int vca_ca_work(vca_ca* ca, uint64_t* tiles_visited, uint64_t* changes, int32_t* tile_cells);

/* Everything vca_ca_status / vca_ca_histories / vca_ca_work / vca_ca_download give, in one call that reads on a
 * stream of the simulation's own through pinned staging buffers, so it does not wait for other work of the
 * context (another thread may keep advancing a batch meanwhile).  The simulation must have stopped and must not
 * be in a batch (VCA_ERR_STATE otherwise).  Any array pointer may be NULL; histories get min(steps, cap) entries. */
typedef struct {
    int32_t steps, stop_reason;
    uint64_t tiles_visited, changes;
    uint8_t* type;
    int32_t* flow;
    float* temp;
    uint16_t* solid;
    uint16_t* arrival;
    uint64_t* active;
    uint64_t* changed;
    int32_t cap;
    uint64_t tiles_staged; /* visited tiles whose halo held lava (the rest settle without being read) */
    int32_t ff_from;       /* the step the fast-forward started from, -1 when it did not run */
    int32_t ff_steps;      /* steps completed by the fast-forward (steps - ff_steps were launched) */
} vca_ca_result;
int vca_ca_collect(vca_ca* ca, vca_ca_result* r);

/* flowSums of the final state (Simulation.cpp:408-466) on the device: for flowId f = 1..nflows, lava[f] / basalt[f]
 * = LAVA / BASALT cells of the flow and maxd2[f] = the largest (x - vx)^2 + (y - vy)^2 over them, (vx, vy) =
 * vents[f-1] (arrays of nflows + 1 entries; entry 0 and flows without cells are 0).  The reach is hypot of the
 * farthest cell: every cell at that squared distance has the same true distance, and vca_ca_flow_points tells
 * which lattice points of the circle occur when the host libm's hypot differs between them.  Like vca_ca_collect:
 * the simulation must have stopped and must not be in a batch; reads on a stream of its own. */
int vca_ca_flow_stats(vca_ca* ca, const vca_vent* vents, int32_t nflows, uint64_t* lava, uint64_t* basalt,
                      uint64_t* maxd2);
/* seen[a] = 1 when a LAVA / BASALT cell of flow `flow` lies at |x - vent.x| = a, (x - vent.x)^2 + (y - vent.y)^2
 * = d2 (a < cap). */
int vca_ca_flow_points(vca_ca* ca, const vca_vent* vent, int32_t flow, uint64_t d2, uint8_t* seen, int32_t cap);

/* Device memory a simulation of these dims takes (state, change slots, tile lists), for admission control. */
int vca_ca_footprint(const vca_ca_desc* desc, uint64_t* bytes);
/* Autotuner (no counterpart in the reference): runs the first probe_steps steps of this simulation once per
 * candidate launch configuration — every tile shape that suits the dims, each with the active-front compaction
 * (dense = 0) and without it (dense = 1) — and returns the fastest by the device time of a whole step.  Results do
 * not depend on either choice, so the caller may put the winner into the descriptor (tile_shape, dense) of the real
 * run.  The context's kernel timing state is left as it was. */
typedef struct {
    int32_t tile_shape;          /* 1..5 */
    int32_t dense;
    int32_t candidates, probe_steps;
    float ms_per_step[2][5];     /* [dense][tile_shape - 1], 0 = not tried */
} vca_tune_result;
int vca_ca_autotune(vca_ctx* ctx, const vca_ca_desc* desc, const int32_t* height, const int32_t* spill,
                    const vca_vent* vents, int32_t nvents, int32_t probe_steps, vca_tune_result* out);
/* Tiles of the simulation (its share of a batch's tile range). */
int vca_ca_tiles(vca_ca* ca, uint64_t* tiles);

/* ---- batches of CA simulations -------------------------------------------------------------
 * No counterpart in the reference (it runs one LavaCA per process; engine/src/app/Batch.cpp:40-120 spawns
 * one engine per config).  Every step launch serves all running members: the work lists hold
 * (simulation, tile) entries of every member, so many small grids fill the GPU together.  Members keep
 * their own state and stop rules; results are identical to running each alone.  A simulation is in at
 * most one batch; while it is, vca_ca_run / vca_ca_enqueue on it fail and vca_batch_* advance it. */
typedef struct vca_batch vca_batch;
int vca_batch_create(vca_ctx* ctx, int32_t max_sims, uint64_t max_tiles, vca_batch** out);
int vca_batch_destroy(vca_batch* b);           /* detaches the remaining members */
/* VCA_ERR_NOMEM when no slot or tile range is free; *slot (may be NULL) = the member's slot */
int vca_batch_add(vca_batch* b, vca_ca* ca, int32_t* slot);
int vca_batch_remove(vca_batch* b, vca_ca* ca);
/* steps_per_graph > 0: step launches are captured once per membership as a hipGraph of that many steps
 * and replayed (steps past a member's stop are no-ops); 0: direct launches. */
int vca_batch_set_graph(vca_batch* b, int32_t steps_per_graph);
int vca_batch_enqueue(vca_batch* b, int32_t steps);
/* Enqueues up to max_steps steps in chunks and returns when no member is running or the steps are done;
 * *running = members still running. */
int vca_batch_run(vca_batch* b, int32_t max_steps, int32_t* running);
int vca_batch_info(vca_batch* b, int32_t* members, uint64_t* tiles_used, int32_t* running);
/* Step count and stop reason of every slot in one device read: out[2*s] = steps, out[2*s+1] = stop reason
 * (-1 for an empty slot); cap = entries of out / 2.  Synchronises the context stream.  A member that has become
 * quiescent (see vca_ca_run) and has the fast-forward on is completed first: members of a batch run to their stop. */
int vca_batch_status(vca_batch* b, int32_t* out, int32_t cap);

/* ---- Potts MC: replaces PottsMC ------------------------------------------------------------ */
typedef struct {
    int32_t nx, ny, nz;
    int32_t neighborhood;
    int32_t periodic[3];         /* effective periodicity per axis (PottsMC.cpp:18) */
    uint64_t seed;               /* McParams (McRules.hpp:80-85) */
    int32_t Q;
    double J, kT;
    int32_t checkerboard;        /* mc.update: 1 checkerboard (colour passes), 0 random (sequential) */
} vca_mc_desc;

typedef struct {
    uint64_t attempts, accepted, uphill;   /* PottsMC::attempts/accepted/acceptedUphill (PottsMC.hpp:104-106) */
    int32_t sweeps;
} vca_mc_counts;

typedef struct {
    int32_t mcs;
    uint64_t grains, cells, largest;        /* Grains::label (Grains.cpp:33-150) */
    double mean_grain_size, energy;         /* PottsMC::measure / totalEnergy (PottsMC.cpp:535-664) */
} vca_mc_stats;

/* Grain counts by mobility class (McStage.cpp:131-150, threshold 0.5). */
typedef struct {
    uint64_t grains, grain_cells, largest;
} vca_mc_class;

/* PottsMC ctor + setDomain/setDomainAll + setMobility + setCheckerboard + initialize
 * (PottsMC.cpp:14-143).  domain: N bytes (NULL = whole space).  mobility: N floats (NULL = none); it
 * feeds the by-mobility grain statistics, and the trial skip as well when mobility_dynamics != 0
 * (McStage.cpp:205-223: mobilityFromCooling).  Mobility and Metropolis probabilities become exact
 * 53-bit integer thresholds on the host, so the device never evaluates exp. */
int vca_mc_create(vca_ctx* ctx, const vca_mc_desc* desc, const uint8_t* domain, const float* mobility,
                  int32_t mobility_dynamics, vca_mc** out);
/* Same, but the basalt domain and the cooling mobility (CoolingMobility.cpp:8-129) are derived on the device from
 * the final CA state held by `ca` (same dims and device, stopped, not in a batch), without the caller downloading
 * it.  ctx may be another context than the CA's. */
int vca_mc_create_from_ca(vca_ctx* ctx, const vca_mc_desc* desc, vca_ca* ca, int32_t use_mobility,
                          double mobility_min, vca_mc** out);
int vca_mc_destroy(vca_mc* mc);
/* PottsMC::sweep x n (PottsMC.cpp:516-533); counts are cumulative since creation.  Checkerboard sweeps run one
 * launch per colour pass unless vca_mc_set_sweep_mode asked for the persistent kernel. */
int vca_mc_sweep(vca_mc* mc, int32_t n, vca_mc_counts* counts);
/* How vca_mc_sweep launches checkerboard sweeps (no counterpart in the reference; for tests and tuning):
 * VCA_MC_SWEEP_AUTO (default: one launch per colour pass), VCA_MC_SWEEP_PER_COLOUR (the same) or
 * VCA_MC_SWEEP_PERSISTENT (all n sweeps of a call in ONE launch of at most one block per compute unit, a device-wide
 * barrier between colour passes; slower than a launch per pass on a whole MI355X, whose eight XCDs each have their
 * own L2).  Same results in every mode.  VCA_ERR_STATE for the random update. */
#define VCA_MC_SWEEP_AUTO 0
#define VCA_MC_SWEEP_PER_COLOUR 1
#define VCA_MC_SWEEP_PERSISTENT 2
int vca_mc_set_sweep_mode(vca_mc* mc, int32_t mode);
int vca_mc_counts_get(vca_mc* mc, vca_mc_counts* counts);
/* PottsMC::measure (PottsMC.cpp:535-553) on the device.  classes (2 entries: high, low) are filled
 * when the handle has mobility data and classes != NULL; high_cells gets the number of domain cells
 * with mobility >= 0.5. */
int vca_mc_measure(vca_mc* mc, vca_mc_stats* stats, vca_mc_class* classes, uint64_t* high_cells);
/* PottsMC::crystals (PottsMC.cpp:666-697): orientation per cell, 0 outside the domain. */
int vca_mc_download(vca_mc* mc, int32_t* crystal);
/* Grain representative per cell (smallest global index of the grain, 0xFFFFFFFF outside). */
int vca_mc_download_labels(vca_mc* mc, uint32_t* root);
int vca_mc_info(vca_mc* mc, int32_t* colors, uint64_t* domain_cells);
/* Device-side mobility of the handle (NULL when off), for host statistics; N floats. */
int vca_mc_download_mobility(vca_mc* mc, float* mobility);

/* ---- timing (benchmark support) ---------------------------------------------------------- */
typedef struct {
    double ms;          /* summed device time of the kernel */
    uint64_t launches;  /* launches timed */
} vca_kernel_time;
/* Enables per-launch HIP-event timing of the kernels of the CA step and the MC sweep: on = 1 times every launch,
 * on = N > 1 the launches of every N-th CA step and every N-th MC sweep (the averages per launch then come from a
 * regular sample of the run at a fraction of the events' own cost), 0 turns timing off.  Events come from a pool
 * and are read back at the synchronisation points the calls already have (no extra host sync). */
int vca_ctx_set_timing(vca_ctx* ctx, int32_t on);
/* name: "ca_front", "ca_descent", "ca_rules", "ca_finish", "ca_step" (whole step), "mc_color" (one colour launch),
 * "mc_sweep" (whole sweep of colour launches), "mc_sweeps" (one persistent sweep launch) */
int vca_ctx_kernel_time(vca_ctx* ctx, const char* name, vca_kernel_time* out);

#ifdef __cplusplus
}
#endif
#endif /* VCA_H */

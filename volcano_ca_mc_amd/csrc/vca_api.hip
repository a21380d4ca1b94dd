// extern "C" layer of libvca.so: handles, device memory, host-built tables, launch orchestration.
// Declarations and the reference interfaces each entry point replaces: include/vca.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstddef>
#include <cstdio>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <new>
#include <stdexcept>
#include <string>
#include <vector>

#include "vca.h"
#include "vca_internal.h"

using namespace vca;

namespace {

thread_local std::string g_last_error;

struct VcaError : std::runtime_error {
    int code;
    VcaError(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

void check(hipError_t e, const char* what) {
    if (e != hipSuccess) throw VcaError(VCA_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

template <class F>
int guarded(F&& f) {
    try {
        f();
        return VCA_OK;
    } catch (const VcaError& e) {
        g_last_error = e.what();
        return e.code;
    } catch (const std::bad_alloc&) {
        g_last_error = "host allocation failed";
        return VCA_ERR_NOMEM;
    } catch (const std::exception& e) {
        g_last_error = e.what();
        return VCA_ERR_ARG;
    }
}

void require(bool ok, const std::string& msg) {
    if (!ok) throw VcaError(VCA_ERR_ARG, msg);
}

// Stream that device allocations of the current API call are ordered on (set by AllocStream for the call).
thread_local hipStream_t t_alloc_stream = nullptr;
struct AllocStream {
    hipStream_t prev;
    explicit AllocStream(hipStream_t s) : prev(t_alloc_stream) { t_alloc_stream = s; }
    ~AllocStream() { t_alloc_stream = prev; }
};

// Device allocation that remembers its size; freed with the owner.  Inside an AllocStream scope it comes from
// the device's stream-ordered pool (hipMallocAsync / hipFreeAsync on that stream): no device-wide
// synchronisation per allocation, and freed blocks are reused by the next job.
struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
    hipStream_t stream = nullptr;
    bool pooled = false;
    void alloc(size_t n) {
        release();
        if (n == 0) n = 1;
        hipError_t e;
        if (t_alloc_stream) {
            e = hipMallocAsync(&p, n, t_alloc_stream);
            stream = t_alloc_stream;
            pooled = e == hipSuccess;
        } else {
            e = hipMalloc(&p, n);
        }
        if (e != hipSuccess) {
            p = nullptr;
            (void)hipGetLastError();
            throw VcaError(VCA_ERR_NOMEM, "hipMalloc of " + std::to_string(n) + " bytes failed: " + hipGetErrorString(e));
        }
        bytes = n;
    }
    void release() {
        if (p) {
            if (pooled)
                (void)hipFreeAsync(p, stream);
            else
                (void)hipFree(p);
        }
        p = nullptr;
        bytes = 0;
        pooled = false;
    }
    ~DevBuf() { release(); }
    template <class T>
    T* as() const {
        return static_cast<T*>(p);
    }
};

struct KernelTime {
    double ms = 0.0;
    uint64_t launches = 0;
};

// Neighbourhood tables from the reference rules (engine/src/core/Neighborhood.cpp:55-119).
int pos(int dx, int dy, int dz) { return (dx + 1) + 3 * (dy + 1) + 9 * (dz + 1); }

bool keep2d(int kind, int var, int u, int v) {
    switch (kind) {
        case VCA_VON_NEUMANN: return std::abs(u) + std::abs(v) == 1;
        case VCA_MOORE: return true;
        case VCA_PENTAGONAL: return var == 0 ? u != -1 : var == 1 ? u != 1 : var == 2 ? v != -1 : v != 1;
        case VCA_HEXAGONAL:
            return var == 0 ? !((u == -1 && v == -1) || (u == 1 && v == 1)) : !((u == -1 && v == 1) || (u == 1 && v == -1));
    }
    return false;
}
bool keep3d(int kind, int var, int dx, int dy, int dz) {
    switch (kind) {
        case VCA_VON_NEUMANN: return std::abs(dx) + std::abs(dy) + std::abs(dz) == 1;
        case VCA_MOORE: return true;
        case VCA_PENTAGONAL: {
            const bool a[6] = {dx != -1, dx != 1, dy != -1, dy != 1, dz != -1, dz != 1};
            return a[var];
        }
    }
    return false;
}

// Returns the variant masks of kind on dims; throws for the unsupported combinations.
std::vector<uint32_t> variant_masks(int kind, int nx, int ny, int nz) {
    const int sizes[3] = {nx, ny, nz};
    const int axes = (nx > 1) + (ny > 1) + (nz > 1);
    require(axes >= 2, "sąsiedztwo wymaga przestrzeni 2D lub 3D");
    require(kind >= 0 && kind <= 3, "nieznane sąsiedztwo");
    std::vector<uint32_t> out;
    if (axes == 2) {
        int plane[2], k = 0;
        for (int a = 0; a < 3; ++a)
            if (sizes[a] > 1) plane[k++] = a;
        const int count = kind == VCA_PENTAGONAL ? 4 : kind == VCA_HEXAGONAL ? 2 : 1;
        for (int var = 0; var < count; ++var) {
            uint32_t m = 0;
            for (int v = -1; v <= 1; ++v)
                for (int u = -1; u <= 1; ++u) {
                    if ((u == 0 && v == 0) || !keep2d(kind, var, u, v)) continue;
                    int d[3] = {0, 0, 0};
                    d[plane[0]] = u;
                    d[plane[1]] = v;
                    m |= 1u << pos(d[0], d[1], d[2]);
                }
            out.push_back(m);
        }
        return out;
    }
    require(kind != VCA_HEXAGONAL, "sąsiedztwo heksagonalne jest zdefiniowane tylko w 2D");
    const int count = kind == VCA_PENTAGONAL ? 6 : 1;
    for (int var = 0; var < count; ++var) {
        uint32_t m = 0;
        for (int dz = -1; dz <= 1; ++dz)
            for (int dy = -1; dy <= 1; ++dy)
                for (int dx = -1; dx <= 1; ++dx)
                    if (!(dx == 0 && dy == 0 && dz == 0) && keep3d(kind, var, dx, dy, dz)) m |= 1u << pos(dx, dy, dz);
        out.push_back(m);
    }
    return out;
}

NbTable make_nb(int kind, int nx, int ny, int nz) {
    NbTable t{};
    const std::vector<uint32_t> m = variant_masks(kind, nx, ny, nz);
    t.variants = static_cast<int32_t>(m.size());
    for (size_t v = 0; v < m.size(); ++v) {
        t.mask[v] = m[v];
        t.any_mask |= m[v];
        t.max_size = std::max(t.max_size, __builtin_popcount(m[v]));
    }
    for (int k = 9; k < 18; ++k)
        if ((t.any_mask >> k) & 1u) t.side_mask |= 1u << k;
    for (int k = 0; k < 9; ++k)
        if (k != 4 && ((t.any_mask >> k) & 1u)) t.diagonal_falls = 1;  // dz == -1 with dx|dy != 0
    t.conn_mask = variant_masks(kind == VCA_VON_NEUMANN ? VCA_VON_NEUMANN : VCA_MOORE, nx, ny, nz)[0];
    t.var_r32 = static_cast<uint32_t>((uint64_t{1} << 32) % static_cast<uint64_t>(t.variants));
    t.var_inv = 1.0 / static_cast<double>(t.variants);
    return t;
}

int default_shape(int nx, int ny, int nz) {
    (void)nx;
    if (ny == 1) return 3;  // x-z plane: 64 x 1 x 16
    if (nz == 1) return 4;  // x-y plane: 64 x 16 x 1
    return 0;               // 3D: 64 x 4 x 4
}

}  // namespace

struct vca_ctx {
    int device = 0;
    int cus = 0;                  // compute units of the device
    uint64_t clock_hz = 0;        // rate of the device's constant-rate clock (wall_clock64)
    hipStream_t stream = nullptr;
    bool own_stream = false;
    bool timing = false;
    // timing period: the launches of every timing_every-th CA step and MC sweep are timed (1 = all of them)
    uint32_t timing_every = 1;
    uint64_t ca_tick = 0, mc_tick = 0;
    bool time_ca_step() { return timing && (ca_tick++ % timing_every) == 0; }
    bool time_mc_sweep() { return timing && (mc_tick++ % timing_every) == 0; }
    std::map<std::string, KernelTime> times;
    // Event pool for per-launch timing: events are recorded around launches without any host wait and
    // drained (elapsed times summed per kernel name) at the next point where the host synchronises anyway.
    std::vector<hipEvent_t> pool;
    size_t used = 0;
    struct Pending {
        const char* name;
        size_t a, b;
    };
    std::vector<Pending> pending;

    void set_device() const { check(hipSetDevice(device), "hipSetDevice"); }
    hipEvent_t take() {
        if (used == pool.size()) {
            hipEvent_t e;
            check(hipEventCreate(&e), "hipEventCreate");
            pool.push_back(e);
        }
        return pool[used++];
    }
    size_t record() {
        hipEvent_t e = take();
        check(hipEventRecord(e, stream), "hipEventRecord");
        return used - 1;
    }
    void mark(const char* name, size_t a, size_t b) { pending.push_back(Pending{name, a, b}); }
    void drain() {
        if (pending.empty()) {
            used = 0;
            return;
        }
        check(hipEventSynchronize(pool[used - 1]), "hipEventSynchronize");
        for (const Pending& p : pending) {
            float ms = 0.0f;
            check(hipEventElapsedTime(&ms, pool[p.a], pool[p.b]), "hipEventElapsedTime");
            KernelTime& k = times[p.name];
            k.ms += ms;
            k.launches += 1;
        }
        pending.clear();
        used = 0;
    }
};

struct Batch;

// Blocks of the tile kernels' launches (each block strides over the step's tile list; the device's list length is not
// known to the host): 32 per CU, at most one per tile the batch can hold.  Measured on MI355X: blocks beyond a small
// front cost little (caldera caSim 44.2 ms at 32 per CU, 43.6 at 16, 45.0 at 8), while fewer blocks lose the dynamic
// balance of a large front (cone 86 ms at 32 per CU, 91 at 16, 98 at 6, 118 at 4).
inline int tile_grid(const vca_ctx* c, uint32_t cap) {
    const uint64_t g = static_cast<uint64_t>(std::max(1, c->cus)) * 32u;
    return static_cast<int>(std::max<uint64_t>(1, std::min<uint64_t>(cap, g)));
}

namespace {
void noalloc_setup(int device);  // vca_ca_run_noalloc's per-device set-up (below)
}  // namespace

// A region of a simulation's single device allocation.
struct Span {
    void* p = nullptr;
    size_t bytes = 0;
    template <class T>
    T* as() const {
        return static_cast<T*>(p);
    }
};

// Offsets of the regions of one allocation, each 256-byte aligned.
struct Layout {
    size_t total = 0;
    size_t add(size_t bytes) {
        const size_t at = total;
        total += (std::max<size_t>(bytes, 1) + 255) & ~static_cast<size_t>(255);
        return at;
    }
};

struct vca_ca {
    vca_ctx* ctx = nullptr;
    CaParams P{};
    CaDev D{};
    int shape = 0;
    DevBuf arena;  // every region below: one allocation and one free per simulation
    Span type, type2, flow, temp, temp2, solid, arrival, descent, spill, vent_cell, tile_flags, tile_near, tile_halo,
        tile_tchg, listed, dirtied, tile_vent, part, ctrl, hist_a, hist_c;
    uint64_t n = 0;
    std::unique_ptr<Batch> own;   // batch of one: the solo path (made on the first solo enqueue)
    bool flow16 = true;           // every flowId the simulation can hold (its vents') fits int16_t
    hipStream_t side = nullptr;   // during vca_ca_collect: the pooled stream its reads go on
    Batch* member = nullptr;      // the shared batch this simulation is in
    int slot = -1;
    // the fast-forward where it applies (vca_ca_set_fast_forward)
    bool ff = true;
    bool hold = false;  // CaCtrl::hold as last set (ca_set_hold)
    // a fast-forward up to the next solidification costs a few launches and a host wait: after one that covered few
    // steps, the next is not tried before step ff_hold (the hold doubles each time, and resets after a long one)
    int32_t ff_hold = 0, ff_backoff = kFfBackoff;
    static constexpr int32_t kFfBackoff = 64, kFfMaxBackoff = 8192, kFfWorthSteps = 32;
    // after a fast-forward from `from` that left the simulation at `to` (running)
    void ff_done(int32_t from, int32_t to) {
        if (to - from >= kFfWorthSteps) {
            ff_backoff = kFfBackoff;
        } else {
            ff_hold = to + ff_backoff;
            ff_backoff = std::min(2 * ff_backoff, kFfMaxBackoff);
        }
    }
    // the type and temperature copies holding the state after `steps` steps
    const void* type_after(int32_t steps) const { return (steps & 1) ? type2.p : type.p; }
    const void* temp_after(int32_t steps) const { return (steps & 1) ? temp2.p : temp.p; }
};

// Entries of the per-device simulation table (kSimTable) handed out to batches as contiguous ranges.  A range
// is given back with an event recorded on the stream of the batch that used it: it is reused only once that
// event has completed (kernels still queued may read the entries), without the releasing thread waiting.
class SimTable {
public:
    static uint32_t take(int device, int n) {
        std::lock_guard<std::mutex> lock(mu());
        std::vector<char>& used = map()[device];
        used.resize(kSimTable, 0);
        for (int pass = 0; pass < 2; ++pass) {
            reap(device, pass == 1);  // second pass: wait for the pending releases
            for (int b = 0; b + n <= kSimTable; ++b) {
                bool ok = true;
                for (int k = 0; k < n && ok; ++k) ok = !used[static_cast<size_t>(b + k)];
                if (!ok) continue;
                for (int k = 0; k < n; ++k) used[static_cast<size_t>(b + k)] = 1;
                return static_cast<uint32_t>(b);
            }
        }
        throw VcaError(VCA_ERR_NOMEM, "simulation table full (" + std::to_string(kSimTable) + " entries per device)");
    }
    // entries become free when `after` completes (nullptr: now); takes ownership of the event
    static void give(int device, uint32_t base, int n, hipEvent_t after) {
        std::lock_guard<std::mutex> lock(mu());
        pending().push_back(Pending{device, base, n, after});
        reap(device, false);
    }

private:
    struct Pending {
        int device;
        uint32_t base;
        int n;
        hipEvent_t ev;
    };
    static void reap(int device, bool wait) {
        std::vector<Pending>& p = pending();
        std::vector<char>& used = map()[device];
        used.resize(kSimTable, 0);
        for (size_t k = 0; k < p.size();) {
            if (p[k].device != device) {
                ++k;
                continue;
            }
            if (p[k].ev) {
                const hipError_t e = wait ? hipEventSynchronize(p[k].ev) : hipEventQuery(p[k].ev);
                if (e == hipErrorNotReady) {
                    ++k;
                    continue;
                }
                (void)hipGetLastError();
                (void)hipEventDestroy(p[k].ev);
            }
            for (int q = 0; q < p[k].n && p[k].base + q < used.size(); ++q) used[p[k].base + q] = 0;
            p[k] = p.back();
            p.pop_back();
        }
    }
    static std::vector<Pending>& pending() {
        static std::vector<Pending> v;
        return v;
    }
    static std::mutex& mu() {
        static std::mutex m;
        return m;
    }
    static std::map<int, std::vector<char>>& map() {
        static std::map<int, std::vector<char>> m;
        return m;
    }
};

// Simulations advanced together (see vca_batch_* in vca.h).  The solo path of a vca_ca is a batch of one.
// Entries each sub-list of a batch holds at most: a simulation of n tiles has at most ceil(n / kSub) tiles of one
// residue, and the batch's simulations share at most `tiles` tiles.
inline uint32_t sub_capacity(uint32_t tiles, int slots) {
    return (tiles + kSub - 1) / kSub + static_cast<uint32_t>(slots) + 1u;
}

struct Batch {
    vca_ctx* ctx = nullptr;
    int max_slots = 0;
    uint32_t cap = 0;                             // tiles
    DevBuf owner, tile_base, active, desc, ctrl, status;
    uint32_t sim_base = 0;
    bool has_table = false;
    std::vector<CaSim> host_sims;                 // what the table entries hold (the copies read from here)
    std::vector<vca_ca*> members;                 // slot -> simulation
    std::vector<std::pair<uint32_t, uint32_t>> free_ranges;  // (first tile, count), sorted
    uint32_t hwm = 0;
    BatchDev B{};
    int grid_tiles = 1;
    int steps_per_graph = 0;
    hipGraphExec_t graph = nullptr;
    int graph_steps = 0;
    int capture_failures = 0;
    int par = 0;                  // parity of the next step (the list copy it reads)
    bool fin_prev = false;        // the last step's finish is still to be done (never true between enqueues)
    SlotMask purge{};             // slots whose simulation left while the pending list copy may hold its entries
    bool purge_any = false;
    bool external = false;        // lists and counts in caller-provided device memory (vca_ca_run_noalloc)
    // the descent lists of the next step are complete: every member's last step ran its rules kernel in this batch
    // (not after a member joined or was fast-forwarded)
    bool desc_ok = false;
    bool desc_ready() const { return desc_ok || B.desc_list == nullptr; }

    void init(vca_ctx* c, int slots, uint32_t tiles) {
        ctx = c;
        max_slots = slots;
        cap = std::max<uint32_t>(tiles, 1u);
        hipStream_t s = ctx->stream;
        sim_base = SimTable::take(ctx->device, slots);
        has_table = true;
        host_sims.assign(static_cast<size_t>(slots), CaSim{});
        for (int k = 0; k < slots; ++k) ca_write_sim(sim_base + static_cast<uint32_t>(k), &host_sims[static_cast<size_t>(k)], s);
        check(hipGetLastError(), "simulation table");
        owner.alloc(static_cast<size_t>(cap) * 4);
        tile_base.alloc(static_cast<size_t>(slots) * 4);
        const uint32_t subcap = sub_capacity(cap, slots);
        active.alloc(static_cast<size_t>(subcap) * kSub * kShapes * 2 * 8);
        // descent lists and, after them, their counts
        const size_t desc_bytes = (static_cast<size_t>(subcap) * kShapes * 2 * 8 + 255) & ~static_cast<size_t>(255);
        desc.alloc(desc_bytes + 2 * kShapes * 16 * 4);
        ctrl.alloc(sizeof(BatchCtrl));
        check(hipMemsetAsync(owner.p, 0xFF, owner.bytes, s), "memset");  // kNoSlot
        check(hipMemsetAsync(ctrl.p, 0, sizeof(BatchCtrl), s), "memset");
        members.assign(static_cast<size_t>(slots), nullptr);
        par = 0;
        fin_prev = false;
        free_ranges.clear();
        hwm = 0;
        B.sim_base = sim_base;
        B.owner = owner.as<uint32_t>();
        B.tile_base = tile_base.as<uint32_t>();
        B.ntiles = 0;
        B.nslots = 0;
        B.subcap = subcap;
        B.active = active.as<uint64_t>();
        B.desc_list = desc.as<uint64_t>();
        B.desc_count = reinterpret_cast<uint32_t*>(static_cast<char*>(desc.p) + desc_bytes);
        check(hipMemsetAsync(B.desc_count, 0, 2 * kShapes * 16 * 4, s), "memset");
        B.ctrl = ctrl.as<BatchCtrl>();
        desc_ok = false;
        grid_tiles = tile_grid(c, cap);
    }
    // A batch of one slot over device memory the library does not allocate: counts, tile owners and the two list copies
    // carved out of `base` (bytes), for as many tiles as fit; the simulation table entry `entry` is the caller's.
    // Nothing is allocated on the device; the host vectors get their final sizes here.
    void init_external(int device, char* base, size_t bytes, uint32_t entry) {
        external = true;
        max_slots = 1;
        sim_base = entry;
        has_table = false;
        Layout L;
        const size_t o_ctrl = L.add(sizeof(BatchCtrl)), o_base = L.add(4), fixed = L.total;
        // per tile: its owner (4 B) and its entries in both copies of every shape's sub-lists (2 x kShapes x 8 B, plus
        // the sub-lists' rounding), each region 256-byte aligned: the most tiles that fit
        auto need = [&](uint32_t c) {
            return fixed + ((static_cast<size_t>(c) * 4 + 255) & ~static_cast<size_t>(255)) +
                   static_cast<size_t>(sub_capacity(c, 1)) * kSub * kShapes * 2 * 8;
        };
        cap = static_cast<uint32_t>(std::min<size_t>((bytes - fixed) / (4 + 2 * kShapes * 8), kTileMask));
        while (cap > 1 && need(cap) > bytes) --cap;
        const uint32_t subcap = sub_capacity(cap, 1);
        const size_t o_owner = L.add(static_cast<size_t>(cap) * 4);
        const size_t o_active = L.add(static_cast<size_t>(subcap) * kSub * kShapes * 2 * 8);
        if (L.total > bytes) throw VcaError(VCA_ERR_ARG, "scratch layout exceeds its array");
        host_sims.assign(1, CaSim{});
        members.assign(1, nullptr);
        bases.assign(1, 0);
        free_ranges.clear();
        free_ranges.reserve(4);
        hwm = 0;
        par = 0;
        fin_prev = false;
        B.sim_base = sim_base;
        B.owner = reinterpret_cast<uint32_t*>(base + o_owner);
        B.tile_base = reinterpret_cast<uint32_t*>(base + o_base);
        B.ntiles = 0;
        B.nslots = 1;
        B.subcap = subcap;
        B.active = reinterpret_cast<uint64_t*>(base + o_active);
        B.desc_list = nullptr;  // no descent lists: the descent checks every listed tile
        B.desc_count = nullptr;
        B.ctrl = reinterpret_cast<BatchCtrl*>(base + o_ctrl);
        grid_tiles = 0;  // set by reset_external (the context's device)
        (void)device;
    }
    // Ready for the next simulation (external batch): counts cleared by a kernel, host state reset.
    void reset_external(vca_ctx* c) {
        ctx = c;
        grid_tiles = tile_grid(c, cap);
        fill_u32(reinterpret_cast<uint32_t*>(B.ctrl), 0u, sizeof(BatchCtrl) / 4, ctx->stream);
        members[0] = nullptr;
        free_ranges.clear();
        hwm = 0;
        par = 0;
        fin_prev = false;
        purge = SlotMask{};
        purge_any = false;
        steps_per_graph = 0;
    }
    ~Batch() {
        drop_graph();
        if (has_table) {
            hipEvent_t ev = nullptr;
            if (hipEventCreateWithFlags(&ev, hipEventDisableTiming) != hipSuccess ||
                hipEventRecord(ev, ctx->stream) != hipSuccess) {
                (void)hipGetLastError();
                if (ev) (void)hipEventDestroy(ev);
                ev = nullptr;
                (void)hipStreamSynchronize(ctx->stream);
            }
            SimTable::give(ctx->device, sim_base, max_slots, ev);
        }
    }
    void drop_graph() {
        if (graph) (void)hipGraphExecDestroy(graph);
        graph = nullptr;
        graph_steps = 0;
    }
    uint32_t shapes() const {
        uint32_t m = 0;
        for (const vca_ca* c : members)
            if (c) m |= 1u << c->shape;
        return m;
    }
    bool any_lava_mode() const {
        for (const vca_ca* c : members)
            if (c && !c->P.classic) return true;
        return false;
    }
    // some member needs the scan of ca_front (classic mode, or every tile visited)
    bool any_scan() const {
        for (const vca_ca* c : members)
            if (c && (c->P.classic || c->P.dense)) return true;
        return false;
    }
    bool flow16() const {
        for (const vca_ca* c : members)
            if (c && !c->flow16) return false;
        return true;
    }
    int count() const {
        int k = 0;
        for (const vca_ca* c : members) k += c != nullptr;
        return k;
    }
    // First-fit tile range; false when none is free.
    bool take_range(uint32_t n, uint32_t& base) {
        for (size_t k = 0; k < free_ranges.size(); ++k) {
            if (free_ranges[k].second < n) continue;
            base = free_ranges[k].first;
            free_ranges[k].first += n;
            free_ranges[k].second -= n;
            if (free_ranges[k].second == 0) free_ranges.erase(free_ranges.begin() + static_cast<long>(k));
            return true;
        }
        if (static_cast<uint64_t>(hwm) + n > cap) return false;
        base = hwm;
        hwm += n;
        return true;
    }
    void give_range(uint32_t base, uint32_t n) {
        free_ranges.emplace_back(base, n);
        std::sort(free_ranges.begin(), free_ranges.end());
        std::vector<std::pair<uint32_t, uint32_t>> merged;
        for (auto& r : free_ranges) {
            if (!merged.empty() && merged.back().first + merged.back().second == r.first)
                merged.back().second += r.second;
            else
                merged.push_back(r);
        }
        free_ranges.swap(merged);
        while (!free_ranges.empty() && free_ranges.back().first + free_ranges.back().second == hwm) {
            hwm = free_ranges.back().first;
            free_ranges.pop_back();
        }
    }
    void refresh() {
        int top = 0;
        for (int k = 0; k < max_slots; ++k)
            if (members[static_cast<size_t>(k)]) top = k + 1;
        B.nslots = static_cast<uint32_t>(std::max(top, 1));
        B.ntiles = hwm;
        drop_graph();
    }
    int attach(vca_ca* ca) {
        int slot = -1;
        for (int k = 0; k < max_slots; ++k)
            if (!members[static_cast<size_t>(k)]) {
                slot = k;
                break;
            }
        if (slot < 0) throw VcaError(VCA_ERR_NOMEM, "batch: no free simulation slot");
        uint32_t base = 0;
        if (!take_range(ca->P.ntiles, base)) throw VcaError(VCA_ERR_NOMEM, "batch: no free tile range");
        hipStream_t s = ctx->stream;
        CaSim& sim = host_sims[static_cast<size_t>(slot)];
        sim = CaSim{};
        sim.P = ca->P;
        sim.D = ca->D;
        sim.running = 1;
        sim.shape = ca->shape;
        sim.push = !ca->P.classic && !ca->P.dense;
        // entries of simulations that left go before the new one is listed (the pending copy never holds more
        // entries than the batch has tiles)
        flush_purge();
        ca_write_sim(sim_base + static_cast<uint32_t>(slot), &sim, s);
        fill_u32(B.tile_base + slot, base, 1, s);
        ca_launch_owner(B.owner, base, ca->P.ntiles, static_cast<uint32_t>(slot), s);
        // a lava-front simulation's first tiles (later steps list their own)
        if (sim.push) ca_launch_seed(B, static_cast<uint32_t>(slot), ca->P.ntiles, par, s);
        check(hipGetLastError(), "batch add");
        desc_ok = false;  // the new member's tiles are in no descent list
        members[static_cast<size_t>(slot)] = ca;
        bases.resize(static_cast<size_t>(max_slots));
        bases[static_cast<size_t>(slot)] = base;
        refresh();
        return slot;
    }
    void detach(vca_ca* ca, int slot) {
        hipStream_t s = ctx->stream;
        host_sims[static_cast<size_t>(slot)].running = 0;
        ca_write_sim(sim_base + static_cast<uint32_t>(slot), &host_sims[static_cast<size_t>(slot)], s);
        const uint32_t base = bases[static_cast<size_t>(slot)];
        ca_launch_owner(B.owner, base, ca->P.ntiles, kNoSlot, s);
        check(hipGetLastError(), "batch remove");
        give_range(base, ca->P.ntiles);
        members[static_cast<size_t>(slot)] = nullptr;
        // the pending list copy may hold entries the simulation pushed for a step that has not run
        purge.bits[static_cast<uint32_t>(slot) >> 5] |= 1u << (static_cast<uint32_t>(slot) & 31u);
        purge_any = true;
        refresh();
    }
    std::vector<uint32_t> bases;

    void flush_purge() {
        if (!purge_any) return;
        ca_launch_purge(B, purge, par, ctx->stream);
        check(hipGetLastError(), "batch purge");
        purge = SlotMask{};
        purge_any = false;
    }
    StepLaunch step_config() const {
        StepLaunch L{};
        L.shapes = shapes();
        L.lava_mode = any_lava_mode();
        L.flow16 = flow16();
        L.scan = any_scan();
        L.par = par;
        L.fin_prev = fin_prev;
        L.desc_list = desc_ok && B.desc_list != nullptr;
        uint64_t tiles[kShapes] = {};
        for (const vca_ca* c : members)
            if (c) tiles[c->shape] += c->P.ntiles;
        // a multiple of kSub blocks (block b walks sub-list b % kSub)
        for (int k = 0; k < kShapes; ++k)
            L.grid[k] = static_cast<int>(
                (std::min<uint64_t>(static_cast<uint64_t>(grid_tiles), tiles[k]) + kSub - 1) / kSub * kSub);
        return L;
    }
    // One step on stream s (no timing); advances the parity and the pending finish.
    void launch_step(hipStream_t s, hipEvent_t* ev, bool* rec) {
        const StepLaunch L = step_config();
        ca_launch_step(B, L, s, ev, rec);
        par ^= 1;
        desc_ok = true;  // this step's rules kernel listed the next step's descent tiles
        fin_prev = !L.scan;  // with the scan, ca_finish ran after the step
    }
    // The pending finish of the last step, if any (every enqueue ends with it: the host reads the stop flags next).
    void close_steps(hipStream_t s, hipEvent_t* ev) {
        if (fin_prev) ca_launch_finish(B, s, ev);
        fin_prev = false;
    }
    void launch_direct(int steps) {
        vca_ctx* c = ctx;
        static const char* kNames[kCaKernels] = {"ca_front", "ca_descent", "ca_rules", "ca_finish"};
        for (int k = 0; k < steps; ++k) {
            if (!c->time_ca_step()) {
                launch_step(c->stream, nullptr, nullptr);
                continue;
            }
            hipEvent_t ev[2 * kCaKernels];
            size_t idx[2 * kCaKernels];
            for (int e = 0; e < 2 * kCaKernels; ++e) {
                ev[e] = c->take();
                idx[e] = c->used - 1;
            }
            bool rec[kCaKernels] = {};
            launch_step(c->stream, ev, rec);
            int first = -1, last = -1;
            for (int q = 0; q < kCaKernels; ++q) {
                if (!rec[q]) continue;
                c->mark(kNames[q], idx[2 * q], idx[2 * q + 1]);
                if (first < 0) first = q;
                last = q;
            }
            if (first >= 0) c->mark("ca_step", idx[2 * first], idx[2 * last + 1]);
            if (c->used > 2 * kCaKernels * 256) c->drain();  // bound the pool on long enqueues
        }
        if (fin_prev && c->timing) {
            hipEvent_t ev[2] = {c->take(), c->take()};
            const size_t a = c->used - 2;
            close_steps(c->stream, ev);
            c->mark("ca_finish", a, a + 1);
        } else {
            close_steps(c->stream, nullptr);
        }
        check(hipGetLastError(), "ca launch");
    }
    // Captures steps_per_graph steps (an even count: the graph starts and ends on parity 0) and the closing finish
    // of the current membership (no memset or host work inside).
    bool capture() {
        hipStream_t s = ctx->stream;
        if (s == nullptr) return false;  // the legacy stream cannot be captured
        hipGraph_t g = nullptr;
        const int par0 = par;
        const bool fin0 = fin_prev, desc0 = desc_ok;
        // a graph starts on parity 0, after a step that filled the descent lists (see enqueue)
        par = 0;
        fin_prev = false;
        desc_ok = true;
        if (hipStreamBeginCapture(s, hipStreamCaptureModeRelaxed) != hipSuccess) {
            (void)hipGetLastError();
            par = par0;
            fin_prev = fin0;
            desc_ok = desc0;
            return false;
        }
        for (int k = 0; k < steps_per_graph; ++k) launch_step(s, nullptr, nullptr);
        close_steps(s, nullptr);
        par = par0;
        fin_prev = fin0;
        desc_ok = desc0;
        // a capture invalidated by the runtime leaves nothing enqueued: the caller launches the steps directly
        if (hipStreamEndCapture(s, &g) != hipSuccess || g == nullptr) {
            (void)hipGetLastError();
            if (g) (void)hipGraphDestroy(g);
            return false;
        }
        hipGraphExec_t exec = nullptr;
        const hipError_t e = hipGraphInstantiate(&exec, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        if (e != hipSuccess) {
            (void)hipGetLastError();
            return false;
        }
        graph = exec;
        graph_steps = steps_per_graph;
        return true;
    }
    void enqueue(int steps) {
        if (steps <= 0 || count() == 0) return;
        flush_purge();
        if (steps_per_graph > 0 && !ctx->timing) {
            if (!graph && !capture()) ++capture_failures;
            if (capture_failures >= 3) steps_per_graph = 0;  // this process cannot capture: direct launches
            // a graph starts on parity 0, after a step whose rules kernel filled the descent lists
            while (graph && steps >= graph_steps + 1 && (par != 0 || !desc_ready())) {
                launch_direct(1);
                --steps;
            }
            while (graph && par == 0 && desc_ready() && steps >= graph_steps) {
                check(hipGraphLaunch(graph, ctx->stream), "hipGraphLaunch");
                steps -= graph_steps;
                desc_ok = true;
            }
        }
        launch_direct(steps);
    }
};

struct vca_mc {
    vca_ctx* ctx = nullptr;
    McParams P{};
    int spin_bytes = 4;
    bool checkerboard = true;
    double J = 1.0, kT = 0.0;
    uint64_t n = 0;
    uint32_t ndomain = 0;
    int colors = 0;
    std::vector<uint32_t> color_start;
    int sweeps = 0;
    bool has_mob_dyn = false;
    bool has_mob_cells = false;
    int sweep_mode = VCA_MC_SWEEP_AUTO;
    DevBuf spin, list, mob_thr, mob_cells, accept, counters, parent, size, hi, gout, bar;
};

// ===================================================================================================
// library / context
// ===================================================================================================
extern "C" int vca_version(void) { return 1; }

extern "C" int vca_last_error(char* buf, size_t cap) {
    if (buf && cap > 0) {
        std::strncpy(buf, g_last_error.c_str(), cap - 1);
        buf[cap - 1] = '\0';
    }
    return VCA_OK;
}

extern "C" int vca_ctx_create(int device, void* stream, vca_ctx** out) {
    return guarded([&] {
        require(out != nullptr, "out == NULL");
        int count = 0;
        if (hipGetDeviceCount(&count) != hipSuccess || count == 0)
            throw VcaError(VCA_ERR_DEVICE, "no HIP device available");
        require(device >= 0 && device < count, "device index out of range");
        hipDeviceProp_t prop;
        check(hipGetDeviceProperties(&prop, device), "hipGetDeviceProperties");
        if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0)
            throw VcaError(VCA_ERR_DEVICE, std::string("libvca is built for gfx950, device is ") + prop.gcnArchName);
        check(hipSetDevice(device), "hipSetDevice");
        noalloc_setup(device);
        auto* c = new vca_ctx();
        c->device = device;
        c->cus = prop.multiProcessorCount;
        int khz = 0;
        check(hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, device), "wall clock rate");
        c->clock_hz = static_cast<uint64_t>(khz > 0 ? khz : 100000) * 1000u;
        c->set_device();
        if (stream) {
            c->stream = static_cast<hipStream_t>(stream);
        } else {
            check(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking), "hipStreamCreate");
            c->own_stream = true;
        }
        // keep freed blocks of the stream-ordered pool for reuse instead of returning them at each sync
        hipMemPool_t pool = nullptr;
        if (hipDeviceGetDefaultMemPool(&pool, device) == hipSuccess && pool) {
            uint64_t keep = ~0ULL;
            (void)hipMemPoolSetAttribute(pool, hipMemPoolAttrReleaseThreshold, &keep);
        }
        (void)hipGetLastError();
        *out = c;
    });
}

extern "C" int vca_ctx_destroy(vca_ctx* ctx) {
    return guarded([&] {
        if (!ctx) return;
        ctx->set_device();
        (void)hipStreamSynchronize(ctx->stream);
        for (auto& e : ctx->pool) (void)hipEventDestroy(e);
        if (ctx->own_stream) (void)hipStreamDestroy(ctx->stream);
        delete ctx;
    });
}

extern "C" int vca_ctx_synchronize(vca_ctx* ctx) {
    return guarded([&] {
        require(ctx, "ctx == NULL");
        check(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
    });
}

extern "C" int vca_device_memory(vca_ctx* ctx, uint64_t* free_bytes, uint64_t* total_bytes) {
    return guarded([&] {
        require(ctx, "ctx == NULL");
        ctx->set_device();
        size_t f = 0, t = 0;
        check(hipMemGetInfo(&f, &t), "hipMemGetInfo");
        if (free_bytes) *free_bytes = f;
        if (total_bytes) *total_bytes = t;
    });
}

extern "C" int vca_ctx_set_timing(vca_ctx* ctx, int32_t on) {
    return guarded([&] {
        require(ctx, "ctx == NULL");
        ctx->set_device();
        ctx->drain();
        require(on >= 0, "invalid argument");
        ctx->timing = on != 0;
        ctx->timing_every = on > 1 ? static_cast<uint32_t>(on) : 1u;
        ctx->ca_tick = ctx->mc_tick = 0;
        if (ctx->timing) ctx->times.clear();
    });
}

extern "C" int vca_ctx_kernel_time(vca_ctx* ctx, const char* name, vca_kernel_time* out) {
    return guarded([&] {
        require(ctx && name && out, "NULL argument");
        ctx->set_device();
        ctx->drain();
        auto it = ctx->times.find(name);
        out->ms = it == ctx->times.end() ? 0.0 : it->second.ms;
        out->launches = it == ctx->times.end() ? 0 : it->second.launches;
    });
}

// ===================================================================================================
// CA
// ===================================================================================================
namespace {
// Regions of a simulation's single allocation (creation and the footprint estimate use the same plan).  The
// regions cleared at creation come first, so that one memset covers them.  Creation inputs (rock heights, seed
// cells) are staged in the D(n) array when they fit (it is set to "none" after them), else in a region of their
// own.
struct CaPlan {
    Layout L;
    size_t hcap = 1;
    size_t type = 0, type2 = 0, flow = 0, temp = 0, temp2 = 0, solid = 0, arrival = 0, ctrl = 0, tile_near = 0,
           tile_halo = 0, tile_tchg = 0, listed = 0, dirtied = 0, tile_vent = 0, part = 0, hist_a = 0, hist_c = 0;
    size_t zero_bytes = 0;
    size_t descent = 0, vent_cell = 0, tile_flags = 0, spill = 0;
    size_t staging = 0, st_height = 0, st_cells = 0, st_flows = 0, st_temps = 0, st_bytes = 0;
};

CaPlan ca_plan(uint64_t n, uint32_t ntiles, size_t cols, bool lava, size_t vent_slots, size_t nvents, int max_steps) {
    CaPlan c;
    Layout& L = c.L;
    c.hcap = static_cast<size_t>(std::max(1, max_steps));
    c.type = L.add(n);
    c.type2 = L.add(n);
    c.flow = L.add(n * 4);
    c.temp = L.add(n * 4);
    c.temp2 = L.add(n * 4);
    c.solid = L.add(n * 2);
    c.arrival = L.add(n * 2);
    c.ctrl = L.add(sizeof(CaCtrl));
    c.tile_near = L.add(ntiles);
    c.tile_halo = L.add(static_cast<size_t>(ntiles) * kHaloBytes);
    c.tile_tchg = L.add(static_cast<size_t>(ntiles) * 4);
    c.listed = L.add(static_cast<size_t>(ntiles) * 4);
    c.dirtied = L.add(static_cast<size_t>(ntiles) * 4);
    c.tile_vent = L.add(ntiles);
    c.part = L.add(static_cast<size_t>(kPartRows) * sizeof(CaPart));
    c.hist_a = L.add(c.hcap * 8);
    c.hist_c = L.add(c.hcap * 8);
    c.zero_bytes = L.total;
    c.descent = L.add(n);
    c.vent_cell = L.add(vent_slots * 4);
    c.tile_flags = L.add(ntiles);
    if (lava) c.spill = L.add(cols * 4);
    Layout st;
    c.st_height = st.add(lava ? cols * 4 : 0);
    c.st_cells = st.add(nvents * 4);
    c.st_flows = st.add(nvents * 4);
    c.st_temps = st.add(nvents * 4);
    c.st_bytes = st.total;
    c.staging = st.total <= n ? c.descent : L.add(st.total);
    return c;
}
}  // namespace

namespace {
// Streams for work outside a context's stream (creation, reads of stopped simulations), reused per device
// (creating and destroying a stream per job costs ~ms).
class StreamPool {
public:
    static hipStream_t take(int device) {
        {
            std::lock_guard<std::mutex> lock(mu());
            std::vector<hipStream_t>& v = free_list()[device];
            if (!v.empty()) {
                hipStream_t s = v.back();
                v.pop_back();
                return s;
            }
        }
        hipStream_t s = nullptr;
        check(hipStreamCreateWithFlags(&s, hipStreamNonBlocking), "hipStreamCreate");
        return s;
    }
    static void give(int device, hipStream_t s) {
        std::lock_guard<std::mutex> lock(mu());
        free_list()[device].push_back(s);
    }

private:
    static std::mutex& mu() {
        static std::mutex m;
        return m;
    }
    static std::map<int, std::vector<hipStream_t>>& free_list() {
        static std::map<int, std::vector<hipStream_t>> v;
        return v;
    }
};

// A pooled stream of the context's device, for work that must not go on the context's stream (which may be
// advancing a batch, or under capture, on another thread meanwhile); synchronised and returned on scope exit.
struct SideStream {
    int device;
    hipStream_t s;
    explicit SideStream(const vca_ctx* ctx) : device(ctx->device), s(StreamPool::take(ctx->device)) {}
    SideStream(const SideStream&) = delete;
    SideStream& operator=(const SideStream&) = delete;
    ~SideStream() {
        (void)hipStreamSynchronize(s);
        StreamPool::give(device, s);
    }
};

}  // namespace

extern "C" int vca_ca_create(vca_ctx* ctx, const vca_ca_desc* d, const int32_t* height, const int32_t* spill,
                             const vca_vent* vents, int32_t nvents, vca_ca** out) {
    return guarded([&] {
        require(ctx && d && out, "NULL argument");
        require(d->nx >= 1 && d->ny >= 1 && d->nz >= 1, "wymiary muszą być >= 1");
        const uint64_t n = static_cast<uint64_t>(d->nx) * d->ny * d->nz;
        require(n <= 2000000000ULL, "przestrzeń ma za dużo komórek (maksimum 2 000 000 000)");
        require(d->max_steps >= 0 && d->max_steps <= 65535, "ca.maxSteps poza zakresem 0..65535");
        require(nvents >= 0 && (nvents == 0 || vents != nullptr), "vents == NULL");
        const bool classic = d->classic != 0;
        if (!classic) {
            require(height != nullptr && spill != nullptr, "tryb lava wymaga mapy wysokości i poziomów przelewu");
            require(!(((d->nx > 1) + (d->ny > 1) + (d->nz > 1)) == 2 && d->ny != 1),
                    "w trybie lava przestrzeń 2D to przekrój x-z (ny = 1)");
        }
        ctx->set_device();
        auto ca = std::make_unique<vca_ca>();
        ca->ctx = ctx;
        ca->n = n;
        CaParams& P = ca->P;
        P.nx = d->nx;
        P.ny = d->ny;
        P.nz = d->nz;
        P.classic = classic;
        P.per[0] = d->periodic && d->nx > 1;
        P.per[1] = d->periodic && d->ny > 1;
        P.per[2] = d->periodic && classic && d->nz > 1;  // Boundary.cpp:5-11
        P.seed = d->seed;
        P.max_steps = d->max_steps;
        P.eruption_steps = d->eruption_steps;
        P.erupt_temp = d->erupt_temp;
        P.solidus = d->solidus_temp;
        P.cool_amb = d->cooling_ambient;
        P.cool_lat = d->cooling_lateral;
        P.cool_fall = d->cooling_fall;
        P.ponding = d->ponding != 0;
        P.dense = d->dense != 0;
        P.nb = make_nb(d->neighborhood, d->nx, d->ny, d->nz);
        require(d->tile_shape >= 0 && d->tile_shape <= kShapes, "tile_shape poza zakresem 0..5");
        ca->shape = d->tile_shape > 0 ? d->tile_shape - 1 : default_shape(d->nx, d->ny, d->nz);
        P.ty = kShapeTY[ca->shape];
        P.tz = kShapeTZ[ca->shape];
        P.ntx = (d->nx + kTileX - 1) / kTileX;
        P.nty = (d->ny + P.ty - 1) / P.ty;
        P.ntz = (d->nz + P.tz - 1) / P.tz;
        const uint64_t ntiles = static_cast<uint64_t>(P.ntx) * P.nty * P.ntz;
        require(ntiles <= kTileMask, "za dużo kafli dla tej siatki");
        P.ntiles = static_cast<uint32_t>(ntiles);
        magic_div(static_cast<uint32_t>(P.ntx), P.div_ntx_m, P.div_ntx_p);
        magic_div(static_cast<uint32_t>(P.nty), P.div_nty_m, P.div_nty_p);

        // vents: flowId -> cell (isVent, LavaCA.hpp:335-341)
        int32_t maxf = 0;
        std::vector<uint32_t> vcell(static_cast<size_t>(nvents));
        std::vector<int32_t> vflow(static_cast<size_t>(nvents));
        std::vector<float> vtemp(static_cast<size_t>(nvents), classic ? 0.0f : d->erupt_temp);
        for (int k = 0; k < nvents; ++k) {
            const vca_vent& v = vents[k];
            // cells take their flowId from a vent (seed) or from a neighbour's vote: the vents' ids are every
            // flowId the simulation can hold
            ca->flow16 = ca->flow16 && v.flow_id >= INT16_MIN && v.flow_id <= INT16_MAX;
            require(v.x >= 0 && v.x < d->nx && v.y >= 0 && v.y < d->ny && v.z >= 0 && v.z < d->nz,
                    "krater poza siatką");
            vcell[k] = static_cast<uint32_t>(v.x + static_cast<uint64_t>(d->nx) * (v.y + static_cast<uint64_t>(d->ny) * v.z));
            vflow[k] = v.flow_id;
            maxf = std::max(maxf, v.flow_id);
        }
        std::vector<uint32_t> vent_table(static_cast<size_t>(maxf) + 1, 0xFFFFFFFFu);
        if (!classic)
            for (int k = 0; k < nvents; ++k)
                if (vflow[k] >= 0) vent_table[static_cast<size_t>(vflow[k])] = vcell[k];
        P.nvent_slots = maxf + 1;

        // Creation runs on a pooled stream of its own: the context's stream may be advancing (or capturing) a batch
        // on another thread meanwhile.  One allocation holds every region.
        SideStream side(ctx);
        hipStream_t s = side.s;
        const size_t cols = static_cast<size_t>(d->nx) * d->ny;
        const CaPlan plan = ca_plan(n, P.ntiles, cols, !classic, vent_table.size(), static_cast<size_t>(nvents), d->max_steps);
        {
            AllocStream alloc_on(s);
            ca->arena.alloc(plan.L.total);
        }
        // freed on the context's stream, after the simulation's last queued step (creation is synchronised below)
        ca->arena.stream = ctx->stream;
        char* base = static_cast<char*>(ca->arena.p);
        auto region = [&](Span& sp, size_t off, size_t bytes) {
            sp.p = base + off;
            sp.bytes = bytes;
        };
        region(ca->type, plan.type, n);
        region(ca->type2, plan.type2, n);
        region(ca->flow, plan.flow, n * 4);
        region(ca->temp, plan.temp, n * 4);
        region(ca->temp2, plan.temp2, n * 4);
        region(ca->solid, plan.solid, n * 2);
        region(ca->arrival, plan.arrival, n * 2);
        region(ca->ctrl, plan.ctrl, sizeof(CaCtrl));
        region(ca->tile_near, plan.tile_near, P.ntiles);
        region(ca->tile_halo, plan.tile_halo, static_cast<size_t>(P.ntiles) * kHaloBytes);
        region(ca->tile_tchg, plan.tile_tchg, static_cast<size_t>(P.ntiles) * 4);
        region(ca->listed, plan.listed, static_cast<size_t>(P.ntiles) * 4);
        region(ca->dirtied, plan.dirtied, static_cast<size_t>(P.ntiles) * 4);
        region(ca->tile_vent, plan.tile_vent, P.ntiles);
        region(ca->part, plan.part, static_cast<size_t>(kPartRows) * sizeof(CaPart));
        region(ca->hist_a, plan.hist_a, plan.hcap * 8);
        region(ca->hist_c, plan.hist_c, plan.hcap * 8);
        region(ca->descent, plan.descent, n);
        region(ca->vent_cell, plan.vent_cell, vent_table.size() * 4);
        region(ca->tile_flags, plan.tile_flags, P.ntiles);
        if (!classic) region(ca->spill, plan.spill, cols * 4);
        // state, counters, neighbour marks, partials and histories start at zero; descent at "none"
        check(hipMemsetAsync(base, 0, plan.zero_bytes, s), "memset");
        check(hipMemcpyAsync(ca->vent_cell.p, vent_table.data(), vent_table.size() * 4, hipMemcpyHostToDevice, s),
              "copy vents");

        CaDev& D = ca->D;
        D.type[0] = ca->type.as<uint8_t>();
        D.type[1] = ca->type2.as<uint8_t>();
        D.flow = ca->flow.as<int32_t>();
        D.temp[0] = ca->temp.as<float>();
        D.temp[1] = ca->temp2.as<float>();
        D.solid = ca->solid.as<uint16_t>();
        D.arrival = ca->arrival.as<uint16_t>();
        D.descent = ca->descent.as<int8_t>();
        D.vent_cell = ca->vent_cell.as<uint32_t>();
        D.tile_flags = ca->tile_flags.as<uint8_t>();
        D.tile_near = ca->tile_near.as<uint8_t>();
        D.tile_halo = ca->tile_halo.as<uint8_t>();
        D.tile_tchg = ca->tile_tchg.as<uint32_t>();
        D.listed = ca->listed.as<uint32_t>();
        D.dirtied = ca->dirtied.as<uint32_t>();
        D.tile_vent = ca->tile_vent.as<uint8_t>();
        D.part = ca->part.as<CaPart>();
        D.ctrl = ca->ctrl.as<CaCtrl>();
        D.hist_active = ca->hist_a.as<unsigned long long>();
        D.hist_changed = ca->hist_c.as<unsigned long long>();
        D.spill = nullptr;
        ca_launch_vent_tiles(P, D, s);  // after the vent table's upload, on the same stream
        char* staging = base + plan.staging;
        if (!classic) {
            D.spill = ca->spill.as<int32_t>();
            check(hipMemcpyAsync(ca->spill.p, spill, cols * 4, hipMemcpyHostToDevice, s), "copy spill");
            check(hipMemcpyAsync(staging + plan.st_height, height, cols * 4, hipMemcpyHostToDevice, s), "copy height");
            ca_launch_terrain(P, D, reinterpret_cast<const int32_t*>(staging + plan.st_height), s);
            check(hipGetLastError(), "terrain");
        }
        std::vector<char> seeds;  // kept until the synchronisation below (the upload may read it until then)
        if (nvents > 0) {
            // cells, flows and temperatures in one upload
            seeds.assign(plan.st_bytes - plan.st_cells, 0);
            std::memcpy(seeds.data(), vcell.data(), vcell.size() * 4);
            std::memcpy(seeds.data() + (plan.st_flows - plan.st_cells), vflow.data(), vflow.size() * 4);
            std::memcpy(seeds.data() + (plan.st_temps - plan.st_cells), vtemp.data(), vtemp.size() * 4);
            check(hipMemcpyAsync(staging + plan.st_cells, seeds.data(), seeds.size(), hipMemcpyHostToDevice, s), "copy");
            ca_launch_seeds(P, D, reinterpret_cast<const uint32_t*>(staging + plan.st_cells),
                            reinterpret_cast<const int32_t*>(staging + plan.st_flows),
                            reinterpret_cast<const float*>(staging + plan.st_temps), nvents, s);
            check(hipGetLastError(), "seeds");
        }
        // the staged inputs are read: D(n) starts at "none"
        check(hipMemsetAsync(ca->descent.p, 0xFF, n, s), "memset");
        // countInitialState (LavaCA.hpp:1252-1280) and the initial stop rule, on the device
        ca_launch_init_flags(P, D, s);
        ca_launch_init_stop(P, D, s);
        check(hipGetLastError(), "init");
        check(hipStreamSynchronize(s), "init");
        *out = ca.release();
    });
}

extern "C" int vca_ca_destroy(vca_ca* ca) {
    return guarded([&] {
        if (!ca) return;
        ca->ctx->set_device();
        // no host wait: the device buffers are freed in stream order after the work already queued on the
        // context's stream (stream-ordered pool), and the simulation table entries are reused only after it
        if (ca->member) ca->member->detach(ca, ca->slot);
        ca->member = nullptr;
        delete ca;
    });
}

namespace {
// Pinned staging buffers shared by every collecting thread (allocated once, reused).
class PinnedPool {
public:
    static constexpr size_t kBytes = 32u << 20;
    static void* take() {
        {
            std::lock_guard<std::mutex> lock(mu());
            if (!free_list().empty()) {
                void* p = free_list().back();
                free_list().pop_back();
                return p;
            }
        }
        void* p = nullptr;
        check(hipHostMalloc(&p, kBytes, hipHostMallocDefault), "hipHostMalloc");
        return p;
    }
    static void give(void* p) {
        std::lock_guard<std::mutex> lock(mu());
        free_list().push_back(p);
    }

private:
    static std::mutex& mu() {
        static std::mutex m;
        return m;
    }
    static std::vector<void*>& free_list() {
        static std::vector<void*> v;
        return v;
    }
};

// Device -> host copy of `bytes` through two pinned buffers on stream s (copy of one overlaps the host copy
// of the other).
void staged_download(void* dst, const void* src, size_t bytes, hipStream_t s) {
    if (!dst || bytes == 0) return;
    void* buf[2] = {PinnedPool::take(), PinnedPool::take()};
    hipEvent_t ev[2];
    check(hipEventCreateWithFlags(&ev[0], hipEventDisableTiming), "hipEventCreate");
    check(hipEventCreateWithFlags(&ev[1], hipEventDisableTiming), "hipEventCreate");
    size_t pending[2] = {0, 0}, at[2] = {0, 0};
    size_t off = 0;
    int k = 0;
    try {
        while (off < bytes || pending[0] || pending[1]) {
            if (pending[k]) {
                check(hipEventSynchronize(ev[k]), "download");
                std::memcpy(static_cast<char*>(dst) + at[k], buf[k], pending[k]);
                pending[k] = 0;
            }
            if (off < bytes) {
                const size_t n = std::min(PinnedPool::kBytes, bytes - off);
                check(hipMemcpyAsync(buf[k], static_cast<const char*>(src) + off, n, hipMemcpyDeviceToHost, s),
                      "download");
                check(hipEventRecord(ev[k], s), "download");
                pending[k] = n;
                at[k] = off;
                off += n;
            }
            k ^= 1;
        }
    } catch (...) {
        (void)hipStreamSynchronize(s);
        (void)hipEventDestroy(ev[0]);
        (void)hipEventDestroy(ev[1]);
        PinnedPool::give(buf[0]);
        PinnedPool::give(buf[1]);
        throw;
    }
    (void)hipEventDestroy(ev[0]);
    (void)hipEventDestroy(ev[1]);
    PinnedPool::give(buf[0]);
    PinnedPool::give(buf[1]);
}
}  // namespace

extern "C" int vca_ca_collect(vca_ca* ca, vca_ca_result* r) {
    return guarded([&] {
        require(ca && r, "NULL argument");
        if (ca->member) throw VcaError(VCA_ERR_STATE, "the simulation is in a batch: remove it first");
        ca->ctx->set_device();
        SideStream side(ca->ctx);
        ca->side = side.s;
        CaCtrl c;
        check(hipMemcpyAsync(&c, ca->ctrl.p, sizeof(CaCtrl), hipMemcpyDeviceToHost, ca->side), "read ctrl");
        check(hipStreamSynchronize(ca->side), "read ctrl");
        if (c.stop == VCA_RUNNING || c.stop == kStopHeld)
            throw VcaError(VCA_ERR_STATE, "the simulation is still running");
        r->steps = c.steps;
        r->stop_reason = c.stop;
        r->tiles_visited = c.tiles_visited;
        r->tiles_staged = c.tiles_staged;
        r->changes = c.changes_total;
        r->ff_from = c.ff_steps > 0 ? c.ff_from : -1;
        r->ff_steps = c.ff_steps;
        const size_t m = static_cast<size_t>(std::min(c.steps, std::max(0, r->cap)));
        if (m > 0 && r->active) staged_download(r->active, ca->hist_a.p, m * 8, ca->side);
        if (m > 0 && r->changed) staged_download(r->changed, ca->hist_c.p, m * 8, ca->side);
        const uint64_t n = ca->n;
        staged_download(r->type, ca->type_after(c.steps), n, ca->side);
        staged_download(r->flow, ca->flow.p, n * 4, ca->side);
        if (r->temp) ca_launch_final_temps(ca->D, n, c.steps & 1, ca->side);
        staged_download(r->temp, ca->temp_after(c.steps), n * 4, ca->side);
        staged_download(r->solid, ca->solid.p, n * 2, ca->side);
        staged_download(r->arrival, ca->arrival.p, n * 2, ca->side);
        check(hipStreamSynchronize(ca->side), "collect");
        ca->side = nullptr;
    });
}

namespace {
// Returns the step count (its parity selects the type and temperature copies holding the state).
int32_t require_stopped(vca_ca* ca, hipStream_t s) {
    if (ca->member) throw VcaError(VCA_ERR_STATE, "the simulation is in a batch: remove it first");
    CaCtrl c;
    check(hipMemcpyAsync(&c, ca->ctrl.p, sizeof(CaCtrl), hipMemcpyDeviceToHost, s), "read ctrl");
    check(hipStreamSynchronize(s), "read ctrl");
    if (c.stop == VCA_RUNNING || c.stop == kStopHeld)
        throw VcaError(VCA_ERR_STATE, "the simulation is still running");
    return c.steps;
}
}  // namespace

extern "C" int vca_ca_flow_stats(vca_ca* ca, const vca_vent* vents, int32_t nflows, uint64_t* lava, uint64_t* basalt,
                                 uint64_t* maxd2) {
    return guarded([&] {
        require(ca && nflows >= 0, "invalid argument");
        require(nflows == 0 || (vents && lava && basalt && maxd2), "NULL argument");
        ca->ctx->set_device();
        SideStream side(ca->ctx);
        const int par = require_stopped(ca, side.s) & 1;
        const size_t m = static_cast<size_t>(nflows) + 1;
        std::vector<int32_t> xy(2 * m, 0);
        for (int32_t f = 0; f < nflows; ++f) {
            xy[2 * static_cast<size_t>(f)] = vents[f].x;
            xy[2 * static_cast<size_t>(f) + 1] = vents[f].y;
        }
        AllocStream alloc_on(side.s);
        DevBuf dxy, out;
        dxy.alloc(xy.size() * 4);
        out.alloc(3 * m * 8);
        check(hipMemcpyAsync(dxy.p, xy.data(), xy.size() * 4, hipMemcpyHostToDevice, side.s), "copy vents");
        check(hipMemsetAsync(out.p, 0, 3 * m * 8, side.s), "memset");
        unsigned long long* o = out.as<unsigned long long>();
        if (nflows > 0) ca_launch_flow_sums(ca->P, ca->D, par, dxy.as<int32_t>(), nflows, o, o + m, o + 2 * m, side.s);
        check(hipGetLastError(), "flow sums");
        std::vector<uint64_t> h(3 * m);
        check(hipMemcpyAsync(h.data(), out.p, 3 * m * 8, hipMemcpyDeviceToHost, side.s), "read flow sums");
        check(hipStreamSynchronize(side.s), "flow sums");
        for (size_t f = 0; f < m && nflows > 0; ++f) {
            lava[f] = h[f];
            basalt[f] = h[m + f];
            maxd2[f] = h[2 * m + f];
        }
    });
}

extern "C" int vca_ca_flow_points(vca_ca* ca, const vca_vent* vent, int32_t flow, uint64_t d2, uint8_t* seen,
                                  int32_t cap) {
    return guarded([&] {
        require(ca && vent && seen && cap > 0, "invalid argument");
        ca->ctx->set_device();
        SideStream side(ca->ctx);
        const int par = require_stopped(ca, side.s) & 1;
        AllocStream alloc_on(side.s);
        DevBuf buf;
        buf.alloc(static_cast<size_t>(cap) * 4);
        check(hipMemsetAsync(buf.p, 0, static_cast<size_t>(cap) * 4, side.s), "memset");
        ca_launch_flow_points(ca->P, ca->D, par, flow, vent->x, vent->y, d2, buf.as<uint32_t>(), static_cast<uint32_t>(cap),
                              side.s);
        check(hipGetLastError(), "flow points");
        std::vector<uint32_t> h(static_cast<size_t>(cap));
        check(hipMemcpyAsync(h.data(), buf.p, h.size() * 4, hipMemcpyDeviceToHost, side.s), "read flow points");
        check(hipStreamSynchronize(side.s), "flow points");
        for (int32_t k = 0; k < cap; ++k) seen[k] = h[static_cast<size_t>(k)] ? 1 : 0;
    });
}

// Status of a simulation on the context's stream: step count, stop reason, and when the fast-forward may run.
// A held simulation (CaCtrl::hold) reads as running.
struct CaStatus {
    int32_t steps = 0, stop = 0, quiet = 0, qdead = 0;
    bool held = false;
};
static CaStatus ca_status_read(vca_ca* ca) {
    CaCtrl c;
    check(hipMemcpyAsync(&c, ca->ctrl.p, sizeof(CaCtrl), hipMemcpyDeviceToHost, ca->ctx->stream), "read ctrl");
    check(hipStreamSynchronize(ca->ctx->stream), "ca step");
    ca->ctx->drain();
    const bool held = c.stop == kStopHeld;
    return CaStatus{c.steps, held ? VCA_RUNNING : c.stop, c.quiet, c.qdead, held};
}
// Whether the simulation runs with its hold on where a host fast-forward follows its steps (lava front, the
// fast-forward on).
static bool ca_holds(const vca_ca* ca) { return ca->ff && !ca->P.classic && !ca->P.dense; }
// Switches the simulation's hold (CaCtrl::hold) on or off, in stream order; off releases a held simulation.
static void ca_set_hold(vca_ca* ca, bool on) {
    if (ca->hold == on) return;
    ca_launch_hold(ca->ctrl.as<CaCtrl>(), on ? 1 : 0, ca->ctx->stream);
    check(hipGetLastError(), "hold");
    ca->hold = on;
}
static void ca_status_sync(vca_ca* ca, int32_t* steps, int32_t* stop) {
    const CaStatus st = ca_status_read(ca);
    if (steps) *steps = st.steps;
    if (stop) *stop = st.stop;
}

// The fast-forward of a running simulation that is in batch B (slot, list copy par of its next step) from step
// `steps` over at most h steps, on the context's stream (timed as "ca_ff" when the context times kernels): to its stop
// when no AIR cell can get lava any more (quiet), or up to the first solidification while none can be born (qdead).
// Returns false when neither holds.
static bool ca_fast_forward(vca_ca* ca, const CaStatus& st, int64_t h, const BatchDev& B, uint32_t slot, int par) {
    // the fast-forward reads the simulation's tiles from its entries of the next step's list: lava-front simulations
    // (ca.sweep = full lists its tiles only when the step starts)
    if (!ca->ff || st.stop != VCA_RUNNING || ca->P.dense || ca->P.classic) return false;
    h = std::min<int64_t>(h, static_cast<int64_t>(ca->P.max_steps) - st.steps);
    if (h <= 0) return false;
    const bool quiet = st.quiet > 0 && st.steps >= st.quiet;
    const bool dead = st.qdead > 0 && st.steps == st.qdead && st.steps >= ca->ff_hold;
    if (!quiet && !dead) return false;
    vca_ctx* c = ca->ctx;
    if (!quiet) {
        // up to the first solidification only: skipped (and held off) when that is estimated within a few steps
        ca_launch_ff_estimate(ca->P, ca->D, B, slot, ca->shape, par, st.steps, c->stream);
        CaCtrl cc;
        check(hipMemcpyAsync(&cc, ca->ctrl.p, sizeof(CaCtrl), hipMemcpyDeviceToHost, c->stream), "read ctrl");
        check(hipStreamSynchronize(c->stream), "fast-forward estimate");
        if (cc.ff_sol < vca_ca::kFfWorthSteps) {
            ca->ff_done(st.steps, st.steps);
            return false;
        }
    }
    hipEvent_t ev[2] = {nullptr, nullptr};
    size_t a = 0;
    if (c->timing) {
        ev[0] = c->take();
        ev[1] = c->take();
        a = c->used - 2;
    }
    ca_launch_fast_forward(ca->P, ca->D, B, slot, ca->shape, par, st.steps, static_cast<int32_t>(h), !quiet, c->stream,
                           c->timing ? ev : nullptr);
    if (c->timing) c->mark("ca_ff", a, a + 1);
    check(hipGetLastError(), "fast-forward");
    return true;
}

// Steps a simulation of the batch of one `b` (slot 0) until it stops or `max_new_steps` more steps are done: chunks
// of device-side steps with one host check in between (steps after a stop are no-ops), and the fast-forward where it
// applies.  After a fast-forward the chunks start short again (the next one may apply soon).
// With the fast-forward on, a lava-front simulation runs with its hold on: the step that finds it quiet holds it, so
// the rest of the chunk's launches are no-ops and the fast-forward starts from that step.
static CaStatus ca_run_steps(vca_ca* ca, Batch& b, int64_t max_new_steps) {
    ca_set_hold(ca, ca_holds(ca));
    CaStatus st = ca_status_read(ca);
    const int64_t target = static_cast<int64_t>(st.steps) + std::max<int64_t>(0, max_new_steps);
    int chunk = 128;
    while (st.stop == VCA_RUNNING && st.steps < target) {
        if (ca_fast_forward(ca, st, target - st.steps, b.B, 0u, b.par)) {
            b.desc_ok = false;  // the fast-forward changed types no descent list holds
            const int32_t from = st.steps;
            st = ca_status_read(ca);
            ca->ff_done(from, st.steps);
            // after a long one the next may apply right after the solidification that ended it
            if (st.steps - from >= vca_ca::kFfWorthSteps) chunk = 1;
            continue;
        }
        // (a held simulation is quiet, which admits the fast-forward: released here only should that not apply)
        if (st.held) ca_set_hold(ca, false);
        const int n = static_cast<int>(std::min<int64_t>(target - st.steps, chunk));
        b.enqueue(n);
        check(hipGetLastError(), "ca launch");
        chunk = std::min(128, 2 * chunk);
        st = ca_status_read(ca);
    }
    // held at the last step asked for: released, so that any later stepping goes on from it
    if (st.held) ca_set_hold(ca, false);
    return st;
}

extern "C" int vca_ca_enqueue(vca_ca* ca, int32_t new_steps) {
    return guarded([&] {
        require(ca, "ca == NULL");
        if (ca->member) throw VcaError(VCA_ERR_STATE, "the simulation is in a batch: advance it with vca_batch_*");
        ca->ctx->set_device();
        if (!ca->own) {  // the private batch (members[0] is this simulation) is made on the first solo enqueue
            AllocStream alloc_on(ca->ctx->stream);
            auto own = std::make_unique<Batch>();
            own->init(ca->ctx, 1, ca->P.ntiles);
            own->attach(ca);
            ca->own = std::move(own);
        }
        ca_set_hold(ca, false);  // no host fast-forward follows these steps
        ca->own->enqueue(new_steps);
    });
}

extern "C" int vca_ca_run(vca_ca* ca, int32_t max_new_steps, int32_t* steps_done, int32_t* stop_reason) {
    return guarded([&] {
        require(ca, "ca == NULL");
        if (ca->member) throw VcaError(VCA_ERR_STATE, "the simulation is in a batch: advance it with vca_batch_*");
        ca->ctx->set_device();
        if (!ca->own) {  // the private batch (members[0] is this simulation) is made on the first solo enqueue
            AllocStream alloc_on(ca->ctx->stream);
            auto own = std::make_unique<Batch>();
            own->init(ca->ctx, 1, ca->P.ntiles);
            own->attach(ca);
            ca->own = std::move(own);
        }
        const CaStatus st = ca_run_steps(ca, *ca->own, max_new_steps);
        if (steps_done) *steps_done = st.steps;
        if (stop_reason) *stop_reason = st.stop;
    });
}

extern "C" int vca_ca_set_fast_forward(vca_ca* ca, int32_t on) {
    return guarded([&] {
        require(ca, "ca == NULL");
        ca->ff = on != 0;
        // a batch member holds while the fast-forward is on; off releases it (nothing would fast-forward it)
        ca->ctx->set_device();
        ca_set_hold(ca, ca->member != nullptr && ca_holds(ca));
    });
}

extern "C" int vca_ca_fast_forwarded(vca_ca* ca, int32_t* from_step, int32_t* steps) {
    return guarded([&] {
        require(ca, "ca == NULL");
        ca->ctx->set_device();
        CaCtrl c;
        check(hipMemcpyAsync(&c, ca->ctrl.p, sizeof(CaCtrl), hipMemcpyDeviceToHost, ca->ctx->stream), "read ctrl");
        check(hipStreamSynchronize(ca->ctx->stream), "read ctrl");
        if (from_step) *from_step = c.ff_steps > 0 ? c.ff_from : -1;
        if (steps) *steps = c.ff_steps;
    });
}

extern "C" int vca_ca_status(vca_ca* ca, int32_t* steps_done, int32_t* stop_reason) {
    return guarded([&] {
        require(ca, "ca == NULL");
        ca->ctx->set_device();
        ca_status_sync(ca, steps_done, stop_reason);
    });
}

extern "C" int vca_ca_histories(vca_ca* ca, uint64_t* active, uint64_t* changed, int32_t cap, int32_t* n) {
    return guarded([&] {
        require(ca, "ca == NULL");
        ca->ctx->set_device();
        int32_t steps = 0;
        ca_status_sync(ca, &steps, nullptr);
        const int32_t m = std::min(steps, std::max(0, cap));
        if (m > 0 && active)
            check(hipMemcpyAsync(active, ca->hist_a.p, static_cast<size_t>(m) * 8, hipMemcpyDeviceToHost, ca->ctx->stream),
                  "hist");
        if (m > 0 && changed)
            check(hipMemcpyAsync(changed, ca->hist_c.p, static_cast<size_t>(m) * 8, hipMemcpyDeviceToHost, ca->ctx->stream),
                  "hist");
        check(hipStreamSynchronize(ca->ctx->stream), "hist");
        if (n) *n = steps;
    });
}

extern "C" int vca_ca_download(vca_ca* ca, uint8_t* type, int32_t* flow, float* temp, uint16_t* solid,
                               uint16_t* arrival) {
    return guarded([&] {
        require(ca, "ca == NULL");
        ca->ctx->set_device();
        int32_t steps = 0;
        ca_status_sync(ca, &steps, nullptr);
        const uint64_t n = ca->n;
        if (type) check(hipMemcpyAsync(type, ca->type_after(steps), n, hipMemcpyDeviceToHost, ca->ctx->stream), "download");
        if (flow) check(hipMemcpyAsync(flow, ca->flow.p, n * 4, hipMemcpyDeviceToHost, ca->ctx->stream), "download");
        if (temp) {
            ca_launch_final_temps(ca->D, n, steps & 1, ca->ctx->stream);
            check(hipMemcpyAsync(temp, ca->temp_after(steps), n * 4, hipMemcpyDeviceToHost, ca->ctx->stream), "download");
        }
        if (solid) check(hipMemcpyAsync(solid, ca->solid.p, n * 2, hipMemcpyDeviceToHost, ca->ctx->stream), "download");
        if (arrival) check(hipMemcpyAsync(arrival, ca->arrival.p, n * 2, hipMemcpyDeviceToHost, ca->ctx->stream), "download");
        check(hipStreamSynchronize(ca->ctx->stream), "download");
    });
}

extern "C" int vca_ca_work(vca_ca* ca, uint64_t* tiles_visited, uint64_t* changes, int32_t* tile_cells) {
    return guarded([&] {
        require(ca, "ca == NULL");
        ca->ctx->set_device();
        CaCtrl c;
        check(hipMemcpyAsync(&c, ca->ctrl.p, sizeof(CaCtrl), hipMemcpyDeviceToHost, ca->ctx->stream), "read ctrl");
        check(hipStreamSynchronize(ca->ctx->stream), "sync");
        if (tiles_visited) *tiles_visited = c.tiles_visited;
        if (changes) *changes = c.changes_total;
        if (tile_cells) *tile_cells = kTileCells;
    });
}

extern "C" int vca_ca_footprint(const vca_ca_desc* d, uint64_t* bytes) {
    return guarded([&] {
        require(d && bytes, "NULL argument");
        require(d->nx >= 1 && d->ny >= 1 && d->nz >= 1, "wymiary muszą być >= 1");
        const int shape = d->tile_shape > 0 ? d->tile_shape - 1 : default_shape(d->nx, d->ny, d->nz);
        const uint64_t n = static_cast<uint64_t>(d->nx) * d->ny * d->nz;
        const uint64_t tiles = static_cast<uint64_t>((d->nx + kTileX - 1) / kTileX) *
                               static_cast<uint64_t>((d->ny + kShapeTY[shape] - 1) / kShapeTY[shape]) *
                               static_cast<uint64_t>((d->nz + kShapeTZ[shape] - 1) / kShapeTZ[shape]);
        const uint64_t cols = static_cast<uint64_t>(d->nx) * d->ny;
        // the simulation's allocation (vents: a few slots), and the batch-of-one lists a solo run adds
        const CaPlan plan = ca_plan(n, static_cast<uint32_t>(std::min<uint64_t>(tiles, kTileMask)), cols, d->classic == 0,
                                    64, 64, d->max_steps);
        *bytes = plan.L.total + tiles * (4 * kShapes * 8 + 4) + 4096;
    });
}

extern "C" int vca_ca_autotune(vca_ctx* ctx, const vca_ca_desc* desc, const int32_t* height, const int32_t* spill,
                               const vca_vent* vents, int32_t nvents, int32_t probe_steps, vca_tune_result* out) {
    return guarded([&] {
        require(ctx && desc && out && probe_steps > 0, "invalid argument");
        ctx->set_device();
        ctx->drain();
        // the caller's timing state is kept aside and restored
        const bool timing = ctx->timing;
        std::map<std::string, KernelTime> saved;
        saved.swap(ctx->times);
        *out = vca_tune_result{};
        out->probe_steps = probe_steps;
        std::vector<int> shapes;
        if (desc->ny == 1)
            shapes = {4};  // x-z plane: 64 x 1 x 16
        else if (desc->nz == 1)
            shapes = {5};  // x-y plane: 64 x 16 x 1
        else
            shapes = {1, 2, 3, 4, 5};
        float best = 0.0f;
        try {
            ctx->timing = true;
            for (int dense = 0; dense < 2; ++dense)
                for (int sh : shapes) {
                    vca_ca_desc d = *desc;
                    d.tile_shape = sh;
                    d.dense = dense;
                    vca_ca* ca = nullptr;
                    if (vca_ca_create(ctx, &d, height, spill, vents, nvents, &ca) != VCA_OK)
                        throw VcaError(VCA_ERR_ARG, g_last_error);
                    ctx->drain();
                    ctx->times.clear();
                    int32_t steps = 0, stop = 0;
                    const int rc = vca_ca_run(ca, probe_steps, &steps, &stop);
                    ctx->drain();
                    vca_ca_destroy(ca);
                    if (rc != VCA_OK) throw VcaError(rc, g_last_error);
                    const KernelTime& k = ctx->times["ca_step"];
                    const float ms = k.launches ? static_cast<float>(k.ms / static_cast<double>(k.launches)) : 0.0f;
                    out->ms_per_step[dense][sh - 1] = ms;
                    ++out->candidates;
                    if (k.launches && (best == 0.0f || ms < best)) {
                        best = ms;
                        out->tile_shape = sh;
                        out->dense = dense;
                    }
                }
        } catch (...) {
            ctx->timing = timing;
            ctx->times.swap(saved);
            throw;
        }
        ctx->timing = timing;
        ctx->times.swap(saved);
        if (out->tile_shape == 0) {  // nothing ran (the simulation stops before its first step)
            out->tile_shape = default_shape(desc->nx, desc->ny, desc->nz) + 1;
            out->dense = desc->dense;
        }
    });
}

extern "C" int vca_ca_tiles(vca_ca* ca, uint64_t* tiles) {
    return guarded([&] {
        require(ca && tiles, "NULL argument");
        *tiles = ca->P.ntiles;
    });
}

// ===================================================================================================
// vca_ca_run without device allocation
// ===================================================================================================
// The device memory of vca_ca_run_noalloc (its batch of one: list counts, tile owners, both copies of the tile lists)
// is this array, one copy per device, held until the process ends; host code reaches it only through the address
// hipGetSymbolAddress gives for the current device (kept per device by noalloc_setup).
constexpr size_t kScratchBytes = static_cast<size_t>(16) << 20;
static __device__ __align__(256) uint8_t scratch[kScratchBytes];

namespace {
constexpr int kMaxDevices = 64;
struct NoAlloc {
    void* ptr = nullptr;          // this device's copy of `scratch`
    Batch* batch = nullptr;       // the batch of one over it (never freed)
    std::mutex mu;                // one call at a time per device (the calls share the array)
};
std::mutex g_noalloc_mu;
NoAlloc* g_noalloc[kMaxDevices] = {};

// Once per device, with that device current (vca_ctx_create): the array's address on the device, a simulation table
// entry of its own and the host side of its batch -- everything vca_ca_run_noalloc would otherwise have to allocate.
void noalloc_setup(int device) {
    std::lock_guard<std::mutex> lock(g_noalloc_mu);
    if (device < 0 || device >= kMaxDevices || g_noalloc[device]) return;
    void* p = nullptr;
    check(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)");
    auto* na = new NoAlloc();
    na->ptr = p;
    na->batch = new Batch();
    na->batch->init_external(device, static_cast<char*>(p), kScratchBytes, SimTable::take(device, 1));
    g_noalloc[device] = na;
}
}  // namespace

extern "C" int vca_ca_run_noalloc(vca_ca* ca, int32_t max_new_steps, int32_t* steps_done, int32_t* stop_reason) {
    return guarded([&] {
        require(ca, "ca == NULL");
        require(max_new_steps >= 0, "max_new_steps < 0");
        if (ca->member) throw VcaError(VCA_ERR_STATE, "the simulation is in a batch: advance it with vca_batch_*");
        if (ca->own)
            throw VcaError(VCA_ERR_STATE, "the simulation was advanced by vca_ca_run / vca_ca_enqueue, whose private batch "
                                          "holds its tile lists: continue with those");
        vca_ctx* ctx = ca->ctx;
        ctx->set_device();
        NoAlloc* na = ctx->device < kMaxDevices ? g_noalloc[ctx->device] : nullptr;
        if (!na) throw VcaError(VCA_ERR_STATE, "no static device array set up for this device");
        Batch& b = *na->batch;
        if (ca->P.ntiles > b.cap)
            throw VcaError(VCA_ERR_ARG, "vca_ca_run_noalloc: the grid has " + std::to_string(ca->P.ntiles) +
                                            " tiles of 1024 cells; the static device array (" +
                                            std::to_string(kScratchBytes >> 20) + " MiB) holds the lists of at most " +
                                            std::to_string(b.cap) + " tiles");
        std::lock_guard<std::mutex> lock(na->mu);
        // timed launches take events from the context's pool, which may create events: not here
        const bool timing = ctx->timing;
        ctx->timing = false;
        auto leave = [&] {
            if (b.members[0]) {
                b.host_sims[0].running = 0;
                ca_write_sim(b.sim_base, &b.host_sims[0], ctx->stream);
                b.members[0] = nullptr;
            }
            (void)hipStreamSynchronize(ctx->stream);
            ctx->timing = timing;
        };
        try {
            b.reset_external(ctx);
            b.attach(ca);
            const CaStatus st = ca_run_steps(ca, b, max_new_steps);
            leave();
            if (steps_done) *steps_done = st.steps;
            if (stop_reason) *stop_reason = st.stop;
        } catch (...) {
            leave();
            throw;
        }
    });
}

struct vca_batch {
    Batch b;
};

extern "C" int vca_batch_create(vca_ctx* ctx, int32_t max_sims, uint64_t max_tiles, vca_batch** out) {
    return guarded([&] {
        require(ctx && out, "NULL argument");
        require(max_sims >= 1 && max_sims <= kMaxSlots, "max_sims poza zakresem 1..1024");
        require(max_tiles >= 1 && max_tiles <= kTileMask, "max_tiles poza zakresem");
        ctx->set_device();
        AllocStream alloc_on(ctx->stream);
        auto b = std::make_unique<vca_batch>();
        b->b.init(ctx, max_sims, static_cast<uint32_t>(max_tiles));
        check(hipStreamSynchronize(ctx->stream), "batch create");
        *out = b.release();
    });
}

extern "C" int vca_batch_destroy(vca_batch* b) {
    return guarded([&] {
        if (!b) return;
        b->b.ctx->set_device();
        (void)hipStreamSynchronize(b->b.ctx->stream);
        for (vca_ca* c : b->b.members)
            if (c) {
                c->member = nullptr;
                c->slot = -1;
            }
        delete b;
    });
}

extern "C" int vca_batch_add(vca_batch* b, vca_ca* ca, int32_t* slot) {
    return guarded([&] {
        require(b && ca, "NULL argument");
        require(ca->ctx == b->b.ctx, "the simulation and the batch use different contexts");
        if (ca->member) throw VcaError(VCA_ERR_STATE, "the simulation is already in a batch");
        b->b.ctx->set_device();
        // the batch's status read fast-forwards a held member (from the step that found it quiet)
        ca_set_hold(ca, ca_holds(ca));
        ca->slot = b->b.attach(ca);
        ca->member = &b->b;
        // the private batch of earlier solo steps holds lists for a step the shared batch moves past: a later solo run
        // makes a new one
        ca->own.reset();
        if (slot) *slot = ca->slot;
    });
}

extern "C" int vca_batch_status(vca_batch* b, int32_t* out, int32_t cap) {
    return guarded([&] {
        require(b && out && cap >= 0, "invalid argument");
        Batch& bb = b->b;
        bb.ctx->set_device();
        AllocStream alloc_on(bb.ctx->stream);
        if (!bb.status.p) bb.status.alloc(static_cast<size_t>(bb.max_slots) * 16);
        std::vector<int32_t> all(static_cast<size_t>(bb.B.nslots) * 4);
        auto read = [&] {
            ca_launch_status(bb.B, bb.status.as<int32_t>(), bb.ctx->stream);
            check(hipGetLastError(), "batch status");
            check(hipMemcpyAsync(all.data(), bb.status.p, all.size() * 4, hipMemcpyDeviceToHost, bb.ctx->stream),
                  "batch status");
            check(hipStreamSynchronize(bb.ctx->stream), "batch status");
            bb.ctx->drain();
        };
        read();
        // members where no cell can be born any more (to their stop) or until a type changes: the fast-forward
        std::vector<int32_t> from(bb.B.nslots, -1);
        bool ff = false;
        for (uint32_t k = 0; k < bb.B.nslots; ++k) {
            vca_ca* m = bb.members[k];
            const CaStatus st{all[4 * k], all[4 * k + 1], all[4 * k + 2], all[4 * k + 3]};
            if (m && ca_fast_forward(m, st, m->P.max_steps, bb.B, k, bb.par)) {
                from[k] = st.steps;
                ff = true;
                bb.desc_ok = false;  // the fast-forward changed types no descent list holds
            }
        }
        if (ff) {
            read();
            for (uint32_t k = 0; k < bb.B.nslots; ++k)
                if (from[k] >= 0 && bb.members[k]) bb.members[k]->ff_done(from[k], all[4 * k]);
        }
        for (int32_t k = 0; k < cap; ++k) {
            const bool in = static_cast<uint32_t>(k) < bb.B.nslots && bb.members[static_cast<size_t>(k)];
            out[2 * k] = in ? all[4 * static_cast<size_t>(k)] : 0;
            out[2 * k + 1] = in ? all[4 * static_cast<size_t>(k) + 1] : -1;
        }
    });
}

extern "C" int vca_batch_remove(vca_batch* b, vca_ca* ca) {
    return guarded([&] {
        require(b && ca, "NULL argument");
        if (ca->member != &b->b) throw VcaError(VCA_ERR_STATE, "the simulation is not in this batch");
        b->b.ctx->set_device();
        b->b.detach(ca, ca->slot);
        ca->member = nullptr;
        ca->slot = -1;
    });
}

extern "C" int vca_batch_set_graph(vca_batch* b, int32_t steps_per_graph) {
    return guarded([&] {
        require(b && steps_per_graph >= 0 && steps_per_graph <= 4096, "invalid argument");
        b->b.ctx->set_device();
        b->b.drop_graph();
        b->b.steps_per_graph = steps_per_graph + (steps_per_graph & 1);  // graphs start and end on parity 0
    });
}

extern "C" int vca_batch_enqueue(vca_batch* b, int32_t steps) {
    return guarded([&] {
        require(b && steps >= 0, "invalid argument");
        b->b.ctx->set_device();
        b->b.enqueue(steps);
    });
}

namespace {
// Running members (one device read for all of them).
int batch_running(vca_batch* b) {
    Batch& bb = b->b;
    std::vector<int32_t> st(static_cast<size_t>(bb.max_slots) * 2);
    if (vca_batch_status(b, st.data(), bb.max_slots) != VCA_OK) throw VcaError(VCA_ERR_HIP, g_last_error);
    int running = 0;
    for (int k = 0; k < bb.max_slots; ++k) running += st[2 * static_cast<size_t>(k) + 1] == VCA_RUNNING;
    return running;
}
}  // namespace

extern "C" int vca_batch_run(vca_batch* b, int32_t max_steps, int32_t* running) {
    return guarded([&] {
        require(b && max_steps >= 0, "invalid argument");
        b->b.ctx->set_device();
        int live = batch_running(b);
        int remaining = max_steps;
        while (remaining > 0 && live > 0) {
            const int chunk = std::min(remaining, 128);
            b->b.enqueue(chunk);
            remaining -= chunk;
            live = batch_running(b);
        }
        if (running) *running = live;
    });
}

extern "C" int vca_batch_info(vca_batch* b, int32_t* members, uint64_t* tiles_used, int32_t* running) {
    return guarded([&] {
        require(b, "b == NULL");
        b->b.ctx->set_device();
        if (members) *members = b->b.count();
        if (tiles_used) {
            uint64_t t = 0;
            for (vca_ca* c : b->b.members)
                if (c) t += c->P.ntiles;
            *tiles_used = t;
        }
        if (running) *running = batch_running(b);
    });
}

// ===================================================================================================
// MC
// ===================================================================================================
namespace {

// Colour of (x, y, z): buildColors / colorOf (PottsMC.cpp:145-224).
struct Colouring {
    bool parity = false;
    int index[27];
    int count = 0;
    int dims[3];
    int per[3];
    int axis_class(int c, int a) const {
        if (per[a] && (dims[a] % 2) == 1 && c == dims[a] - 1) return 2;
        return c & 1;
    }
    int of(int x, int y, int z) const {
        if (parity) return (x + y + z) & 1;
        return index[axis_class(x, 0) + 3 * axis_class(y, 1) + 9 * axis_class(z, 2)];
    }
};

Colouring make_colouring(int kind, const int dims[3], const int per[3]) {
    Colouring c;
    bool odd = false;
    for (int a = 0; a < 3; ++a) {
        c.dims[a] = dims[a];
        c.per[a] = per[a];
        odd = odd || (per[a] && dims[a] % 2 == 1);
    }
    c.parity = kind == VCA_VON_NEUMANN && !odd;
    std::fill(std::begin(c.index), std::end(c.index), -1);
    if (c.parity) {
        c.count = 2;
        return c;
    }
    int classes[3];
    for (int a = 0; a < 3; ++a) classes[a] = dims[a] == 1 ? 1 : (per[a] && dims[a] % 2 == 1 ? 3 : 2);
    for (int cz = 0; cz < classes[2]; ++cz)
        for (int cy = 0; cy < classes[1]; ++cy)
            for (int cx = 0; cx < classes[0]; ++cx) c.index[cx + 3 * cy + 9 * cz] = c.count++;
    return c;
}


}  // namespace

namespace {

// PottsMC ctor + setDomain + setCheckerboard + initialize (PottsMC.cpp:14-143) with the domain given on the device
// (L.dom: nothing, 0/1 bytes or the CA cell types) and the mobility, when there is one, already in mc->mob_cells.
// The domain lists are built on the device: colour-major, global index ascending within a colour (buildColors),
// or index order for the sequential random mode; one host read of the colour totals.
std::unique_ptr<vca_mc> mc_build(vca_ctx* ctx, const vca_mc_desc* d, McListSrc L, std::unique_ptr<vca_mc> mc,
                                 bool mobility_dynamics) {
    hipStream_t s = ctx->stream;
    McParams& P = mc->P;
    const uint64_t n = mc->n;
    const int dims[3] = {d->nx, d->ny, d->nz};
    const Colouring col = make_colouring(d->neighborhood, dims, P.per);
    L.n = n;
    L.ncol = mc->checkerboard ? col.count : 1;
    L.parity = col.parity ? 1 : 0;
    for (int a = 0; a < 3; ++a) {
        L.dims[a] = dims[a];
        L.per[a] = P.per[a];
    }
    for (int k = 0; k < 27; ++k) L.index[k] = col.index[k];
    mc->colors = mc->checkerboard ? col.count : 0;
    const uint64_t nseg = mc_list_segments(n);
    DevBuf cnt, totals;
    cnt.alloc(static_cast<size_t>(L.ncol) * nseg * 4);
    totals.alloc(static_cast<size_t>(L.ncol) * 4 * 2);
    mc_launch_list_count(L, cnt.as<uint32_t>(), totals.as<uint32_t>(), s);
    check(hipGetLastError(), "mc lists");
    std::vector<uint32_t> tot(static_cast<size_t>(L.ncol));
    check(hipMemcpyAsync(tot.data(), totals.p, tot.size() * 4, hipMemcpyDeviceToHost, s), "read colour totals");
    check(hipStreamSynchronize(s), "mc lists");
    mc->color_start.assign(1, 0);
    uint64_t sum = 0;
    for (uint32_t t : tot) {
        sum += t;
        mc->color_start.push_back(static_cast<uint32_t>(sum));
    }
    mc->ndomain = static_cast<uint32_t>(sum);
    // first slot of each colour, after the totals in the same buffer
    check(hipMemcpyAsync(totals.as<uint32_t>() + L.ncol, mc->color_start.data(), static_cast<size_t>(L.ncol) * 4,
                         hipMemcpyHostToDevice, s),
          "copy colour starts");
    mc->list.alloc(static_cast<size_t>(mc->ndomain) * 4);
    uint64_t* thr = nullptr;
    if (mc->has_mob_cells && mobility_dynamics) {
        mc->mob_thr.alloc(static_cast<size_t>(mc->ndomain) * 8);
        thr = mc->mob_thr.as<uint64_t>();
        mc->has_mob_dyn = true;
    }
    mc_launch_list_scatter(L, cnt.as<uint32_t>(), totals.as<uint32_t>() + L.ncol, mc->list.as<uint32_t>(),
                           mc->has_mob_cells ? mc->mob_cells.as<float>() : nullptr, thr, s);
    check(hipGetLastError(), "mc lists");
    mc->spin.alloc(n * static_cast<uint64_t>(mc->spin_bytes));
    check(hipMemsetAsync(mc->spin.p, 0, n * static_cast<uint64_t>(mc->spin_bytes), s), "memset");
    // Metropolis table T[before][after] (PottsMC.cpp:431-436): dE = J*after - J*before in double;
    // accept-always for dE <= 0, else (r >> 11) < ceil(exp(-dE/kT) * 2^53) (exp from the host libm).
    std::vector<uint64_t> acc(27 * 27, 0);
    for (int b = 0; b < 27; ++b)
        for (int a = 0; a < 27; ++a) {
            const double dE = d->J * a - d->J * b;
            uint64_t t = 0;
            if (dE <= 0.0) {
                t = kDownhill;
            } else if (d->kT > 0.0) {
                const double v = std::ceil(std::exp(-dE / d->kT) * 9007199254740992.0);
                t = v >= 9007199254740992.0 ? (1ULL << 53) : static_cast<uint64_t>(v);
            }
            acc[static_cast<size_t>(b) * 27 + a] = t;
        }
    mc->accept.alloc(acc.size() * 8);
    check(hipMemcpyAsync(mc->accept.p, acc.data(), acc.size() * 8, hipMemcpyHostToDevice, s), "copy table");
    mc->counters.alloc(kMcRows * kMcCounters * 8);
    check(hipMemsetAsync(mc->counters.p, 0, kMcRows * kMcCounters * 8, s), "memset");
    if (mc->checkerboard) {
        mc->bar.alloc(sizeof(McGridBar));
        check(hipMemsetAsync(mc->bar.p, 0, sizeof(McGridBar), s), "memset");
    }
    mc->gout.alloc(kMcRows * kMcSlots * 8);
    mc_launch_init_spins(P, mc->spin.p, mc->spin_bytes, mc->list.as<uint32_t>(), mc->ndomain, s);
    check(hipStreamSynchronize(s), "mc init");  // also: the temporary buffers and the caller's inputs are free
    return mc;
}

std::unique_ptr<vca_mc> mc_new(vca_ctx* ctx, const vca_mc_desc* d) {
    require(d->Q >= 2, "MC wymaga co najmniej 2 orientacji (Q >= 2)");
    require(d->nx >= 1 && d->ny >= 1 && d->nz >= 1, "wymiary muszą być >= 1");
    const uint64_t n = static_cast<uint64_t>(d->nx) * d->ny * d->nz;
    require(n <= 2000000000ULL, "przestrzeń ma za dużo komórek");
    auto mc = std::make_unique<vca_mc>();
    mc->ctx = ctx;
    mc->n = n;
    McParams& P = mc->P;
    P.nx = d->nx;
    P.ny = d->ny;
    P.nz = d->nz;
    const int dims[3] = {d->nx, d->ny, d->nz};
    for (int a = 0; a < 3; ++a) P.per[a] = d->periodic[a] && dims[a] > 1;
    P.seed = d->seed;
    P.Q = d->Q;
    P.nb = make_nb(d->neighborhood, d->nx, d->ny, d->nz);
    mc->checkerboard = d->checkerboard != 0;
    mc->J = d->J;
    mc->kT = d->kT;
    mc->spin_bytes = d->Q <= 255 ? 1 : (d->Q <= 65535 ? 2 : 4);
    return mc;
}

}  // namespace

extern "C" int vca_mc_create(vca_ctx* ctx, const vca_mc_desc* d, const uint8_t* domain, const float* mobility,
                             int32_t mobility_dynamics, vca_mc** out) {
    return guarded([&] {
        require(ctx && d && out, "NULL argument");
        ctx->set_device();
        AllocStream alloc_on(ctx->stream);
        auto mc = mc_new(ctx, d);
        hipStream_t s = ctx->stream;
        McListSrc L{};
        DevBuf dom;
        if (domain) {
            dom.alloc(mc->n);
            check(hipMemcpyAsync(dom.p, domain, mc->n, hipMemcpyHostToDevice, s), "copy domain");
            L.mode = 1;
            L.dom = dom.as<uint8_t>();
        }
        if (mobility) {
            mc->has_mob_cells = true;
            mc->mob_cells.alloc(mc->n * 4);
            check(hipMemcpyAsync(mc->mob_cells.p, mobility, mc->n * 4, hipMemcpyHostToDevice, s), "copy mobility");
        }
        *out = mc_build(ctx, d, L, std::move(mc), mobility_dynamics != 0).release();
    });
}

extern "C" int vca_mc_create_from_ca(vca_ctx* ctx, const vca_mc_desc* d, vca_ca* ca, int32_t use_mobility,
                                     double mobility_min, vca_mc** out) {
    return guarded([&] {
        require(ctx && d && ca && out, "NULL argument");
        require(ca->P.nx == d->nx && ca->P.ny == d->ny && ca->P.nz == d->nz, "wymiary MC różnią się od CA");
        require(ca->ctx->device == ctx->device, "MC i CA na różnych urządzeniach");
        require(!ca->member, "the simulation is in a batch: remove it first");
        ctx->set_device();
        AllocStream alloc_on(ctx->stream);
        auto mc = mc_new(ctx, d);
        hipStream_t s = ctx->stream;
        // a stopped simulation's state is final (later steps are no-ops), so ctx may be another context whose
        // stream is not ordered after the CA's
        int32_t steps = 0;
        {
            SideStream side(ca->ctx);
            steps = require_stopped(ca, side.s);
        }
        const uint8_t* types = static_cast<const uint8_t*>(ca->type_after(steps));
        // CoolingMobility (CoolingMobility.cpp:8-129) from the CA result on the device
        mc->has_mob_cells = true;
        mc->mob_cells.alloc(mc->n * 4);
        DevBuf red;
        red.alloc(16);
        const uint32_t init[4] = {0, 0xFFFFFFFFu, 0, 0};
        check(hipMemcpyAsync(red.p, init, sizeof init, hipMemcpyHostToDevice, s), "copy");
        mc_launch_cooling(mc->n, types, ca->solid.as<uint16_t>(), ca->arrival.as<uint16_t>(),
                          red.as<uint32_t>(), mobility_min, mc->mob_cells.as<float>(), s);
        check(hipGetLastError(), "cooling mobility");
        McListSrc L{};
        L.mode = 2;
        L.dom = types;
        *out = mc_build(ctx, d, L, std::move(mc), use_mobility != 0).release();
    });
}

extern "C" int vca_mc_destroy(vca_mc* mc) {
    return guarded([&] {
        if (!mc) return;
        mc->ctx->set_device();
        (void)hipStreamSynchronize(mc->ctx->stream);
        delete mc;
    });
}

namespace {

// Blocks of the persistent sweep kernel for nsweeps sweeps of `mc`, or 0 for one launch per colour pass: one block per
// 256 cells of the largest colour, at most one per CU (so that all blocks are resident together, next to other work
// on the device).  Only on request: on MI355X a device-wide barrier between passes (eight XCDs, each with its own L2)
// costs more than the launch boundary it replaces -- 13 to 20 us per pass against 8.5 us for a launch per colour on
// the 818 k-cell basalt of the 1024 x 1024 x 384 caldera.
unsigned mc_persistent_grid(const vca_mc* mc, int32_t nsweeps) {
    if (!mc->checkerboard || mc->colors < 1 || nsweeps < 1 || mc->sweep_mode != VCA_MC_SWEEP_PERSISTENT) return 0;
    uint32_t most = 0;
    for (int c = 0; c < mc->colors; ++c)
        most = std::max(most, mc->color_start[static_cast<size_t>(c) + 1] - mc->color_start[static_cast<size_t>(c)]);
    if (most == 0) return 0;
    const uint64_t need = (static_cast<uint64_t>(most) + kBlock - 1) / kBlock;
    const uint64_t cus = static_cast<uint64_t>(std::max(1, mc->ctx->cus));
    return static_cast<unsigned>(std::min(need, cus));
}

// A persistent sweep launch whose barrier wait gave up leaves wrong spins: reported at the next synchronisation.
void mc_check_abort(vca_mc* mc, const uint32_t& flag) {
    if (flag == 0) return;
    check(hipMemsetAsync(mc->bar.p, 0, sizeof(McGridBar), mc->ctx->stream), "memset");
    check(hipStreamSynchronize(mc->ctx->stream), "sync");
    throw VcaError(VCA_ERR_HIP, "MC: a barrier wait of the persistent sweep kernel exceeded its time limit (its blocks "
                                "were not resident together); the spins are invalid");
}

}  // namespace

extern "C" int vca_mc_counts_get(vca_mc* mc, vca_mc_counts* out) {
    return guarded([&] {
        require(mc && out, "NULL argument");
        mc->ctx->set_device();
        std::vector<unsigned long long> c(kMcRows * kMcCounters);
        check(hipMemcpyAsync(c.data(), mc->counters.p, c.size() * 8, hipMemcpyDeviceToHost, mc->ctx->stream),
              "counters");
        uint32_t aborted = 0;
        if (mc->bar.p)
            check(hipMemcpyAsync(&aborted, mc->bar.as<McGridBar>()->abort, 4, hipMemcpyDeviceToHost, mc->ctx->stream),
                  "barrier flag");
        check(hipStreamSynchronize(mc->ctx->stream), "sync");
        mc->ctx->drain();
        mc_check_abort(mc, aborted);
        out->attempts = out->accepted = out->uphill = 0;
        for (int r = 0; r < kMcRows; ++r) {
            out->attempts += c[r * kMcCounters + 0];
            out->accepted += c[r * kMcCounters + 1];
            out->uphill += c[r * kMcCounters + 2];
        }
        out->sweeps = mc->sweeps;
    });
}


extern "C" int vca_mc_set_sweep_mode(vca_mc* mc, int32_t mode) {
    return guarded([&] {
        require(mc, "mc == NULL");
        require(mode == VCA_MC_SWEEP_AUTO || mode == VCA_MC_SWEEP_PER_COLOUR || mode == VCA_MC_SWEEP_PERSISTENT,
                "unknown sweep mode");
        if (!mc->checkerboard && mode != VCA_MC_SWEEP_AUTO)
            throw VcaError(VCA_ERR_STATE, "the random update has no colour passes");
        mc->sweep_mode = mode;
    });
}

extern "C" int vca_mc_sweep(vca_mc* mc, int32_t nsweeps, vca_mc_counts* counts) {
    return guarded([&] {
        require(mc && nsweeps >= 0, "invalid argument");
        vca_ctx* c = mc->ctx;
        c->set_device();
        hipStream_t s = c->stream;
        const uint64_t* thr = mc->has_mob_dyn ? mc->mob_thr.as<uint64_t>() : nullptr;
        if (const unsigned grid = mc_persistent_grid(mc, nsweeps)) {
            McColours C{};
            C.count = mc->colors;
            for (int k = 0; k <= mc->colors; ++k) C.start[k] = mc->color_start[static_cast<size_t>(k)];
            const size_t t0 = c->timing ? c->record() : 0;
            mc_launch_sweeps(mc->P, mc->spin.p, mc->spin_bytes, mc->list.as<uint32_t>(), thr, C,
                             static_cast<uint64_t>(mc->sweeps), nsweeps, mc->accept.as<uint64_t>(),
                             mc->counters.as<unsigned long long>(), mc->bar.as<McGridBar>(), 10 * c->clock_hz, grid,
                             s);
            if (c->timing) {
                c->mark("mc_sweeps", t0, c->record());
                if (c->used > 4096) c->drain();
            }
            mc->sweeps += nsweeps;
            nsweeps = 0;
        }
        for (int k = 0; k < nsweeps; ++k) {
            const uint64_t prefix = hash_prefix(mc->P.seed, static_cast<uint64_t>(mc->sweeps));
            const bool tm = c->time_mc_sweep();
            const size_t t0 = tm ? c->record() : 0;
            if (mc->checkerboard) {
                for (int col = 0; col < mc->colors; ++col) {
                    const uint32_t b = mc->color_start[static_cast<size_t>(col)];
                    const uint32_t e = mc->color_start[static_cast<size_t>(col) + 1];
                    const size_t k0 = tm ? c->record() : 0;
                    mc_launch_color(mc->P, mc->spin.p, mc->spin_bytes, mc->list.as<uint32_t>() + b, thr ? thr + b : nullptr,
                                    e - b, prefix, mc->accept.as<uint64_t>(),
                                    mc->counters.as<unsigned long long>(), s);
                    if (tm && e > b) c->mark("mc_color", k0, c->record());
                }
            } else {
                mc_launch_random(mc->P, mc->spin.p, mc->spin_bytes, mc->list.as<uint32_t>(), thr, mc->ndomain, prefix,
                                 mc->accept.as<uint64_t>(), mc->counters.as<unsigned long long>(), s);
            }
            if (tm) {
                c->mark("mc_sweep", t0, c->record());
                if (c->used > 4096) c->drain();
            }
            ++mc->sweeps;
        }
        check(hipGetLastError(), "mc launch");
        if (counts) {
            if (vca_mc_counts_get(mc, counts) != VCA_OK) throw VcaError(VCA_ERR_HIP, g_last_error);
        }
    });
}

extern "C" int vca_mc_measure(vca_mc* mc, vca_mc_stats* st, vca_mc_class* classes, uint64_t* high_cells) {
    return guarded([&] {
        require(mc && st, "NULL argument");
        mc->ctx->set_device();
        AllocStream alloc_on(mc->ctx->stream);
        hipStream_t s = mc->ctx->stream;
        const uint64_t n = mc->n;
        if (!mc->parent.p) {
            mc->parent.alloc(n * 4);
            mc->size.alloc(n * 4);
            if (mc->has_mob_cells) mc->hi.alloc(n * 4);
        }
        check(hipMemsetAsync(mc->gout.p, 0, kMcRows * kMcSlots * 8, s), "memset");
        const uint64_t prefix = hash_prefix(mc->P.seed, static_cast<uint64_t>(mc->sweeps));
        mc_launch_energy(mc->P, mc->spin.p, mc->spin_bytes, mc->list.as<uint32_t>(), mc->ndomain, prefix,
                         mc->gout.as<unsigned long long>(), s);
        mc_launch_grains(mc->P, mc->spin.p, mc->spin_bytes, mc->list.as<uint32_t>(), mc->ndomain,
                         mc->parent.as<uint32_t>(), mc->size.as<uint32_t>(),
                         mc->has_mob_cells ? mc->hi.as<uint32_t>() : nullptr, mc->mob_cells.as<float>(),
                         mc->gout.as<unsigned long long>(), s);
        std::vector<unsigned long long> rows(kMcRows * kMcSlots);
        check(hipMemcpyAsync(rows.data(), mc->gout.p, rows.size() * 8, hipMemcpyDeviceToHost, s), "grains");
        uint32_t aborted = 0;
        if (mc->bar.p)
            check(hipMemcpyAsync(&aborted, mc->bar.as<McGridBar>()->abort, 4, hipMemcpyDeviceToHost, s), "barrier flag");
        check(hipStreamSynchronize(s), "measure");
        mc_check_abort(mc, aborted);
        unsigned long long g[kMcSlots] = {};
        for (int r = 0; r < kMcRows; ++r)
            for (int i = 0; i < kMcSlots; ++i) {
                const unsigned long long x = rows[static_cast<size_t>(r) * kMcSlots + i];
                g[i] = (i == 1 || i == 4 || i == 7) ? std::max(g[i], x) : g[i] + x;
            }
        const unsigned long long unlike = g[9];
        st->mcs = mc->sweeps;
        st->grains = g[0];
        st->cells = mc->ndomain;
        st->largest = g[1];
        st->mean_grain_size = g[0] > 0 ? double(mc->ndomain) / double(g[0]) : 0.0;
        // totalEnergy: J * P for symmetric neighbourhoods, 0.5 * J * sum for random ones (PottsMC.cpp:642, 663)
        st->energy = mc->P.nb.variants == 1 ? mc->J * static_cast<double>(unlike) : 0.5 * mc->J * static_cast<double>(unlike);
        if (classes && mc->has_mob_cells) {
            classes[0] = vca_mc_class{g[2], g[3], g[4]};
            classes[1] = vca_mc_class{g[5], g[6], g[7]};
        }
        if (high_cells) *high_cells = g[8];
    });
}

extern "C" int vca_mc_download(vca_mc* mc, int32_t* crystal) {
    return guarded([&] {
        require(mc && crystal, "NULL argument");
        mc->ctx->set_device();
        check(hipStreamSynchronize(mc->ctx->stream), "sync");
        const uint64_t n = mc->n;
        if (mc->spin_bytes == 4) {
            check(hipMemcpyAsync(crystal, mc->spin.p, n * 4, hipMemcpyDeviceToHost, mc->ctx->stream), "download");
            check(hipStreamSynchronize(mc->ctx->stream), "download");
            return;
        }
        std::vector<uint8_t> raw(n * static_cast<uint64_t>(mc->spin_bytes));
        check(hipMemcpyAsync(raw.data(), mc->spin.p, raw.size(), hipMemcpyDeviceToHost, mc->ctx->stream), "download");
        check(hipStreamSynchronize(mc->ctx->stream), "download");
        if (mc->spin_bytes == 1)
            for (uint64_t i = 0; i < n; ++i) crystal[i] = raw[i];
        else
            for (uint64_t i = 0; i < n; ++i) crystal[i] = reinterpret_cast<const uint16_t*>(raw.data())[i];
    });
}

extern "C" int vca_mc_download_labels(vca_mc* mc, uint32_t* root) {
    return guarded([&] {
        require(mc && root, "NULL argument");
        require(mc->parent.p != nullptr, "vca_mc_measure musi poprzedzać pobranie etykiet");
        mc->ctx->set_device();
        DevBuf tmp;
        tmp.alloc(mc->n * 4);
        mc_launch_labels(mc->P, mc->spin.p, mc->spin_bytes, mc->parent.as<uint32_t>(), tmp.as<uint32_t>(),
                         mc->ctx->stream);
        check(hipMemcpyAsync(root, tmp.p, mc->n * 4, hipMemcpyDeviceToHost, mc->ctx->stream), "labels");
        check(hipStreamSynchronize(mc->ctx->stream), "labels");
    });
}

extern "C" int vca_mc_info(vca_mc* mc, int32_t* colors, uint64_t* domain_cells) {
    return guarded([&] {
        require(mc, "mc == NULL");
        if (colors) *colors = mc->colors;
        if (domain_cells) *domain_cells = mc->ndomain;
    });
}

extern "C" int vca_mc_download_mobility(vca_mc* mc, float* mobility) {
    return guarded([&] {
        require(mc && mobility, "NULL argument");
        require(mc->has_mob_cells, "brak mobilności");
        mc->ctx->set_device();
        check(hipMemcpyAsync(mobility, mc->mob_cells.p, mc->n * 4, hipMemcpyDeviceToHost, mc->ctx->stream), "mobility");
        check(hipStreamSynchronize(mc->ctx->stream), "mobility");
    });
}

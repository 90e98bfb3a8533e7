// Host side of the C ABI in include/ropephysics.h: validation, colouring,
// layout, tile packing, the tile-shape autotuner, HIP-graph replay of step
// sequences, and bulk transfers.
//
// Reference counterpart: crs::gpu::Batch (src/gpu/gpu_batch.cpp:121-484). The
// differences are deliberate: rods may differ in topology (ragged CSR instead
// of one replicated prototype), every transfer is one bulk copy per field, the
// batch owns a stream (so synchronisation is per batch, not device-wide), and
// every argument is checked and every HIP status reported.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <cstdlib>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <memory>
#include <set>
#include <string>
#include <vector>

#include "../../include/ropephysics.h"
#include "rp_layout.h"

namespace rp {
int plain_lds_bytes(int nP, int nF, int nS, int nB, int nPrim, int self, int table, int cap);
int tiled_lds_bytes(int nP, int nF, int nS, int nPrim, int self, int table, int cap);
hipError_t launch_plain(const DeviceArrays& d, const StepArgs& sa, size_t lds, hipStream_t s);
hipError_t allow_dynamic_lds(int bytes, int* granted);
hipError_t launch_tiled(int tp, int tc, bool self, const DeviceArrays& d, const StepArgs& sa, const Tile* tiles,
                        const TileSlot* slots, int numTiles, int threads, size_t lds, hipStream_t s);
hipError_t launch_fixed_frames(float* q, const float* src, const float* invInertia4, int f0, int count,
                               hipStream_t s);
}  // namespace rp

namespace {

using namespace rp;

thread_local std::string g_error;

int fail(int code, const std::string& msg) {
    g_error = msg;
    return code;
}

#define RP_HIP(call)                                                                                     \
    do {                                                                                                 \
        const hipError_t e_ = (call);                                                                    \
        if (e_ != hipSuccess) return fail(RP_ERR_DEVICE, std::string(#call) + ": " + hipGetErrorString(e_)); \
    } while (0)

// ------------------------------------------------------------------ tile configurations

struct TileConfig {
    int tp, tc, waves;
    const char* name;
};

// id = index + 1. Every (tp, tc) here is instantiated in rp_kernels.hip.
const TileConfig kConfigs[] = {
    {2, 1, 1, "tp2-tc1-w1"}, {3, 1, 1, "tp3-tc1-w1"}, {3, 2, 1, "tp3-tc2-w1"}, {5, 2, 1, "tp5-tc2-w1"},
    {2, 1, 2, "tp2-tc1-w2"}, {3, 1, 2, "tp3-tc1-w2"}, {3, 2, 2, "tp3-tc2-w2"}, {5, 2, 2, "tp5-tc2-w2"},
    {2, 1, 4, "tp2-tc1-w4"}, {3, 1, 4, "tp3-tc1-w4"}, {1, 1, 2, "tp1-tc1-w2"}, {1, 1, 4, "tp1-tc1-w4"},
};
constexpr int kNumConfigs = int(sizeof(kConfigs) / sizeof(kConfigs[0]));

// ------------------------------------------------------------------ device buffers

struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
    ~DevBuf() {
        if (p) (void)hipFree(p);
    }
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
};

struct PackedTiles {
    int config = 0;  // id
    std::vector<Tile> tiles;
    std::vector<TileSlot> slots;
    int maxLds = 0;
    int lanesBusy = 0;
    DevBuf dTiles, dSlots;
};

struct GraphKey {
    int strategy, config, steps;
    float dt, gx, gy, gz, lin, ang;
    int substeps, iterations, interval;
    bool operator<(const GraphKey& o) const {
        return std::memcmp(this, &o, sizeof(GraphKey)) < 0;
    }
};

}  // namespace

struct rp_batch_impl {
    int numRods = 0;
    std::vector<int> pStart, fStart, sStart, bStart;
    std::vector<RodInfo> rods;
    std::vector<unsigned char> tileable;
    int strategy = RP_STRATEGY_PLAIN;
    int requestedConfig = 0;
    int plainLds = 0;
    size_t ldsLimit = 65536;  // dynamic LDS a workgroup may take
    size_t cuLds = 65536;     // LDS of a compute unit
    hipStream_t stream = nullptr;

    // device
    std::vector<std::unique_ptr<DevBuf>> owned;
    DeviceArrays d{};
    float* stateBackup = nullptr;  // autotuner scratch (x, v, q, omega)
    float* fixedStage = nullptr;
    size_t deviceBytes = 0;

    std::unique_ptr<PackedTiles> tiles;
    bool tuned = false;
    int stepsPerLaunch = 64;
    long long stepsTaken = 0;
    std::map<GraphKey, hipGraphExec_t> graphs;
    std::vector<int> colStart;  // host copy of the colour-start table
    std::vector<std::pair<int, float>> tuneLog;  // (config, ms for the trial) of the last autotune
    bool timing = false;
    bool anySelf = false;  // some rod collides with itself: run the SELF kernel variant
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    bool evPending = false;
    int lastLaunches = 0;
    float lastMs = 0.0f;
    PathRec* paths = nullptr;
    ObsRec* obs = nullptr;
    int* clock = nullptr;

    // Device pointers are baked into captured graphs; drop them when one changes.
    void drop_graphs() {
        if (stream) (void)hipStreamSynchronize(stream);
        for (auto& kv : graphs) (void)hipGraphExecDestroy(kv.second);
        graphs.clear();
    }

    ~rp_batch_impl() {
        for (auto& kv : graphs) (void)hipGraphExecDestroy(kv.second);
        if (ev0) (void)hipEventDestroy(ev0);
        if (ev1) (void)hipEventDestroy(ev1);
        if (stream) (void)hipStreamDestroy(stream);
    }

    template <typename T>
    int alloc(T*& out, size_t count) {
        auto b = std::make_unique<DevBuf>();
        b->bytes = std::max<size_t>(count, 1) * sizeof(T);
        const hipError_t e = hipMalloc(&b->p, b->bytes);
        if (e != hipSuccess)
            return fail(RP_ERR_MEMORY, std::string("hipMalloc of ") + std::to_string(b->bytes) + " bytes failed: " +
                                           hipGetErrorString(e));
        deviceBytes += b->bytes;
        out = static_cast<T*>(b->p);
        owned.push_back(std::move(b));
        return RP_OK;
    }
    template <typename T>
    int upload(T*& out, const std::vector<T>& host) {
        int rc = alloc(out, host.size());
        if (rc) return rc;
        if (!host.empty()) RP_HIP(hipMemcpy(out, host.data(), host.size() * sizeof(T), hipMemcpyHostToDevice));
        return RP_OK;
    }
};

namespace {

// Greedy colouring, lowest free colour in constraint index order: the
// reference's colorConstraints (src/core/coloring.cpp:11-57). `res` lists the
// state ids each constraint writes.
int greedy_colours(int count, const std::vector<std::array<int, 3>>& res, int width, int numIds,
                   std::vector<int>& order, std::vector<int>& start) {
    std::vector<std::vector<char>> used(numIds);
    std::vector<int> colOf(count);
    int nCol = 0;
    for (int k = 0; k < count; ++k) {
        int c = 0;
        for (;; ++c) {
            bool clash = false;
            for (int m = 0; m < width; ++m) {
                const std::vector<char>& u = used[res[k][m]];
                if (c < int(u.size()) && u[c]) {
                    clash = true;
                    break;
                }
            }
            if (!clash) break;
        }
        colOf[k] = c;
        nCol = std::max(nCol, c + 1);
        for (int m = 0; m < width; ++m) {
            std::vector<char>& u = used[res[k][m]];
            if (int(u.size()) <= c) u.resize(c + 1, 0);
            u[c] = 1;
        }
    }
    start.assign(nCol + 1, 0);
    for (int k = 0; k < count; ++k) ++start[colOf[k] + 1];
    for (int c = 0; c < nCol; ++c) start[c + 1] += start[c];
    order.resize(count);
    std::vector<int> cur(start.begin(), start.end() - 1);
    for (int k = 0; k < count; ++k) order[cur[colOf[k]]++] = k;
    return nCol;
}

// Lanes a rod needs under (tp, tc): every particle and frame within tp trips,
// every colour within tc trips.
int group_lanes(const RodInfo& r, const std::vector<int>& colStart, int tp, int tc) {
    int g = std::max((r.nP + tp - 1) / tp, (r.nF + tp - 1) / tp);
    const int* cs = colStart.data() + r.colOff;
    for (int c = 0; c < r.nSCol; ++c) g = std::max(g, (cs[c + 1] - cs[c] + tc - 1) / tc);
    const int* cb = cs + r.nSCol + 1;
    for (int c = 0; c < r.nBCol; ++c) g = std::max(g, (cb[c + 1] - cb[c] + tc - 1) / tc);
    return std::max(g, 1);
}

int rod_tiled_lds(const RodInfo& r) {
    return tiled_lds_bytes(r.nP, r.nF, r.nS, r.nPrim, r.self, r.table, r.selfCap);
}

}  // namespace

namespace {

// Dynamic LDS budget of one workgroup when `perCU` workgroups are to share a
// compute unit's LDS: a 512-byte allocation granule is left for the static
// part.
int lds_budget(const rp_batch_impl& b, int perCU) {
    return std::min(int(b.ldsLimit), (int(b.cuLds / size_t(perCU)) & ~511) - 512);
}

// Pack rods into workgroups of `waves` wavefronts and at most `ldsCap` bytes
// of LDS: best fit over open workgroups, rods taken widest first; a workgroup
// closes when neither lanes nor LDS are left for the narrowest remaining rod.
int pack_tiles(rp_batch_impl& b, int configId, int ldsCap, const std::vector<int>& colStart, PackedTiles& out) {
    const TileConfig& cfg = kConfigs[configId - 1];
    const int laneCap = 64 * cfg.waves;
    std::vector<int> lanes(b.numRods), lds(b.numRods), order(b.numRods);
    for (int r = 0; r < b.numRods; ++r) {
        if (!b.tileable[r]) return RP_ERR_CAPACITY;
        lanes[r] = group_lanes(b.rods[r], colStart, cfg.tp, cfg.tc);
        lds[r] = rod_tiled_lds(b.rods[r]);
        if (lanes[r] > laneCap || lds[r] > ldsCap) return RP_ERR_CAPACITY;
        order[r] = r;
    }
    std::stable_sort(order.begin(), order.end(), [&](int a, int c) { return lanes[a] > lanes[c]; });
    struct Open {
        int lanesUsed, ldsUsed;
        std::vector<int> rods;
    };
    std::vector<Open> bins;
    // Open bins keyed by free lanes, for best fit.
    std::multimap<int, int> byFree;
    for (int r : order) {
        int chosen = -1;
        for (auto it = byFree.lower_bound(lanes[r]); it != byFree.end(); ++it) {
            Open& o = bins[it->second];
            if (o.ldsUsed + lds[r] <= ldsCap) {
                chosen = it->second;
                byFree.erase(it);
                break;
            }
        }
        if (chosen < 0) {
            bins.push_back(Open{0, 0, {}});
            chosen = int(bins.size()) - 1;
        }
        Open& o = bins[chosen];
        o.rods.push_back(r);
        o.lanesUsed += lanes[r];
        o.ldsUsed += lds[r];
        if (laneCap - o.lanesUsed > 0) byFree.emplace(laneCap - o.lanesUsed, chosen);
    }
    out.config = configId;
    out.tiles.clear();
    out.slots.clear();
    out.maxLds = 0;
    out.lanesBusy = 0;
    for (const Open& o : bins) {
        Tile t;
        t.firstSlot = int(out.slots.size());
        t.numSlots = int(o.rods.size());
        int lane = 0, off = 0;
        for (int r : o.rods) {
            out.slots.push_back(TileSlot{r, lane, lanes[r], off});
            lane += lanes[r];
            off += lds[r];
        }
        t.ldsBytes = off;
        t.pad = 0;
        out.maxLds = std::max(out.maxLds, off);
        out.lanesBusy += lane;
        out.tiles.push_back(t);
    }
    return RP_OK;
}

int upload_tiles(rp_batch_impl& b, PackedTiles& t) {
    auto put = [&](DevBuf& buf, const void* src, size_t bytes) -> int {
        if (buf.p) (void)hipFree(buf.p);
        buf.p = nullptr;
        buf.bytes = std::max<size_t>(bytes, 1);
        RP_HIP(hipMalloc(&buf.p, buf.bytes));
        if (bytes) RP_HIP(hipMemcpy(buf.p, src, bytes, hipMemcpyHostToDevice));
        return RP_OK;
    };
    int rc = put(t.dTiles, t.tiles.data(), t.tiles.size() * sizeof(Tile));
    if (rc) return rc;
    rc = put(t.dSlots, t.slots.data(), t.slots.size() * sizeof(TileSlot));
    (void)b;
    return rc;
}

StepArgs step_args(const rp_step_params& p, int steps) {
    StepArgs a;
    a.h = p.dt / float(p.substeps);
    a.invH2 = 1.0f / (a.h * a.h);
    a.gx = p.gravity[0];
    a.gy = p.gravity[1];
    a.gz = p.gravity[2];
    a.linDecay = std::exp(-p.linear_damping * a.h);
    a.angDecay = std::exp(-p.angular_damping * a.h);
    a.substeps = p.substeps;
    a.iterations = p.iterations;
    a.contactInterval = p.contact_interval > 0 ? p.contact_interval : 1;
    a.steps = steps;
    return a;
}

int launch_steps(rp_batch_impl& b, const rp_step_params& p, int steps, int strategy, PackedTiles* t) {
    if (strategy == RP_STRATEGY_PLAIN) {
        const StepArgs a = step_args(p, 1);
        for (int s = 0; s < steps; ++s) RP_HIP(launch_plain(b.d, a, size_t(b.plainLds), b.stream));
        return RP_OK;
    }
    const TileConfig& cfg = kConfigs[t->config - 1];
    int left = steps;
    while (left > 0) {
        const int n = std::min(left, b.stepsPerLaunch);
        const StepArgs a = step_args(p, n);
        RP_HIP(launch_tiled(cfg.tp, cfg.tc, b.anySelf, b.d, a, static_cast<const Tile*>(t->dTiles.p),
                            static_cast<const TileSlot*>(t->dSlots.p), int(t->tiles.size()), 64 * cfg.waves,
                            size_t(t->maxLds), b.stream));
        left -= n;
    }
    return RP_OK;
}

// Device state a trial run changes: positions, velocities, orientations,
// angular velocities, clocks, overflow counters and observers.
struct StateBackup {
    std::vector<std::pair<void*, size_t>> fields;
    void* store = nullptr;
    size_t bytes = 0;
};

size_t backup_bytes(const rp_batch_impl& b) {
    const size_t P = size_t(b.pStart[b.numRods]), F = size_t(b.fStart[b.numRods]), R = size_t(b.numRods);
    return (P * 6 + F * 7) * sizeof(float) + 2 * R * sizeof(int) + (b.obs ? R * sizeof(ObsRec) : 0);
}

int backup_state(rp_batch_impl& b, bool restore) {
    const size_t P = size_t(b.pStart[b.numRods]), F = size_t(b.fStart[b.numRods]), R = size_t(b.numRods);
    char* s = reinterpret_cast<char*>(b.stateBackup);
    std::pair<void*, size_t> fields[] = {{b.d.x, P * 12},          {b.d.v, P * 12},
                                         {b.d.q, F * 16},          {b.d.omega, F * 12},
                                         {b.clock, R * 4},         {b.d.selfOverflow, R * 4},
                                         {b.obs, b.obs ? R * sizeof(ObsRec) : 0}};
    for (const auto& f : fields) {
        if (!f.first || !f.second) continue;
        if (restore)
            RP_HIP(hipMemcpyAsync(f.first, s, f.second, hipMemcpyDeviceToDevice, b.stream));
        else
            RP_HIP(hipMemcpyAsync(s, f.first, f.second, hipMemcpyDeviceToDevice, b.stream));
        s += f.second;
    }
    RP_HIP(hipStreamSynchronize(b.stream));
    return RP_OK;
}

// Time every feasible tile configuration on this batch's own state for a few
// steps, keep the fastest, and restore the state so the tuning leaves no trace.
// Results do not depend on the configuration: each rod's arithmetic is the
// same whatever lanes and workgroup it lands in.
// Time each tile shape (or only `onlyConfig`) at LDS budgets for 1 to 4
// workgroups per compute unit: the budget decides how many workgroups there
// are and how many are resident at once, so a batch whose workgroups spill
// just past a multiple of the resident count pays a whole extra round.
int autotune(rp_batch_impl& b, const rp_step_params& p, int onlyConfig) {
    if (!b.stateBackup) {
        int rc = b.alloc(b.stateBackup, (backup_bytes(b) + sizeof(float) * 2 + 0) / sizeof(float));
        if (rc) return rc;
    }
    int rc = backup_state(b, false);
    if (rc) return rc;
    double best = 1e300;
    std::unique_ptr<PackedTiles> bestTiles;
    hipEvent_t e0, e1;
    RP_HIP(hipEventCreate(&e0));
    RP_HIP(hipEventCreate(&e1));
    const int trialSteps = 8;
    b.tuneLog.clear();
    for (int id = onlyConfig > 0 ? onlyConfig : 1; id <= (onlyConfig > 0 ? onlyConfig : kNumConfigs); ++id)
      for (int perCU = 1, lastTiles = -1; perCU <= 4; ++perCU) {
        auto t = std::make_unique<PackedTiles>();
        if (pack_tiles(b, id, lds_budget(b, perCU), b.colStart, *t) != RP_OK) continue;
        // A budget that does not change the packing needs no second trial.
        if (int(t->tiles.size()) == lastTiles) continue;
        lastTiles = int(t->tiles.size());
        if (upload_tiles(b, *t) != RP_OK) continue;
        // A warm-up launch (code object load), then the best of two timed ones.
        if (launch_steps(b, p, 1, RP_STRATEGY_TILED, t.get()) != RP_OK) continue;
        float ms = 1e30f;
        for (int rep = 0; rep < 2; ++rep) {
            RP_HIP(hipEventRecord(e0, b.stream));
            if (launch_steps(b, p, trialSteps, RP_STRATEGY_TILED, t.get()) != RP_OK) break;
            RP_HIP(hipEventRecord(e1, b.stream));
            RP_HIP(hipEventSynchronize(e1));
            float m = 0;
            RP_HIP(hipEventElapsedTime(&m, e0, e1));
            ms = std::min(ms, m);
        }
        b.tuneLog.push_back(std::make_pair(id, ms));
        if (ms < best) {
            best = ms;
            bestTiles = std::move(t);
        }
      }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    rc = backup_state(b, true);
    if (rc) return rc;
    if (!bestTiles) return fail(RP_ERR_CAPACITY, "no tile configuration fits this batch");
    b.tiles = std::move(bestTiles);
    return RP_OK;
}

}  // namespace

namespace {

int resolve_tiles(rp_batch_impl& b, const rp_step_params& p) {
    if (b.tiles) return RP_OK;
    if (b.requestedConfig > 0) {
        PackedTiles probe;
        const int rc = pack_tiles(b, b.requestedConfig, lds_budget(b, 1), b.colStart, probe);
        if (rc) return fail(rc, std::string("tile configuration ") + kConfigs[b.requestedConfig - 1].name +
                                    " does not fit this batch");
        return autotune(b, p, b.requestedConfig);
    }
    return autotune(b, p, 0);
}

int check_range(const rp_batch_impl& b, int first, int count) {
    if (first < 0 || count < 0 || first + count > b.numRods)
        return fail(RP_ERR_INVALID, "rod range [" + std::to_string(first) + ", " + std::to_string(first + count) +
                                        ") outside batch of " + std::to_string(b.numRods));
    return RP_OK;
}

}  // namespace

extern "C" {

const char* rp_last_error(void) { return g_error.c_str(); }

int rp_tile_config_count(void) { return kNumConfigs; }

const char* rp_tile_config_name(int id) {
    if (id < 1 || id > kNumConfigs) return nullptr;
    return kConfigs[id - 1].name;
}

int rp_device_query(rp_device_info* out) {
    if (!out) return fail(RP_ERR_INVALID, "null output");
    std::memset(out, 0, sizeof(*out));
    int n = 0;
    const hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess || n == 0)
        return fail(RP_ERR_DEVICE, std::string("no usable HIP device: ") +
                                       (e != hipSuccess ? hipGetErrorString(e) : "device count is 0"));
    hipDeviceProp_t prop;
    RP_HIP(hipGetDeviceProperties(&prop, 0));
    std::snprintf(out->name, sizeof(out->name), "%s", prop.name);
    std::snprintf(out->arch, sizeof(out->arch), "%s", prop.gcnArchName);
    out->compute_units = prop.multiProcessorCount;
    out->wavefront_size = prop.warpSize;
    out->lds_per_workgroup = prop.sharedMemPerBlock;
    out->total_global_mem = prop.totalGlobalMem;
    out->clock_khz = prop.clockRate;
    // Longest clamped chain (n + 1 particles, n + 1 frames, n stretch and n
    // bend constraints, two colours of about n / 2 + 1) some configuration
    // takes.
    int best = 0;
    for (int n2 = 4; n2 <= 4096; ++n2) {
        const int lds = tiled_lds_bytes(n2 + 1, n2 + 1, n2, 0, 0, 1, 0);
        if (size_t(lds) > prop.sharedMemPerBlock) break;
        for (const TileConfig& c : kConfigs) {
            const int lanes = std::max((n2 + 1 + c.tp - 1) / c.tp, (n2 / 2 + 1 + c.tc - 1) / c.tc);
            if (lanes <= 64 * c.waves) best = n2;
        }
    }
    out->max_tiled_segments = best;
    return RP_OK;
}

int rp_batch_create(const rp_batch_desc* desc, void** out_batch) {
    if (!desc || !out_batch) return fail(RP_ERR_INVALID, "null argument");
    *out_batch = nullptr;
    const int R = desc->num_rods;
    if (R <= 0) return fail(RP_ERR_INVALID, "num_rods must be positive");
    if (!desc->particle_start || !desc->frame_start || !desc->stretch_start || !desc->bend_start)
        return fail(RP_ERR_INVALID, "missing CSR start arrays");
    if (desc->strategy < RP_STRATEGY_PLAIN || desc->strategy > RP_STRATEGY_AUTO)
        return fail(RP_ERR_INVALID, "unknown strategy");
    if (desc->tile_config < 0 || desc->tile_config > kNumConfigs)
        return fail(RP_ERR_INVALID, "unknown tile configuration");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return fail(RP_ERR_DEVICE, "no usable HIP device");

    auto b = std::make_unique<rp_batch_impl>();
    b->numRods = R;
    b->requestedConfig = desc->tile_config;
    b->pStart.assign(desc->particle_start, desc->particle_start + R + 1);
    b->fStart.assign(desc->frame_start, desc->frame_start + R + 1);
    b->sStart.assign(desc->stretch_start, desc->stretch_start + R + 1);
    b->bStart.assign(desc->bend_start, desc->bend_start + R + 1);
    for (const std::vector<int>* v : {&b->pStart, &b->fStart, &b->sStart, &b->bStart}) {
        if ((*v)[0] != 0) return fail(RP_ERR_INVALID, "CSR start arrays must begin at 0");
        for (int r = 0; r < R; ++r)
            if ((*v)[r + 1] < (*v)[r]) return fail(RP_ERR_INVALID, "CSR start arrays must be non-decreasing");
    }
    const int P = b->pStart[R], F = b->fStart[R], S = b->sStart[R], B = b->bStart[R];
    if (P > 0 && (!desc->inv_mass)) return fail(RP_ERR_INVALID, "missing inv_mass");
    if (F > 0 && (!desc->inv_inertia || !desc->inertia)) return fail(RP_ERR_INVALID, "missing inertia");
    if (S > 0 && (!desc->stretch_p0 || !desc->stretch_p1 || !desc->stretch_seg || !desc->stretch_rest ||
                  !desc->stretch_compliance))
        return fail(RP_ERR_INVALID, "missing stretch constraint arrays");
    if (B > 0 && (!desc->bend_a || !desc->bend_b || !desc->bend_rest || !desc->bend_compliance ||
                  !desc->bend_rest_darboux))
        return fail(RP_ERR_INVALID, "missing bend constraint arrays");
    if (!desc->radius) return fail(RP_ERR_INVALID, "missing radius");
    if (desc->num_worlds < 0 || (desc->num_worlds > 0 && !desc->worlds))
        return fail(RP_ERR_INVALID, "bad world table");
    for (int w = 0; w < desc->num_worlds; ++w) {
        const rp_world_desc& wd = desc->worlds[w];
        if (wd.num_prims < 0 || wd.num_prims > kMaxPrims || (wd.num_prims > 0 && !wd.prims))
            return fail(RP_ERR_INVALID, "world " + std::to_string(w) + ": 0.." + std::to_string(kMaxPrims) +
                                            " primitives supported");
        for (int k = 0; k < wd.num_prims; ++k)
            if (wd.prims[k].type < RP_PLANE || wd.prims[k].type > RP_BOX)
                return fail(RP_ERR_INVALID, "world " + std::to_string(w) + ": unknown primitive type");
        if (wd.self_collision && wd.hash_table_size < 1)
            return fail(RP_ERR_INVALID, "world " + std::to_string(w) + ": hash_table_size must be positive");
    }

    // ---- per-rod colouring, constraint reordering and rod info
    std::vector<int4> sIdx(S), bIdx(B);
    std::vector<float4> sPar(S), bPar(B), bRest4(B);
    std::vector<int> segEnd(2 * size_t(S));
    std::vector<int>& colStart = b->colStart;
    colStart.clear();
    b->rods.resize(R);
    b->tileable.assign(R, 1);
    std::vector<Prim> prims;
    std::vector<int> primOffOfWorld(std::max(desc->num_worlds, 1), 0);
    for (int w = 0; w < desc->num_worlds; ++w) {
        primOffOfWorld[w] = int(prims.size());
        for (int k = 0; k < desc->worlds[w].num_prims; ++k) {
            const rp_primitive& s = desc->worlds[w].prims[k];
            Prim p;
            p.type = s.type;
            for (int c = 0; c < 3; ++c) {
                p.a[c] = s.a[c];
                p.b[c] = s.b[c];
                p.normal[c] = s.normal[c];
                p.half[c] = s.half_extents[c];
            }
            for (int c = 0; c < 4; ++c) p.rot[c] = s.rotation[c];
            p.radius = s.radius;
            p.friction = s.friction;
            prims.push_back(p);
        }
    }
    std::vector<int> order, start;
    std::vector<std::array<int, 3>> res;
    int maxPlainLds = 0;
    for (int r = 0; r < R; ++r) {
        RodInfo& ri = b->rods[r];
        std::memset(&ri, 0, sizeof(ri));
        ri.pStart = b->pStart[r];
        ri.nP = b->pStart[r + 1] - b->pStart[r];
        ri.fStart = b->fStart[r];
        ri.nF = b->fStart[r + 1] - b->fStart[r];
        ri.sStart = b->sStart[r];
        ri.nS = b->sStart[r + 1] - b->sStart[r];
        ri.bStart = b->bStart[r];
        ri.nB = b->bStart[r + 1] - b->bStart[r];
        if (ri.nP < 1) return fail(RP_ERR_INVALID, "rod " + std::to_string(r) + " has no particles");
        ri.radius = desc->radius[r];
        const int w = desc->num_worlds == 0 ? -1 : (desc->world_of_rod ? desc->world_of_rod[r] : 0);
        if (w >= desc->num_worlds) return fail(RP_ERR_INVALID, "rod " + std::to_string(r) + ": world out of range");
        ri.world = w;
        // Stretch: validate, colour, reorder.
        res.resize(ri.nS);
        for (int k = 0; k < ri.nS; ++k) {
            const int g = ri.sStart + k;
            const int a = desc->stretch_p0[g], c = desc->stretch_p1[g], f = desc->stretch_seg[g];
            if (a < 0 || a >= ri.nP || c < 0 || c >= ri.nP || f < 0 || f >= ri.nF)
                return fail(RP_ERR_INVALID, "rod " + std::to_string(r) + ": stretch constraint " +
                                                std::to_string(k) + " index out of range");
            res[k] = {a, c, ri.nP + f};
            segEnd[2 * size_t(g)] = a;
            segEnd[2 * size_t(g) + 1] = c;
        }
        ri.colOff = int(colStart.size());
        ri.nSCol = greedy_colours(ri.nS, res, 3, ri.nP + ri.nF, order, start);
        for (int k = 0; k < ri.nS; ++k) {
            const int src = ri.sStart + order[k], dst = ri.sStart + k;
            sIdx[dst] = make_int4(desc->stretch_p0[src], desc->stretch_p1[src], desc->stretch_seg[src], order[k]);
            sPar[dst] = make_float4(desc->stretch_rest[src], desc->stretch_compliance[3 * size_t(src)],
                                    desc->stretch_compliance[3 * size_t(src) + 1],
                                    desc->stretch_compliance[3 * size_t(src) + 2]);
        }
        colStart.insert(colStart.end(), start.begin(), start.end());
        // Bend.
        res.resize(ri.nB);
        for (int k = 0; k < ri.nB; ++k) {
            const int g = ri.bStart + k;
            const int a = desc->bend_a[g], c = desc->bend_b[g];
            if (a < 0 || a >= ri.nF || c < 0 || c >= ri.nF)
                return fail(RP_ERR_INVALID, "rod " + std::to_string(r) + ": bend constraint " + std::to_string(k) +
                                                " index out of range");
            res[k] = {ri.nP + a, ri.nP + c, ri.nP + c};
        }
        ri.nBCol = greedy_colours(ri.nB, res, 2, ri.nP + ri.nF, order, start);
        for (int k = 0; k < ri.nB; ++k) {
            const int src = ri.bStart + order[k], dst = ri.bStart + k;
            bIdx[dst] = make_int4(desc->bend_a[src], desc->bend_b[src], order[k], 0);
            bPar[dst] = make_float4(desc->bend_rest[src], desc->bend_compliance[3 * size_t(src)],
                                    desc->bend_compliance[3 * size_t(src) + 1],
                                    desc->bend_compliance[3 * size_t(src) + 2]);
            bRest4[dst] = make_float4(desc->bend_rest_darboux[3 * size_t(src)],
                                      desc->bend_rest_darboux[3 * size_t(src) + 1],
                                      desc->bend_rest_darboux[3 * size_t(src) + 2], 0.0f);
        }
        colStart.insert(colStart.end(), start.begin(), start.end());
        if (ri.nSCol > kMaxColours || ri.nBCol > kMaxColours)
            return fail(RP_ERR_CAPACITY, "rod " + std::to_string(r) + " needs more than " +
                                             std::to_string(kMaxColours) + " colours");
        if (ri.nSCol > kTiledColours || ri.nBCol > kTiledColours) b->tileable[r] = 0;
        // Contacts.
        if (w >= 0) {
            const rp_world_desc& wd = desc->worlds[w];
            ri.nPrim = wd.num_prims;
            ri.primOff = primOffOfWorld[w];
            // contactMarginRadii * radius, in FP32 as the CPU path forms it.
            ri.margin = wd.contact_margin_radii * ri.radius;
            ri.self = wd.self_collision ? 1 : 0;
            ri.selfFriction = wd.self_friction;
            ri.table = wd.self_collision ? wd.hash_table_size : 1;
            if (ri.self) {
                // selfCollisionIndexGap (src/core/collision.cpp:67-74), in FP32.
                float lMin = ri.nS > 0 ? desc->stretch_rest[ri.sStart] : 1.0f;
                for (int k = 0; k < ri.nS; ++k) lMin = std::min(lMin, desc->stretch_rest[ri.sStart + k]);
                const int geometric = int(std::floor(2.0f * ri.radius / lMin)) + 2;
                ri.selfGap = std::max(wd.self_skip, geometric);
                ri.selfCap = wd.self_capacity > 0 ? wd.self_capacity : std::max(16, ri.nS / 2);
            }
        } else {
            ri.table = 1;
        }
        maxPlainLds =
            std::max(maxPlainLds, plain_lds_bytes(ri.nP, ri.nF, ri.nS, ri.nB, ri.nPrim, ri.self, ri.table, ri.selfCap));
    }

    hipDeviceProp_t prop;
    RP_HIP(hipGetDeviceProperties(&prop, 0));
    // Workgroups may use the whole LDS of a compute unit; where the runtime
    // refuses the larger dynamic allocation, stay within 64 KiB.
    b->cuLds = prop.sharedMemPerBlock;
    int granted = 0;
    if (allow_dynamic_lds(int(prop.sharedMemPerBlock), &granted) == hipSuccess) {
        b->ldsLimit = size_t(granted);
    } else {
        (void)hipGetLastError();
        b->cuLds = b->ldsLimit = 65536 - 1024;
    }
    b->plainLds = maxPlainLds;
    for (int r = 0; r < R; ++r) b->anySelf = b->anySelf || b->rods[r].self != 0;
    bool allTileable = true;
    for (int r = 0; r < R; ++r) {
        if (b->tileable[r] && size_t(rod_tiled_lds(b->rods[r])) > b->ldsLimit) b->tileable[r] = 0;
        allTileable = allTileable && b->tileable[r];
    }
    int strategy = desc->strategy;
    if (strategy == RP_STRATEGY_AUTO) strategy = allTileable ? RP_STRATEGY_TILED : RP_STRATEGY_PLAIN;
    if (strategy == RP_STRATEGY_TILED && !allTileable)
        return fail(RP_ERR_CAPACITY, "a rod does not fit the TILED strategy (more than 2 colours or LDS budget)");
    if (strategy == RP_STRATEGY_PLAIN && size_t(maxPlainLds) > b->ldsLimit)
        return fail(RP_ERR_CAPACITY, "a rod needs " + std::to_string(maxPlainLds) + " bytes of LDS; the device has " +
                                         std::to_string(b->ldsLimit) + " per workgroup");
    b->strategy = strategy;
    if (const char* e = std::getenv("RP_STEPS_PER_LAUNCH")) b->stepsPerLaunch = std::max(1, std::atoi(e));

    // ---- device arrays
    RP_HIP(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking));
    std::vector<float> invMass(desc->inv_mass, desc->inv_mass + P);
    std::vector<float4> iI4(F), I4(F);
    for (int j = 0; j < F; ++j) {
        iI4[j] = make_float4(desc->inv_inertia[3 * size_t(j)], desc->inv_inertia[3 * size_t(j) + 1],
                             desc->inv_inertia[3 * size_t(j) + 2], 0.0f);
        I4[j] = make_float4(desc->inertia[3 * size_t(j)], desc->inertia[3 * size_t(j) + 1],
                            desc->inertia[3 * size_t(j) + 2], 0.0f);
    }
    int rc = 0;
    float *x, *v, *q, *om, *force, *kin, *torque, *cs, *fs;
    int* ovf;
    if ((rc = b->alloc(x, size_t(P) * 3)) || (rc = b->alloc(v, size_t(P) * 3)) || (rc = b->alloc(q, size_t(F) * 4)) ||
        (rc = b->alloc(om, size_t(F) * 3)) || (rc = b->alloc(force, size_t(P) * 3)) ||
        (rc = b->alloc(kin, size_t(P) * 3)) || (rc = b->alloc(torque, size_t(F) * 3)) || (rc = b->alloc(cs, R)) ||
        (rc = b->alloc(fs, R)) || (rc = b->alloc(ovf, R)) || (rc = b->alloc(b->fixedStage, size_t(F) * 4)) ||
        (rc = b->alloc(b->clock, R)))
        return rc;
    RP_HIP(hipMemset(b->clock, 0, size_t(R) * 4));
    RP_HIP(hipMemset(x, 0, size_t(P) * 3 * 4));
    RP_HIP(hipMemset(v, 0, size_t(P) * 3 * 4));
    RP_HIP(hipMemset(om, 0, size_t(F) * 3 * 4));
    RP_HIP(hipMemset(force, 0, size_t(P) * 3 * 4));
    RP_HIP(hipMemset(kin, 0, size_t(P) * 3 * 4));
    RP_HIP(hipMemset(torque, 0, size_t(F) * 3 * 4));
    RP_HIP(hipMemset(ovf, 0, size_t(R) * 4));
    {
        std::vector<float> qid(size_t(F) * 4, 0.0f), ones(R, 1.0f);
        for (int j = 0; j < F; ++j) qid[4 * size_t(j)] = 1.0f;
        RP_HIP(hipMemcpy(q, qid.data(), qid.size() * 4, hipMemcpyHostToDevice));
        RP_HIP(hipMemcpy(cs, ones.data(), R * 4, hipMemcpyHostToDevice));
        RP_HIP(hipMemcpy(fs, ones.data(), R * 4, hipMemcpyHostToDevice));
    }
    DeviceArrays& d = b->d;
    d.x = x;
    d.v = v;
    d.q = q;
    d.omega = om;
    d.force = force;
    d.kinVel = kin;
    d.torque = torque;
    d.complianceScale = cs;
    d.frictionScale = fs;
    d.selfOverflow = ovf;
    d.clock = b->clock;
    d.paths = nullptr;
    d.obs = nullptr;
    d.numRods = R;
    float *dInvMass, *dI4, *dII4, *dsPar, *dbPar, *dbRest;
    int *dsIdx, *dbIdx, *dSegEnd, *dCol;
    RodInfo* dRods;
    Prim* dPrims;
    std::vector<float> sParF(size_t(S) * 4), bParF(size_t(B) * 4), bRestF(size_t(B) * 4), iIF(size_t(F) * 4),
        IF(size_t(F) * 4);
    std::vector<int> sIdxI(size_t(S) * 4), bIdxI(size_t(B) * 4);
    std::memcpy(sParF.data(), sPar.data(), sParF.size() * 4);
    std::memcpy(bParF.data(), bPar.data(), bParF.size() * 4);
    std::memcpy(bRestF.data(), bRest4.data(), bRestF.size() * 4);
    std::memcpy(iIF.data(), iI4.data(), iIF.size() * 4);
    std::memcpy(IF.data(), I4.data(), IF.size() * 4);
    std::memcpy(sIdxI.data(), sIdx.data(), sIdxI.size() * 4);
    std::memcpy(bIdxI.data(), bIdx.data(), bIdxI.size() * 4);
    if ((rc = b->upload(dInvMass, invMass)) || (rc = b->upload(dII4, iIF)) || (rc = b->upload(dI4, IF)) ||
        (rc = b->upload(dsIdx, sIdxI)) || (rc = b->upload(dsPar, sParF)) || (rc = b->upload(dbIdx, bIdxI)) ||
        (rc = b->upload(dbPar, bParF)) || (rc = b->upload(dbRest, bRestF)) || (rc = b->upload(dSegEnd, segEnd)) ||
        (rc = b->upload(dCol, colStart)) || (rc = b->upload(dRods, b->rods)) || (rc = b->upload(dPrims, prims)))
        return rc;
    d.invMass = dInvMass;
    d.invInertia4 = dII4;
    d.inertia4 = dI4;
    d.sIdx = dsIdx;
    d.sPar = dsPar;
    d.bIdx = dbIdx;
    d.bPar = dbPar;
    d.bRest4 = dbRest;
    d.segEnd = dSegEnd;
    d.colStart = dCol;
    d.rods = dRods;
    d.prims = dPrims;
    *out_batch = b.release();
    return RP_OK;
}

void rp_batch_destroy(void* batch) {
    if (!batch) return;
    rp_batch_impl* b = static_cast<rp_batch_impl*>(batch);
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    delete b;
}

static int copy_range(rp_batch_impl& b, int first, int count, float* dev, const float* host, float* hostOut,
                      const std::vector<int>& startArr, int comps) {
    const size_t e0 = size_t(startArr[first]), e1 = size_t(startArr[first + count]);
    const size_t bytes = (e1 - e0) * comps * sizeof(float);
    if (bytes == 0) return RP_OK;
    if (host) RP_HIP(hipMemcpyAsync(dev + e0 * comps, host, bytes, hipMemcpyHostToDevice, b.stream));
    if (hostOut) RP_HIP(hipMemcpyAsync(hostOut, dev + e0 * comps, bytes, hipMemcpyDeviceToHost, b.stream));
    return RP_OK;
}

int rp_batch_set_state(void* batch, int first_rod, int num_rods, const float* x, const float* v, const float* q,
                       const float* omega) {
    if (!batch) return fail(RP_ERR_INVALID, "null batch");
    rp_batch_impl& b = *static_cast<rp_batch_impl*>(batch);
    int rc = check_range(b, first_rod, num_rods);
    if (rc) return rc;
    if ((rc = copy_range(b, first_rod, num_rods, b.d.x, x, nullptr, b.pStart, 3)) ||
        (rc = copy_range(b, first_rod, num_rods, b.d.v, v, nullptr, b.pStart, 3)) ||
        (rc = copy_range(b, first_rod, num_rods, b.d.q, q, nullptr, b.fStart, 4)) ||
        (rc = copy_range(b, first_rod, num_rods, b.d.omega, omega, nullptr, b.fStart, 3)))
        return rc;
    RP_HIP(hipStreamSynchronize(b.stream));
    return RP_OK;
}

int rp_batch_set_loads(void* batch, int first_rod, int num_rods, const float* force, const float* kin_vel,
                       const float* torque, const float* fixed_q) {
    if (!batch) return fail(RP_ERR_INVALID, "null batch");
    rp_batch_impl& b = *static_cast<rp_batch_impl*>(batch);
    int rc = check_range(b, first_rod, num_rods);
    if (rc) return rc;
    if ((rc = copy_range(b, first_rod, num_rods, const_cast<float*>(b.d.force), force, nullptr, b.pStart, 3)) ||
        (rc = copy_range(b, first_rod, num_rods, const_cast<float*>(b.d.kinVel), kin_vel, nullptr, b.pStart, 3)) ||
        (rc = copy_range(b, first_rod, num_rods, const_cast<float*>(b.d.torque), torque, nullptr, b.fStart, 3)))
        return rc;
    if (fixed_q) {
        const int f0 = b.fStart[first_rod], f1 = b.fStart[first_rod + num_rods];
        if (f1 > f0) {
            RP_HIP(hipMemcpyAsync(b.fixedStage, fixed_q, size_t(f1 - f0) * 16, hipMemcpyHostToDevice, b.stream));
            RP_HIP(launch_fixed_frames(b.d.q, b.fixedStage, b.d.invInertia4, f0, f1 - f0, b.stream));
        }
    }
    RP_HIP(hipStreamSynchronize(b.stream));
    return RP_OK;
}

int rp_batch_set_scales(void* batch, const float* youngs_scale, const float* friction_scale) {
    if (!batch) return fail(RP_ERR_INVALID, "null batch");
    rp_batch_impl& b = *static_cast<rp_batch_impl*>(batch);
    if (youngs_scale) {
        std::vector<float> c(b.numRods);
        for (int r = 0; r < b.numRods; ++r) {
            if (!(youngs_scale[r] > 0.0f))
                return fail(RP_ERR_INVALID, "youngs_scale[" + std::to_string(r) + "] must be positive");
            // Compliance scales as the inverse of Young's modulus
            // (src/gpu/gpu_batch.cpp:437-439).
            c[r] = 1.0f / youngs_scale[r];
        }
        RP_HIP(hipMemcpyAsync(b.d.complianceScale, c.data(), b.numRods * 4, hipMemcpyHostToDevice, b.stream));
        RP_HIP(hipStreamSynchronize(b.stream));
    }
    if (friction_scale) {
        RP_HIP(hipMemcpyAsync(b.d.frictionScale, friction_scale, b.numRods * 4, hipMemcpyHostToDevice, b.stream));
        RP_HIP(hipStreamSynchronize(b.stream));
    }
    return RP_OK;
}

int rp_batch_step(void* batch, const rp_step_params* params, int num_steps) {
    if (!batch || !params) return fail(RP_ERR_INVALID, "null argument");
    rp_batch_impl& b = *static_cast<rp_batch_impl*>(batch);
    const rp_step_params& p = *params;
    if (num_steps < 0) return fail(RP_ERR_INVALID, "num_steps must be non-negative");
    if (p.substeps < 1 || p.iterations < 0 || !(p.dt > 0.0f))
        return fail(RP_ERR_INVALID, "need substeps >= 1, iterations >= 0 and dt > 0");
    if (num_steps == 0) return RP_OK;
    int rc = 0;
    if (b.strategy == RP_STRATEGY_TILED && (rc = resolve_tiles(b, p))) return rc;
    // Steps run in chunks of stepsPerLaunch: one TILED launch, or that many
    // PLAIN launches. Each chunk shape is captured once into a HIP graph
    // (cached per parameter set) and replayed.
    const int chunk = b.stepsPerLaunch;
    const int full = num_steps / chunk, rem = num_steps % chunk;
    hipGraphExec_t gFull = nullptr, gRem = nullptr;
    if (b.graphs.size() > 62) {  // bound the cache; a call adds at most two graphs
        RP_HIP(hipStreamSynchronize(b.stream));
        for (auto& kv : b.graphs) (void)hipGraphExecDestroy(kv.second);
        b.graphs.clear();
    }
    for (int pass = 0; pass < 2; ++pass) {
        const int n = pass == 0 ? chunk : rem;
        if ((pass == 0 && full == 0) || n == 0) continue;
        GraphKey key;
        std::memset(&key, 0, sizeof(key));
        key.strategy = b.strategy;
        key.config = b.tiles ? b.tiles->config : 0;
        key.steps = n;
        key.dt = p.dt;
        key.gx = p.gravity[0];
        key.gy = p.gravity[1];
        key.gz = p.gravity[2];
        key.lin = p.linear_damping;
        key.ang = p.angular_damping;
        key.substeps = p.substeps;
        key.iterations = p.iterations;
        key.interval = p.contact_interval;
        auto it = b.graphs.find(key);
        if (it == b.graphs.end()) {
            RP_HIP(hipStreamBeginCapture(b.stream, hipStreamCaptureModeThreadLocal));
            rc = launch_steps(b, p, n, b.strategy, b.tiles.get());
            hipGraph_t captured = nullptr;
            const hipError_t ec = hipStreamEndCapture(b.stream, &captured);
            if (rc) {
                if (captured) (void)hipGraphDestroy(captured);
                return rc;
            }
            if (ec != hipSuccess) return fail(RP_ERR_DEVICE, std::string("graph capture: ") + hipGetErrorString(ec));
            hipGraphExec_t exec;
            const hipError_t ei = hipGraphInstantiate(&exec, captured, nullptr, nullptr, 0);
            (void)hipGraphDestroy(captured);
            if (ei != hipSuccess)
                return fail(RP_ERR_DEVICE, std::string("graph instantiate: ") + hipGetErrorString(ei));
            it = b.graphs.emplace(key, exec).first;
        }
        (pass == 0 ? gFull : gRem) = it->second;
    }
    if (b.timing) RP_HIP(hipEventRecord(b.ev0, b.stream));
    for (int i = 0; i < full; ++i) RP_HIP(hipGraphLaunch(gFull, b.stream));
    if (rem) RP_HIP(hipGraphLaunch(gRem, b.stream));
    if (b.timing) {
        RP_HIP(hipEventRecord(b.ev1, b.stream));
        b.evPending = true;
    }
    b.lastLaunches = b.strategy == RP_STRATEGY_PLAIN ? num_steps
                                                     : (num_steps + b.stepsPerLaunch - 1) / b.stepsPerLaunch;
    b.stepsTaken += num_steps;
    return RP_OK;
}

int rp_batch_set_paths(void* batch, int first_rod, int num_rods, const rp_path_desc* paths) {
    if (!batch) return fail(RP_ERR_INVALID, "null batch");
    rp_batch_impl& b = *static_cast<rp_batch_impl*>(batch);
    int rc = check_range(b, first_rod, num_rods);
    if (rc) return rc;
    if (!b.paths) {
        if ((rc = b.alloc(b.paths, b.numRods))) return rc;
        std::vector<PathRec> none(b.numRods);
        for (PathRec& p : none) {
            std::memset(&p, 0, sizeof(p));
            p.particle = -1;
        }
        RP_HIP(hipMemcpy(b.paths, none.data(), none.size() * sizeof(PathRec), hipMemcpyHostToDevice));
        b.drop_graphs();
        b.d.paths = b.paths;
    }
    std::vector<PathRec> recs(num_rods);
    for (int i = 0; i < num_rods; ++i) {
        PathRec& p = recs[i];
        std::memset(&p, 0, sizeof(p));
        p.particle = -1;
        if (!paths || paths[i].particle < 0) continue;
        const rp_path_desc& s = paths[i];
        const RodInfo& ri = b.rods[first_rod + i];
        if (s.particle >= ri.nP)
            return fail(RP_ERR_INVALID, "path " + std::to_string(i) + ": particle out of range");
        if (s.num_waypoints < 1 || s.num_waypoints > kMaxWaypoints || !s.waypoints || !s.leg_seconds)
            return fail(RP_ERR_INVALID, "path " + std::to_string(i) + ": 1.." + std::to_string(kMaxWaypoints) +
                                            " waypoints with leg durations required");
        if (!(s.control_period > 0.0) || s.steps_per_control < 1)
            return fail(RP_ERR_INVALID, "path " + std::to_string(i) + ": control period and steps per control "
                                            "must be positive");
        p.particle = s.particle;
        p.nWay = s.num_waypoints;
        p.stepsPerControl = s.steps_per_control;
        p.height = s.height;
        p.period = s.control_period;
        p.sx = s.start[0];
        p.sy = s.start[1];
        for (int k = 0; k < s.num_waypoints; ++k) {
            if (!(s.leg_seconds[k] > 0.0))
                return fail(RP_ERR_INVALID, "path " + std::to_string(i) + ": leg durations must be positive");
            p.legs[k] = s.leg_seconds[k];
            p.wx[k] = s.waypoints[2 * k];
            p.wy[k] = s.waypoints[2 * k + 1];
        }
    }
    RP_HIP(hipMemcpyAsync(b.paths + first_rod, recs.data(), recs.size() * sizeof(PathRec), hipMemcpyHostToDevice,
                          b.stream));
    RP_HIP(hipStreamSynchronize(b.stream));
    return RP_OK;
}

int rp_batch_set_clock(void* batch, int first_rod, int num_rods, const int* clock) {
    if (!batch) return fail(RP_ERR_INVALID, "null batch");
    rp_batch_impl& b = *static_cast<rp_batch_impl*>(batch);
    int rc = check_range(b, first_rod, num_rods);
    if (rc) return rc;
    std::vector<int> v(num_rods, 0);
    if (clock)
        for (int i = 0; i < num_rods; ++i) {
            if (clock[i] < 0) return fail(RP_ERR_INVALID, "clock values must be non-negative");
            v[i] = clock[i];
        }
    RP_HIP(hipMemcpyAsync(b.clock + first_rod, v.data(), v.size() * 4, hipMemcpyHostToDevice, b.stream));
    RP_HIP(hipStreamSynchronize(b.stream));
    return RP_OK;
}

int rp_batch_get_clock(void* batch, int first_rod, int num_rods, int* clock) {
    if (!batch || !clock) return fail(RP_ERR_INVALID, "null argument");
    rp_batch_impl& b = *static_cast<rp_batch_impl*>(batch);
    int rc = check_range(b, first_rod, num_rods);
    if (rc) return rc;
    RP_HIP(hipMemcpyAsync(clock, b.clock + first_rod, size_t(num_rods) * 4, hipMemcpyDeviceToHost, b.stream));
    RP_HIP(hipStreamSynchronize(b.stream));
    return RP_OK;
}

int rp_batch_set_observers(void* batch, int first_rod, int num_rods, const rp_observer_desc* observers) {
    if (!batch) return fail(RP_ERR_INVALID, "null batch");
    rp_batch_impl& b = *static_cast<rp_batch_impl*>(batch);
    int rc = check_range(b, first_rod, num_rods);
    if (rc) return rc;
    if (!b.obs) {
        if ((rc = b.alloc(b.obs, b.numRods))) return rc;
        std::vector<ObsRec> none(b.numRods, ObsRec{-1, 1, 0.0, 0.0});
        RP_HIP(hipMemcpy(b.obs, none.data(), none.size() * sizeof(ObsRec), hipMemcpyHostToDevice));
        b.drop_graphs();
        b.d.obs = b.obs;
    }
    std::vector<ObsRec> recs(num_rods, ObsRec{-1, 1, 0.0, 0.0});
    for (int i = 0; observers && i < num_rods; ++i) {
        const rp_observer_desc& o = observers[i];
        if (o.particle < 0) continue;
        if (o.particle >= b.rods[first_rod + i].nP || o.every < 1)
            return fail(RP_ERR_INVALID, "observer " + std::to_string(i) + ": bad particle or interval");
        recs[i] = ObsRec{o.particle, o.every, o.z0, 0.0};
    }
    RP_HIP(hipMemcpyAsync(b.obs + first_rod, recs.data(), recs.size() * sizeof(ObsRec), hipMemcpyHostToDevice,
                          b.stream));
    RP_HIP(hipStreamSynchronize(b.stream));
    return RP_OK;
}

int rp_batch_get_observations(void* batch, int first_rod, int num_rods, double* max_rise) {
    if (!batch || !max_rise) return fail(RP_ERR_INVALID, "null argument");
    rp_batch_impl& b = *static_cast<rp_batch_impl*>(batch);
    int rc = check_range(b, first_rod, num_rods);
    if (rc) return rc;
    if (!b.obs) return fail(RP_ERR_STATE, "no observers installed");
    std::vector<ObsRec> recs(num_rods);
    RP_HIP(hipMemcpyAsync(recs.data(), b.obs + first_rod, recs.size() * sizeof(ObsRec), hipMemcpyDeviceToHost,
                          b.stream));
    RP_HIP(hipStreamSynchronize(b.stream));
    for (int i = 0; i < num_rods; ++i) max_rise[i] = recs[i].maxRise;
    return RP_OK;
}

int rp_batch_tune_report(void* batch, int* ids, float* ms, int capacity) {
    if (!batch) return fail(RP_ERR_INVALID, "null batch");
    rp_batch_impl& b = *static_cast<rp_batch_impl*>(batch);
    const int n = int(b.tuneLog.size());
    for (int i = 0; i < n && i < capacity; ++i) {
        if (ids) ids[i] = b.tuneLog[i].first;
        if (ms) ms[i] = b.tuneLog[i].second;
    }
    return n;
}

int rp_batch_set_timing(void* batch, int enable) {
    if (!batch) return fail(RP_ERR_INVALID, "null batch");
    rp_batch_impl& b = *static_cast<rp_batch_impl*>(batch);
    if (enable && !b.ev0) {
        RP_HIP(hipEventCreate(&b.ev0));
        RP_HIP(hipEventCreate(&b.ev1));
    }
    b.timing = enable != 0;
    return RP_OK;
}

int rp_batch_sync(void* batch) {
    if (!batch) return fail(RP_ERR_INVALID, "null batch");
    rp_batch_impl& b = *static_cast<rp_batch_impl*>(batch);
    RP_HIP(hipStreamSynchronize(b.stream));
    return RP_OK;
}

int rp_batch_get_state(void* batch, int first_rod, int num_rods, float* x, float* v, float* q, float* omega) {
    if (!batch) return fail(RP_ERR_INVALID, "null batch");
    rp_batch_impl& b = *static_cast<rp_batch_impl*>(batch);
    int rc = check_range(b, first_rod, num_rods);
    if (rc) return rc;
    if ((rc = copy_range(b, first_rod, num_rods, b.d.x, nullptr, x, b.pStart, 3)) ||
        (rc = copy_range(b, first_rod, num_rods, b.d.v, nullptr, v, b.pStart, 3)) ||
        (rc = copy_range(b, first_rod, num_rods, b.d.q, nullptr, q, b.fStart, 4)) ||
        (rc = copy_range(b, first_rod, num_rods, b.d.omega, nullptr, omega, b.fStart, 3)))
        return rc;
    RP_HIP(hipStreamSynchronize(b.stream));
    return RP_OK;
}

int rp_batch_counters(void* batch, rp_counters* out) {
    if (!batch || !out) return fail(RP_ERR_INVALID, "null argument");
    rp_batch_impl& b = *static_cast<rp_batch_impl*>(batch);
    std::memset(out, 0, sizeof(*out));
    std::vector<int> ovf(b.numRods);
    RP_HIP(hipStreamSynchronize(b.stream));
    RP_HIP(hipMemcpy(ovf.data(), b.d.selfOverflow, b.numRods * 4, hipMemcpyDeviceToHost));
    for (int v : ovf) out->self_contact_overflow += v;
    if (b.timing && b.evPending) {
        RP_HIP(hipEventElapsedTime(&b.lastMs, b.ev0, b.ev1));
        b.evPending = false;
    }
    out->last_step_gpu_ms = b.lastMs;
    out->last_step_launches = b.lastLaunches;
    out->strategy = b.strategy;
    out->steps_taken = b.stepsTaken;
    out->device_bytes = b.deviceBytes;
    if (b.strategy == RP_STRATEGY_PLAIN) {
        out->workgroups = b.numRods;
        out->lanes_total = b.numRods * 128;
        out->lds_bytes_per_workgroup = size_t(b.plainLds);
        out->launches_per_step = 1;
    } else if (b.tiles) {
        const TileConfig& c = kConfigs[b.tiles->config - 1];
        out->tile_config = b.tiles->config;
        out->workgroups = int(b.tiles->tiles.size());
        out->lanes_busy = b.tiles->lanesBusy;
        out->lanes_total = out->workgroups * 64 * c.waves;
        out->lds_bytes_per_workgroup = size_t(b.tiles->maxLds);
        out->launches_per_step = 0;
    }
    return RP_OK;
}

}  // extern "C"

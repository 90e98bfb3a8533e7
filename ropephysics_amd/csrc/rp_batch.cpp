// Host side of the C ABI in include/ropephysics.h: validation, colouring,
// layout, tile packing, the tile-shape autotuner, HIP-graph replay of step
// sequences, and bulk transfers, written once over the scalar for the FP32
// and FP64 batches.
//
// Reference counterpart: crs::gpu::Batch (src/gpu/gpu_batch.cpp:121-484). The
// differences are deliberate: rods may differ in topology (ragged CSR instead
// of one replicated prototype), every transfer is one bulk copy per field, the
// batch owns a stream (so synchronisation is per batch, not device-wide), and
// every argument is checked and every HIP status reported.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <memory>
#include <string>
#include <vector>

#include "rp_host.h"

namespace {

using namespace rp;
using namespace rp::host;

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

struct PackedTiles : Packing {
    DevBuf dTiles, dSlots;
    DevBuf dProgress;  // one step count per workgroup (launch pairing)
};

struct GraphKey {
    int strategy, config, steps;
    double dt, gx, gy, gz, lin, ang;
    int substeps, iterations, interval;
    bool operator<(const GraphKey& o) const {
        return std::memcmp(this, &o, sizeof(GraphKey)) < 0;
    }
};

}  // namespace

// What a batch holds whatever its precision.
struct rp_batch_impl {
    int bits = 32;  // arithmetic precision
    int device = 0;  // current when the batch was created
    int numRods = 0;
    std::vector<int> pStart, fStart, sStart, bStart;
    std::vector<unsigned char> tileable;
    int strategy = RP_STRATEGY_PLAIN;
    int requestedConfig = 0;
    int plainLds = 0;
    size_t ldsLimit = 65536;  // dynamic LDS a workgroup may take
    size_t cuLds = 65536;     // LDS of a compute unit
    int computeUnits = 0;
    hipStream_t stream = nullptr;

    // device
    std::vector<std::unique_ptr<DevBuf>> owned;
    void* stateBackup = nullptr;  // autotuner scratch (x, v, q, omega, clocks, counters, observers)
    size_t deviceBytes = 0;
    int* selfOverflow = nullptr;
    int* clock = nullptr;
    PathRec* paths = nullptr;
    ObsRec* obs = nullptr;
    // Trajectory recording: per-rod records, the FP32 frame buffer and a host
    // copy of each rod's origin clock.
    RecRec* rec = nullptr;
    std::unique_ptr<DevBuf> frameBuf;
    int recEvery = 0, recMaxFrames = 0;
    std::vector<int> recOrigin;

    std::unique_ptr<PackedTiles> tiles;
    // TILED: the most steps one launch runs. A launch lasts as long as its
    // slowest workgroup, so a call's steps go into as few launches as this
    // allows: the bursts of work of different rods (self-contact folds, stretched
    // segments) then overlap instead of each setting one short launch's length.
    int stepsPerLaunch = 8192;
    bool stepsPerLaunchSet = false;  // set from RP_STEPS_PER_LAUNCH: no time bound
    // Device time of one substep sweep of the whole batch in the autotuner's
    // best trial: a launch is also held to about kLaunchBudgetMs by it.
    double tileMsPerSweep = 0.0;
    // PLAIN: launches (one step each) captured per graph.
    int plainStepsPerGraph = 64;
    // GLOBAL: lanes per workgroup, per-rod offsets of the HBM working arrays
    // and their total; the arrays are allocated at the first step.
    int globalThreads = 128;
    int globalXqLds = 0;  // dynamic LDS holding positions and orientations (0: they stay in HBM)
    std::vector<long long> scratchOffHost;
    size_t scratchBytes = 0;
    void* scratch = nullptr;
    long long* scratchOff = nullptr;
    // FP32 motion margin: the largest coordinate magnitude ever uploaded, and
    // over the steps taken the smallest |g| h^2 / spacing at it and the steps
    // taken below the reference's bar of 2.
    double coordScale = 0.0;
    double minMargin = 0.0;
    long long lowMarginSteps = 0;
    long long stepsTaken = 0;
    std::map<GraphKey, hipGraphExec_t> graphs;
    std::vector<int> colStart;  // host copy of the colour-start table
    std::vector<std::pair<int, float>> tuneLog;  // (config, ms for the trial) of the last autotune
    bool timing = false;
    bool anySelf = false;      // some rod collides with itself: run a self-collision kernel variant
    bool anyColoured = false;  // ... and some asks for the coloured self order: the variant that has it
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    bool evPending = false;
    int lastLaunches = 0;
    float lastMs = 0.0f;

    // Device pointers are baked into captured graphs; drop them when one changes.
    void drop_graphs() {
        if (stream) (void)hipStreamSynchronize(stream);
        for (auto& kv : graphs) (void)hipGraphExecDestroy(kv.second);
        graphs.clear();
    }

    // Point the kernels' argument block at the path and observer tables.
    virtual void bind_tables() = 0;

    virtual ~rp_batch_impl() {
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

// A batch whose kernels compute in Real.
template <typename Real>
struct BatchT final : rp_batch_impl {
    DeviceArraysT<Real> d{};
    std::vector<RodInfoT<Real>> rods;
    Real* fixedStage = nullptr;

    BatchT() { bits = Prec<Real>::bits; }
    void bind_tables() override {
        d.paths = paths;
        d.obs = obs;
        d.rec = rec;
        d.frames = frameBuf ? static_cast<float*>(frameBuf->p) : nullptr;
        d.framePitch = 3LL * pStart[numRods];
    }
};

// The batch behind a handle, which must compute in Real.
template <typename Real>
int as_precision(void* batch, BatchT<Real>*& out) {
    out = nullptr;
    if (!batch) return fail(RP_ERR_INVALID, "null batch");
    rp_batch_impl* base = static_cast<rp_batch_impl*>(batch);
    if (base->bits != Prec<Real>::bits)
        return fail(RP_ERR_STATE, "an FP" + std::to_string(base->bits) + " batch takes the " +
                                      (base->bits == 64 ? "*64 calls" : "FP32 calls") + " for state and steps");
    out = static_cast<BatchT<Real>*>(base);
    return RP_OK;
}

int lds_budget(const rp_batch_impl& b, int perCU) { return lds_budget(LdsLimits{b.ldsLimit, b.cuLds}, perCU); }

// All rods of the batch packed at one LDS budget; rods that cannot run TILED
// make it fail.
template <typename Real>
int pack_tiles(BatchT<Real>& b, int configId, int ldsCap, PackedTiles& out) {
    if (!tile_config_available<Real>(configId)) return RP_ERR_CAPACITY;
    std::vector<int> ids(b.numRods);
    for (int r = 0; r < b.numRods; ++r) {
        if (!b.tileable[r]) return RP_ERR_CAPACITY;
        ids[r] = r;
    }
    const int rc = pack_rods(b.rods, ids, b.colStart, configId, ldsCap, out);
    if (rc != RP_OK) return rc;
    const TileConfig& cfg = tile_config(configId);
    const int selfMode = b.anySelf ? (b.anyColoured ? 2 : 1) : 0;
    const int resident = Prec<Real>::tiled_resident(cfg.tp, cfg.tc, selfMode, 64 * cfg.waves, size_t(out.maxLds));
    out.pairUnits = pair_units(int(out.tiles.size()), b.computeUnits, resident);
    place_tiles(out, out.pairUnits);
    return RP_OK;
}

int upload_tiles(PackedTiles& t) {
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
    if ((rc = put(t.dSlots, t.slots.data(), t.slots.size() * sizeof(TileSlot)))) return rc;
    const std::vector<int> zeros(t.tiles.size(), 0);
    return put(t.dProgress, zeros.data(), zeros.size() * sizeof(int));
}

// Estimated device time one TILED launch is held to.
constexpr double kLaunchBudgetMs = 4000.0;

template <typename Real>
int launch_steps(BatchT<Real>& b, const typename Prec<Real>::StepParams& p, int steps, int strategy,
                 PackedTiles* t) {
    if (strategy == RP_STRATEGY_PLAIN) {
        const StepArgsT<Real> a = step_args<Real>(p, 1);
        for (int s = 0; s < steps; ++s) RP_HIP(Prec<Real>::launch_plain(b.d, a, size_t(b.plainLds), b.stream));
        return RP_OK;
    }
    if (strategy == RP_STRATEGY_GLOBAL) {
        for (int left = steps; left > 0;) {
            const int n = std::min(left, b.stepsPerLaunch);
            RP_HIP(Prec<Real>::launch_global(b.d, step_args<Real>(p, n), b.globalThreads, size_t(b.globalXqLds),
                                             b.stream));
            left -= n;
        }
        return RP_OK;
    }
    const TileConfig& cfg = tile_config(t->config);
    int left = steps;
    while (left > 0) {
        const int n = std::min(left, b.stepsPerLaunch);
        StepArgsT<Real> a = step_args<Real>(p, n);
        a.progress = static_cast<int*>(t->dProgress.p);
        a.pairUnits = t->pairUnits;
        const int selfMode = b.anySelf ? (b.anyColoured ? 2 : 1) : 0;
        RP_HIP(Prec<Real>::launch_tiled(cfg.tp, cfg.tc, selfMode, b.d, a, static_cast<const Tile*>(t->dTiles.p),
                                        static_cast<const TileSlot*>(t->dSlots.p), int(t->tiles.size()),
                                        64 * cfg.waves, size_t(t->maxLds), b.stream));
        left -= n;
    }
    return RP_OK;
}

// Device state a trial run changes: positions, velocities, orientations,
// angular velocities, clocks, overflow counters and observers.
template <typename Real>
size_t backup_bytes(const BatchT<Real>& b) {
    const size_t P = size_t(b.pStart[b.numRods]), F = size_t(b.fStart[b.numRods]), R = size_t(b.numRods);
    return (P * 6 + F * 7) * sizeof(Real) + 2 * R * sizeof(int) + (b.obs ? R * sizeof(ObsRec) : 0);
}

template <typename Real>
int backup_state(BatchT<Real>& b, bool restore) {
    const size_t P = size_t(b.pStart[b.numRods]), F = size_t(b.fStart[b.numRods]), R = size_t(b.numRods);
    const size_t w = sizeof(Real);
    char* s = static_cast<char*>(b.stateBackup);
    std::pair<void*, size_t> fields[] = {{b.d.x, P * 3 * w},  {b.d.v, P * 3 * w},         {b.d.q, F * 4 * w},
                                         {b.d.omega, F * 3 * w}, {b.clock, R * 4},       {b.selfOverflow, R * 4},
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

// Time each tile shape (or only `onlyConfig`) at LDS budgets for 1 to 4
// workgroups per compute unit on this batch's own state for a few steps, keep
// the fastest, and restore the state so the tuning leaves no trace. The
// budget decides how many workgroups there are and how many are resident at
// once, so a batch whose workgroups spill just past a multiple of the
// resident count pays a whole extra round. Results do not depend on the
// shape: each rod's arithmetic is the same whatever lanes and workgroup it
// lands in.
template <typename Real>
int autotune(BatchT<Real>& b, const typename Prec<Real>::StepParams& p, int onlyConfig) {
    if (!b.stateBackup) {
        char* store = nullptr;
        int rc = b.alloc(store, backup_bytes(b));
        if (rc) return rc;
        b.stateBackup = store;
    }
    int rc = backup_state(b, false);
    if (rc) return rc;
    // Trial steps record no frames.
    const RecRec* recording = b.d.rec;
    b.d.rec = nullptr;
    double best = 1e300, second = 1e300;
    std::unique_ptr<PackedTiles> bestTiles, secondTiles;
    hipEvent_t e0, e1;
    RP_HIP(hipEventCreate(&e0));
    RP_HIP(hipEventCreate(&e1));
    const int trialSteps = 8;
    b.tuneLog.clear();
    const int first = onlyConfig > 0 ? onlyConfig : 1, last = onlyConfig > 0 ? onlyConfig : num_tile_configs();
    for (int id = first; id <= last && rc == RP_OK; ++id)
        for (int perCU = 1, lastTiles = -1; perCU <= 4; ++perCU) {
            auto t = std::make_unique<PackedTiles>();
            if (pack_tiles(b, id, lds_budget(b, perCU), *t) != RP_OK) continue;
            // A budget that does not change the packing needs no second trial.
            if (int(t->tiles.size()) == lastTiles) continue;
            lastTiles = int(t->tiles.size());
            if (upload_tiles(*t) != RP_OK) continue;
            // A warm-up launch (code object load), then the best of two timed
            // ones. Every timed launch starts from the saved state, so each
            // shape is timed over the same steps.
            if (launch_steps(b, p, 1, RP_STRATEGY_TILED, t.get()) != RP_OK) continue;
            float ms = 1e30f;
            for (int rep = 0; rep < 2 && rc == RP_OK; ++rep) {
                if ((rc = backup_state(b, true))) break;
                RP_HIP(hipEventRecord(e0, b.stream));
                if (launch_steps(b, p, trialSteps, RP_STRATEGY_TILED, t.get()) != RP_OK) break;
                RP_HIP(hipEventRecord(e1, b.stream));
                RP_HIP(hipEventSynchronize(e1));
                float m = 0;
                RP_HIP(hipEventElapsedTime(&m, e0, e1));
                ms = std::min(ms, m);
            }
            if (rc) break;
            b.tuneLog.push_back(std::make_pair(id, ms));
            if (ms < best) {
                second = best;
                secondTiles = std::move(bestTiles);
                best = ms;
                bestTiles = std::move(t);
            } else if (ms < second) {
                second = ms;
                secondTiles = std::move(t);
            }
        }
    // Two shapes within 10% of each other in the short trials: time both over
    // four times as many steps (best of two) and keep the faster.
    if (rc == RP_OK && bestTiles && secondTiles && second < 1.1 * best) {
        float ms[2] = {1e30f, 1e30f};
        PackedTiles* cand[2] = {bestTiles.get(), secondTiles.get()};
        for (int c = 0; c < 2; ++c) {
            if (upload_tiles(*cand[c]) != RP_OK) continue;
            for (int rep = 0; rep < 2 && rc == RP_OK; ++rep) {
                if ((rc = backup_state(b, true))) break;
                RP_HIP(hipEventRecord(e0, b.stream));
                if (launch_steps(b, p, 4 * trialSteps, RP_STRATEGY_TILED, cand[c]) != RP_OK) break;
                RP_HIP(hipEventRecord(e1, b.stream));
                RP_HIP(hipEventSynchronize(e1));
                float m = 0;
                RP_HIP(hipEventElapsedTime(&m, e0, e1));
                ms[c] = std::min(ms[c], m);
            }
            b.tuneLog.push_back(std::make_pair(cand[c]->config, ms[c] / 4.0f));
        }
        if (ms[1] < ms[0]) std::swap(bestTiles, secondTiles);
        if (std::min(ms[0], ms[1]) < 1e30f) best = std::min(ms[0], ms[1]) / 4.0;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    b.d.rec = recording;
    if (rc) return rc;
    rc = backup_state(b, true);
    if (rc) return rc;
    if (!bestTiles) return fail(RP_ERR_CAPACITY, "no tile configuration fits this batch");
    b.tiles = std::move(bestTiles);
    b.tileMsPerSweep = best / (double(trialSteps) * p.substeps * std::max(1, p.iterations));
    return RP_OK;
}

template <typename Real>
int resolve_tiles(BatchT<Real>& b, const typename Prec<Real>::StepParams& p) {
    if (b.tiles) return RP_OK;
    if (b.requestedConfig > 0) {
        PackedTiles probe;
        const int rc = pack_tiles(b, b.requestedConfig, lds_budget(b, 1), probe);
        if (rc) return fail(rc, std::string("tile configuration ") + tile_config(b.requestedConfig).name +
                                    " does not fit this batch in FP" + std::to_string(b.bits));
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

template <typename Real>
int create(const typename Prec<Real>::BatchDesc* desc, void** out_batch) {
    if (!desc || !out_batch) return fail(RP_ERR_INVALID, "null argument");
    *out_batch = nullptr;
    if (desc->strategy < RP_STRATEGY_PLAIN || desc->strategy > RP_STRATEGY_GLOBAL)
        return fail(RP_ERR_INVALID, "unknown strategy");
    if (desc->tile_config < 0 || desc->tile_config > num_tile_configs())
        return fail(RP_ERR_INVALID, "unknown tile configuration");
    if (desc->tile_config > 0 && !tile_config_available<Real>(desc->tile_config))
        return fail(RP_ERR_INVALID, std::string("tile configuration ") + tile_config(desc->tile_config).name +
                                        " is not built for FP" + std::to_string(Prec<Real>::bits));
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return fail(RP_ERR_DEVICE, "no usable HIP device");

    RecordsT<Real> rec;
    int rc = build_records<Real>(desc, rec);
    if (rc) return rc;
    const int R = rec.R, P = rec.P, F = rec.F;
    auto b = std::make_unique<BatchT<Real>>();
    RP_HIP(hipGetDevice(&b->device));
    b->numRods = R;
    b->requestedConfig = desc->tile_config;
    b->pStart = rec.pStart;
    b->fStart = rec.fStart;
    b->sStart = rec.sStart;
    b->bStart = rec.bStart;
    b->rods = rec.rods;
    b->colStart = rec.colStart;
    b->tileable = rec.colourFit;
    b->anySelf = rec.anySelf;
    b->anyColoured = rec.anyColoured;
    b->plainLds = rec.maxPlainLds;

    LdsLimits lim;
    if ((rc = lds_limits(lim))) return rc;
    b->ldsLimit = lim.perWorkgroup;
    b->cuLds = lim.perCU;
    b->computeUnits = lim.computeUnits;
    // A rod runs TILED when it fits the LDS budget and some tile shape built
    // for this precision has lanes enough for it.
    bool allTileable = true;
    for (int r = 0; r < R; ++r) {
        // Checked against the budget the packer uses (one workgroup per CU).
        if (b->tileable[r] && rod_tiled_lds(b->rods[r]) > lds_budget(*b, 1)) b->tileable[r] = 0;
        if (b->tileable[r]) {
            bool shape = false;
            for (int id = 1; id <= num_tile_configs() && !shape; ++id) {
                const TileConfig& c = tile_config(id);
                shape = tile_config_available<Real>(id) &&
                        group_lanes(b->rods[r], b->colStart, c.tp, c.tc) <= 64 * c.waves;
            }
            b->tileable[r] = shape ? 1 : 0;
        }
        allTileable = allTileable && b->tileable[r];
    }
    int strategy = desc->strategy;
    // AUTO: the ragged stepper when every rod fits it, else the one-rod-per-
    // workgroup stepper in LDS, else the same in HBM.
    if (strategy == RP_STRATEGY_AUTO)
        strategy = allTileable ? RP_STRATEGY_TILED
                               : (size_t(rec.maxPlainLds) <= b->ldsLimit ? RP_STRATEGY_PLAIN : RP_STRATEGY_GLOBAL);
    if (strategy == RP_STRATEGY_TILED && !allTileable)
        return fail(RP_ERR_CAPACITY,
                    "a rod does not fit the TILED strategy (more than 2 colours, LDS budget or lanes)");
    if (strategy == RP_STRATEGY_PLAIN && size_t(rec.maxPlainLds) > b->ldsLimit)
        return fail(RP_ERR_CAPACITY, "a rod needs " + std::to_string(rec.maxPlainLds) +
                                         " bytes of LDS; the device has " + std::to_string(b->ldsLimit) +
                                         " per workgroup (the GLOBAL strategy takes it)");
    b->strategy = strategy;
    if (strategy == RP_STRATEGY_GLOBAL) {
        // Each rod's working arrays in HBM, 256-byte aligned; the workgroup as
        // wide as the longest rod's elements (64 to 512 lanes).
        b->scratchOffHost.resize(R);
        size_t off = 0;
        int widest = 1;
        for (int r = 0; r < R; ++r) {
            const RodInfoT<Real>& ri = b->rods[r];
            b->scratchOffHost[r] = (long long)off;
            off += (size_t(Prec<Real>::plain_lds_bytes(ri.nP, ri.nF, ri.nS, ri.nB, ri.nPrim, ri.self, ri.table,
                                                      ri.selfCap)) + 255) & ~size_t(255);
            widest = std::max(widest, std::max(ri.nP, ri.nF));
        }
        b->scratchBytes = off;
        b->globalThreads = std::min(Prec<Real>::global_max_threads(), std::max(64, (widest + 63) / 64 * 64));
        // Positions and orientations in LDS when every rod's fit the budget of
        // two workgroups per compute unit (a 2000-segment FP32 rod: 64 KB).
        size_t xq = 0;
        for (int r = 0; r < R; ++r)
            xq = std::max(xq, size_t(b->rods[r].nP + b->rods[r].nF) * 4 * sizeof(Real));
        b->globalXqLds = xq <= size_t(lds_budget(*b, 2)) ? int(xq) : 0;
    }
    if (const char* e = std::getenv("RP_STEPS_PER_LAUNCH")) {
        b->stepsPerLaunch = std::max(1, std::atoi(e));
        b->stepsPerLaunchSet = true;
    }

    // ---- device arrays
    RP_HIP(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking));
    Real *x, *v, *q, *om, *force, *kin, *torque, *cs, *fs;
    if ((rc = b->alloc(x, size_t(P) * 3)) || (rc = b->alloc(v, size_t(P) * 3)) || (rc = b->alloc(q, size_t(F) * 4)) ||
        (rc = b->alloc(om, size_t(F) * 3)) || (rc = b->alloc(force, size_t(P) * 3)) ||
        (rc = b->alloc(kin, size_t(P) * 3)) || (rc = b->alloc(torque, size_t(F) * 3)) || (rc = b->alloc(cs, R)) ||
        (rc = b->alloc(fs, R)) || (rc = b->alloc(b->selfOverflow, R)) ||
        (rc = b->alloc(b->fixedStage, size_t(F) * 4)) || (rc = b->alloc(b->clock, R)))
        return rc;
    const size_t w = sizeof(Real);
    RP_HIP(hipMemset(b->clock, 0, size_t(R) * 4));
    RP_HIP(hipMemset(b->selfOverflow, 0, size_t(R) * 4));
    RP_HIP(hipMemset(x, 0, size_t(P) * 3 * w));
    RP_HIP(hipMemset(v, 0, size_t(P) * 3 * w));
    RP_HIP(hipMemset(om, 0, size_t(F) * 3 * w));
    RP_HIP(hipMemset(force, 0, size_t(P) * 3 * w));
    RP_HIP(hipMemset(kin, 0, size_t(P) * 3 * w));
    RP_HIP(hipMemset(torque, 0, size_t(F) * 3 * w));
    {
        std::vector<Real> qid(size_t(F) * 4, Real(0)), ones(R, Real(1));
        for (int j = 0; j < F; ++j) qid[4 * size_t(j)] = Real(1);
        RP_HIP(hipMemcpy(q, qid.data(), qid.size() * w, hipMemcpyHostToDevice));
        RP_HIP(hipMemcpy(cs, ones.data(), R * w, hipMemcpyHostToDevice));
        RP_HIP(hipMemcpy(fs, ones.data(), R * w, hipMemcpyHostToDevice));
    }
    DeviceArraysT<Real>& d = b->d;
    d.x = x;
    d.v = v;
    d.q = q;
    d.omega = om;
    d.force = force;
    d.kinVel = kin;
    d.torque = torque;
    d.complianceScale = cs;
    d.frictionScale = fs;
    d.selfOverflow = b->selfOverflow;
    d.clock = b->clock;
    d.paths = nullptr;
    d.obs = nullptr;
    d.stepLimit = nullptr;
    d.rec = nullptr;
    d.frames = nullptr;
    d.framePitch = 0;
    d.numRods = R;
    using Vec4 = typename Prec<Real>::Vec4;
    Real* dInvMass;
    int *dSegEnd, *dCol;
    RodInfoT<Real>* dRods;
    PrimT<Real>* dPrims;
    int4 *sIdx4, *bIdx4;
    Vec4 *iI4, *I4, *sPar4, *bPar4, *bRest4;
    if ((rc = b->upload(dInvMass, rec.invMass)) || (rc = b->upload(iI4, rec.iI4)) || (rc = b->upload(I4, rec.I4)) ||
        (rc = b->upload(sIdx4, rec.sIdx)) || (rc = b->upload(sPar4, rec.sPar)) || (rc = b->upload(bIdx4, rec.bIdx)) ||
        (rc = b->upload(bPar4, rec.bPar)) || (rc = b->upload(bRest4, rec.bRest4)) ||
        (rc = b->upload(dSegEnd, rec.segEnd)) || (rc = b->upload(dCol, rec.colStart)) ||
        (rc = b->upload(dRods, b->rods)) || (rc = b->upload(dPrims, rec.prims)))
        return rc;
    d.invMass = dInvMass;
    d.invInertia4 = reinterpret_cast<const Real*>(iI4);
    d.inertia4 = reinterpret_cast<const Real*>(I4);
    d.sIdx = reinterpret_cast<const int*>(sIdx4);
    d.sPar = reinterpret_cast<const Real*>(sPar4);
    d.bIdx = reinterpret_cast<const int*>(bIdx4);
    d.bPar = reinterpret_cast<const Real*>(bPar4);
    d.bRest4 = reinterpret_cast<const Real*>(bRest4);
    d.segEnd = dSegEnd;
    d.colStart = dCol;
    d.rods = dRods;
    d.prims = dPrims;
    if (strategy == RP_STRATEGY_GLOBAL) {
        if ((rc = b->upload(b->scratchOff, b->scratchOffHost))) return rc;
        d.scratchOff = b->scratchOff;
    }
    *out_batch = static_cast<rp_batch_impl*>(b.release());
    return RP_OK;
}

template <typename T>
int copy_range(rp_batch_impl& b, int first, int count, T* dev, const T* host, T* hostOut,
               const std::vector<int>& startArr, int comps) {
    const size_t e0 = size_t(startArr[first]), e1 = size_t(startArr[first + count]);
    const size_t bytes = (e1 - e0) * comps * sizeof(T);
    if (bytes == 0) return RP_OK;
    if (host) RP_HIP(hipMemcpyAsync(dev + e0 * comps, host, bytes, hipMemcpyHostToDevice, b.stream));
    if (hostOut) RP_HIP(hipMemcpyAsync(hostOut, dev + e0 * comps, bytes, hipMemcpyDeviceToHost, b.stream));
    return RP_OK;
}

template <typename Real>
int set_state(void* batch, int first_rod, int num_rods, const Real* x, const Real* v, const Real* q,
              const Real* omega) {
    BatchT<Real>* bp;
    int rc = as_precision(batch, bp);
    if (rc) return rc;
    BatchT<Real>& b = *bp;
    if ((rc = check_range(b, first_rod, num_rods))) return rc;
    if (x) {
        const size_t n = size_t(b.pStart[first_rod + num_rods] - b.pStart[first_rod]) * 3;
        for (size_t i = 0; i < n; ++i) b.coordScale = std::max(b.coordScale, std::fabs(double(x[i])));
    }
    if ((rc = copy_range<Real>(b, first_rod, num_rods, b.d.x, x, nullptr, b.pStart, 3)) ||
        (rc = copy_range<Real>(b, first_rod, num_rods, b.d.v, v, nullptr, b.pStart, 3)) ||
        (rc = copy_range<Real>(b, first_rod, num_rods, b.d.q, q, nullptr, b.fStart, 4)) ||
        (rc = copy_range<Real>(b, first_rod, num_rods, b.d.omega, omega, nullptr, b.fStart, 3)))
        return rc;
    RP_HIP(hipStreamSynchronize(b.stream));
    return RP_OK;
}

template <typename Real>
int get_state(void* batch, int first_rod, int num_rods, Real* x, Real* v, Real* q, Real* omega) {
    BatchT<Real>* bp;
    int rc = as_precision(batch, bp);
    if (rc) return rc;
    BatchT<Real>& b = *bp;
    if ((rc = check_range(b, first_rod, num_rods))) return rc;
    if ((rc = copy_range<Real>(b, first_rod, num_rods, b.d.x, nullptr, x, b.pStart, 3)) ||
        (rc = copy_range<Real>(b, first_rod, num_rods, b.d.v, nullptr, v, b.pStart, 3)) ||
        (rc = copy_range<Real>(b, first_rod, num_rods, b.d.q, nullptr, q, b.fStart, 4)) ||
        (rc = copy_range<Real>(b, first_rod, num_rods, b.d.omega, nullptr, omega, b.fStart, 3)))
        return rc;
    RP_HIP(hipStreamSynchronize(b.stream));
    return RP_OK;
}

template <typename Real>
int set_loads(void* batch, int first_rod, int num_rods, const Real* force, const Real* kin_vel, const Real* torque,
              const Real* fixed_q) {
    BatchT<Real>* bp;
    int rc = as_precision(batch, bp);
    if (rc) return rc;
    BatchT<Real>& b = *bp;
    if ((rc = check_range(b, first_rod, num_rods))) return rc;
    if ((rc = copy_range<Real>(b, first_rod, num_rods, const_cast<Real*>(b.d.force), force, nullptr, b.pStart, 3)) ||
        (rc = copy_range<Real>(b, first_rod, num_rods, const_cast<Real*>(b.d.kinVel), kin_vel, nullptr, b.pStart, 3)) ||
        (rc = copy_range<Real>(b, first_rod, num_rods, const_cast<Real*>(b.d.torque), torque, nullptr, b.fStart, 3)))
        return rc;
    if (fixed_q) {
        const int f0 = b.fStart[first_rod], f1 = b.fStart[first_rod + num_rods];
        if (f1 > f0) {
            RP_HIP(hipMemcpyAsync(b.fixedStage, fixed_q, size_t(f1 - f0) * 4 * sizeof(Real), hipMemcpyHostToDevice,
                                  b.stream));
            RP_HIP(Prec<Real>::launch_fixed_frames(b.d.q, b.fixedStage, b.d.invInertia4, f0, f1 - f0, b.stream));
        }
    }
    RP_HIP(hipStreamSynchronize(b.stream));
    return RP_OK;
}

template <typename Real>
int set_scales(void* batch, const Real* youngs_scale, const Real* friction_scale) {
    BatchT<Real>* bp;
    int rc = as_precision(batch, bp);
    if (rc) return rc;
    BatchT<Real>& b = *bp;
    if (youngs_scale) {
        std::vector<Real> c(b.numRods);
        for (int r = 0; r < b.numRods; ++r) {
            if (!(youngs_scale[r] > Real(0)))
                return fail(RP_ERR_INVALID, "youngs_scale[" + std::to_string(r) + "] must be positive");
            // Compliance scales as the inverse of Young's modulus
            // (src/gpu/gpu_batch.cpp:437-439).
            c[r] = Real(1) / youngs_scale[r];
        }
        RP_HIP(hipMemcpyAsync(b.d.complianceScale, c.data(), b.numRods * sizeof(Real), hipMemcpyHostToDevice,
                              b.stream));
        RP_HIP(hipStreamSynchronize(b.stream));
    }
    if (friction_scale) {
        RP_HIP(hipMemcpyAsync(b.d.frictionScale, friction_scale, b.numRods * sizeof(Real), hipMemcpyHostToDevice,
                              b.stream));
        RP_HIP(hipStreamSynchronize(b.stream));
    }
    return RP_OK;
}

// GLOBAL: about 2^22 element-substeps of the longest rod per launch (a few
// hundred milliseconds), at least one step; RP_STEPS_PER_LAUNCH replaces it.
template <typename Real>
int global_steps_per_launch(const BatchT<Real>& b, const typename Prec<Real>::StepParams& p) {
    if (b.stepsPerLaunchSet) return b.stepsPerLaunch;
    int widest = 1;
    for (const RodInfoT<Real>& ri : b.rods) widest = std::max(widest, std::max(ri.nP, ri.nF));
    const double perStep = double(widest) * p.substeps * std::max(1, p.iterations);
    return int(std::max(1.0, std::min(double(b.stepsPerLaunch), std::floor(double(1 << 22) / perStep))));
}

// The reference's float motion margin (HangingCable::floatMotionMargin,
// src/apps/cable_hanging.cpp:75-81, required to be at least 2 by
// application_cases.cpp:196-197): gravity's displacement per substep, |g| h^2,
// over the spacing of the working precision at the largest coordinate the
// batch holds. Below about 1, a resting rod cannot start to move under
// gravity: rounding acts as static friction. Recorded per step call.
template <typename Real>
void note_motion_margin(rp_batch_impl& b, const typename Prec<Real>::StepParams& p, int steps) {
    const double g = std::sqrt(double(p.gravity[0]) * p.gravity[0] + double(p.gravity[1]) * p.gravity[1] +
                               double(p.gravity[2]) * p.gravity[2]);
    if (!(g > 0.0)) return;
    const Real c = Real(b.coordScale);
    const double spacing = double(std::nextafter(c, Real(INFINITY)) - c);
    const double h = double(p.dt) / p.substeps;
    const double margin = g * h * h / spacing;
    if (b.minMargin == 0.0 || margin < b.minMargin) b.minMargin = margin;
    if (margin < 2.0) b.lowMarginSteps += steps;
}

template <typename Real>
int step(void* batch, const typename Prec<Real>::StepParams* params, int num_steps) {
    if (!params) return fail(RP_ERR_INVALID, "null argument");
    BatchT<Real>* bp;
    int rc = as_precision(batch, bp);
    if (rc) return rc;
    BatchT<Real>& b = *bp;
    const auto& p = *params;
    if (num_steps < 0) return fail(RP_ERR_INVALID, "num_steps must be non-negative");
    if (p.substeps < 1 || p.iterations < 0 || !(p.dt > 0))
        return fail(RP_ERR_INVALID, "need substeps >= 1, iterations >= 0 and dt > 0");
    if (num_steps == 0) return RP_OK;
    if (b.strategy == RP_STRATEGY_TILED && (rc = resolve_tiles(b, p))) return rc;
    if (b.strategy == RP_STRATEGY_GLOBAL && !b.scratch) {
        char* region = nullptr;
        if ((rc = b.alloc(region, b.scratchBytes))) return rc;
        b.scratch = region;
        b.d.scratch = region;
    }
    note_motion_margin<Real>(b, p, num_steps);
    // Steps run in chunks: one TILED launch of at most stepsPerLaunch steps,
    // or plainStepsPerGraph PLAIN launches. Each chunk shape is captured once
    // into a HIP graph (cached per parameter set and length) and replayed.
    int chunk = b.strategy == RP_STRATEGY_PLAIN ? b.plainStepsPerGraph : b.stepsPerLaunch;
    if (b.strategy == RP_STRATEGY_GLOBAL) chunk = global_steps_per_launch(b, p);
    if (b.strategy == RP_STRATEGY_TILED && !b.stepsPerLaunchSet && b.tileMsPerSweep > 0.0) {
        // Hold a launch to about kLaunchBudgetMs of estimated device time (at
        // least 64 steps), so a very large batch still returns to the host
        // every few seconds.
        const double perStep = b.tileMsPerSweep * p.substeps * std::max(1, p.iterations);
        chunk = int(std::max(64.0, std::min(double(chunk), std::floor(kLaunchBudgetMs / perStep))));
    }
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
    b.lastLaunches = b.strategy == RP_STRATEGY_PLAIN ? num_steps : full + (rem ? 1 : 0);
    b.stepsTaken += num_steps;
    return RP_OK;
}

}  // namespace

extern "C" {

const char* rp_last_error(void) { return last_error(); }

int rp_tile_config_count(void) { return num_tile_configs(); }

const char* rp_tile_config_name(int id) {
    if (id < 1 || id > num_tile_configs()) return nullptr;
    return tile_config(id).name;
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
    // bend constraints, each family in two colours of at most ceil(n / 2))
    // some FP32 configuration takes.
    int best = 0;
    for (int n2 = 4; n2 <= 4096; ++n2) {
        const int lds = Prec<float>::tiled_lds_bytes(n2 + 1, n2 + 1, n2, 0, 0, 1, 0);
        if (size_t(lds) > prop.sharedMemPerBlock) break;
        for (int id = 1; id <= num_tile_configs(); ++id) {
            const TileConfig& c = tile_config(id);
            const int lanes = std::max((n2 + 1 + c.tp - 1) / c.tp, ((n2 + 1) / 2 + c.tc - 1) / c.tc);
            if (tile_config_available<float>(id) && lanes <= 64 * c.waves) best = n2;
        }
    }
    out->max_tiled_segments = best;
    return RP_OK;
}

int rp_batch_create(const rp_batch_desc* desc, void** out_batch) { return create<float>(desc, out_batch); }
int rp_batch_create64(const rp_batch_desc64* desc, void** out_batch) { return create<double>(desc, out_batch); }

void rp_batch_destroy(void* batch) {
    if (!batch) return;
    rp_batch_impl* b = static_cast<rp_batch_impl*>(batch);
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    delete b;
}

int rp_batch_precision(void* batch) {
    if (!batch) return fail(RP_ERR_INVALID, "null batch");
    return static_cast<rp_batch_impl*>(batch)->bits;
}

int rp_batch_set_state(void* batch, int first_rod, int num_rods, const float* x, const float* v, const float* q,
                       const float* omega) {
    return set_state<float>(batch, first_rod, num_rods, x, v, q, omega);
}
int rp_batch_set_state64(void* batch, int first_rod, int num_rods, const double* x, const double* v,
                         const double* q, const double* omega) {
    return set_state<double>(batch, first_rod, num_rods, x, v, q, omega);
}

int rp_batch_set_loads(void* batch, int first_rod, int num_rods, const float* force, const float* kin_vel,
                       const float* torque, const float* fixed_q) {
    return set_loads<float>(batch, first_rod, num_rods, force, kin_vel, torque, fixed_q);
}
int rp_batch_set_loads64(void* batch, int first_rod, int num_rods, const double* force, const double* kin_vel,
                         const double* torque, const double* fixed_q) {
    return set_loads<double>(batch, first_rod, num_rods, force, kin_vel, torque, fixed_q);
}

int rp_batch_set_scales(void* batch, const float* youngs_scale, const float* friction_scale) {
    return set_scales<float>(batch, youngs_scale, friction_scale);
}
int rp_batch_set_scales64(void* batch, const double* youngs_scale, const double* friction_scale) {
    return set_scales<double>(batch, youngs_scale, friction_scale);
}

int rp_batch_step(void* batch, const rp_step_params* params, int num_steps) {
    return step<float>(batch, params, num_steps);
}
int rp_batch_step64(void* batch, const rp_step_params64* params, int num_steps) {
    return step<double>(batch, params, num_steps);
}

// rp_batch_step for a GLOBAL FP32 batch with the working arrays in the
// library's fixed per-device region instead of the batch's own: every launch
// goes straight onto the batch's stream (no graph is captured or
// instantiated) and nothing is allocated.
int rp_batch_step_noalloc(void* batch, const rp_step_params* params, int num_steps) {
    if (!params) return fail(RP_ERR_INVALID, "null argument");
    BatchT<float>* bp;
    int rc = as_precision(batch, bp);
    if (rc) return rc;
    BatchT<float>& b = *bp;
    const rp_step_params& p = *params;
    if (b.strategy != RP_STRATEGY_GLOBAL)
        return fail(RP_ERR_STATE, "rp_batch_step_noalloc steps GLOBAL batches only (create with RP_STRATEGY_GLOBAL)");
    if (num_steps < 0) return fail(RP_ERR_INVALID, "num_steps must be non-negative");
    if (p.substeps < 1 || p.iterations < 0 || !(p.dt > 0))
        return fail(RP_ERR_INVALID, "need substeps >= 1, iterations >= 0 and dt > 0");
    int dev = 0;
    RP_HIP(hipGetDevice(&dev));
    if (dev != b.device)
        return fail(RP_ERR_STATE, "the batch's device (" + std::to_string(b.device) + ") must be current, not " +
                                      std::to_string(dev));
    char* region = nullptr;
    size_t cap = 0;
    RP_HIP(f32::noalloc_scratch(&region, &cap));
    if (b.scratchBytes > cap)
        return fail(RP_ERR_CAPACITY, "the batch's working arrays need " + std::to_string(b.scratchBytes) +
                                         " bytes; rp_batch_step_noalloc has " + std::to_string(cap) +
                                         " (RP_NOALLOC_SCRATCH_BYTES)");
    if (num_steps == 0) return RP_OK;
    note_motion_margin<float>(b, p, num_steps);
    DeviceArraysT<float> d = b.d;
    d.scratch = region;
    const int chunk = global_steps_per_launch(b, p);
    if (b.timing) RP_HIP(hipEventRecord(b.ev0, b.stream));
    int launches = 0;
    for (int left = num_steps; left > 0; ++launches) {
        const int n = std::min(left, chunk);
        RP_HIP(Prec<float>::launch_global(d, step_args<float>(p, n), b.globalThreads, size_t(b.globalXqLds),
                                          b.stream));
        left -= n;
    }
    if (b.timing) {
        RP_HIP(hipEventRecord(b.ev1, b.stream));
        b.evPending = true;
    }
    b.lastLaunches = launches;
    b.stepsTaken += num_steps;
    return RP_OK;
}

int rp_batch_get_state(void* batch, int first_rod, int num_rods, float* x, float* v, float* q, float* omega) {
    return get_state<float>(batch, first_rod, num_rods, x, v, q, omega);
}
int rp_batch_get_state64(void* batch, int first_rod, int num_rods, double* x, double* v, double* q,
                         double* omega) {
    return get_state<double>(batch, first_rod, num_rods, x, v, q, omega);
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
        b.bind_tables();
    }
    std::vector<PathRec> recs(num_rods);
    for (int i = 0; i < num_rods; ++i) {
        PathRec& p = recs[i];
        std::memset(&p, 0, sizeof(p));
        p.particle = -1;
        if (!paths || paths[i].particle < 0) continue;
        const rp_path_desc& s = paths[i];
        const int nP = b.pStart[first_rod + i + 1] - b.pStart[first_rod + i];
        if (s.particle >= nP) return fail(RP_ERR_INVALID, "path " + std::to_string(i) + ": particle out of range");
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
        b.bind_tables();
    }
    std::vector<ObsRec> recs(num_rods, ObsRec{-1, 1, 0.0, 0.0});
    for (int i = 0; observers && i < num_rods; ++i) {
        const rp_observer_desc& o = observers[i];
        if (o.particle < 0) continue;
        const int nP = b.pStart[first_rod + i + 1] - b.pStart[first_rod + i];
        if (o.particle >= nP || o.every < 1)
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

int rp_batch_record(void* batch, int every, int max_frames) {
    if (!batch) return fail(RP_ERR_INVALID, "null batch");
    rp_batch_impl& b = *static_cast<rp_batch_impl*>(batch);
    if (every < 0 || max_frames < 0 || (every > 0 && max_frames < 1))
        return fail(RP_ERR_INVALID, "need every >= 1 and max_frames >= 1, or every = 0 to stop recording");
    const size_t pitch = 3 * size_t(b.pStart[b.numRods]);
    if (every > 0 && size_t(max_frames) > (size_t(1) << 62) / (pitch * sizeof(float) + 1))
        return fail(RP_ERR_INVALID, "frame buffer size overflows");
    const size_t bytes = every > 0 ? size_t(max_frames) * pitch * sizeof(float) : 0;
    int rc;
    if (!b.rec && every > 0) {
        if ((rc = b.alloc(b.rec, b.numRods))) return rc;
        b.drop_graphs();
        b.bind_tables();
    }
    if (every == 0 || !b.frameBuf || b.frameBuf->bytes < bytes) {
        // The kernels' argument block holds the buffer's address: graphs
        // captured with the old one must go before it is freed.
        b.drop_graphs();
        if (b.frameBuf) b.deviceBytes -= b.frameBuf->bytes;
        b.frameBuf.reset();
        if (bytes) {
            auto buf = std::make_unique<DevBuf>();
            const hipError_t e = hipMalloc(&buf->p, bytes);
            if (e != hipSuccess)
                return fail(RP_ERR_MEMORY, "hipMalloc of " + std::to_string(bytes) + " bytes of frames failed: " +
                                               hipGetErrorString(e));
            buf->bytes = bytes;
            b.deviceBytes += bytes;
            b.frameBuf = std::move(buf);
        }
        b.bind_tables();
    }
    if (bytes) RP_HIP(hipMemsetAsync(b.frameBuf->p, 0, bytes, b.stream));
    b.recOrigin.assign(b.numRods, 0);
    if (b.rec) {
        RP_HIP(hipMemcpyAsync(b.recOrigin.data(), b.clock, size_t(b.numRods) * 4, hipMemcpyDeviceToHost, b.stream));
        RP_HIP(hipStreamSynchronize(b.stream));
        std::vector<RecRec> recs(b.numRods);
        for (int r = 0; r < b.numRods; ++r) recs[r] = RecRec{b.recOrigin[r], every, max_frames, 0};
        RP_HIP(hipMemcpyAsync(b.rec, recs.data(), recs.size() * sizeof(RecRec), hipMemcpyHostToDevice, b.stream));
    }
    RP_HIP(hipStreamSynchronize(b.stream));
    b.recEvery = every;
    b.recMaxFrames = every > 0 ? max_frames : 0;
    return RP_OK;
}

int rp_batch_recorded_frames(void* batch, int* frames) {
    if (!batch || !frames) return fail(RP_ERR_INVALID, "null argument");
    rp_batch_impl& b = *static_cast<rp_batch_impl*>(batch);
    *frames = 0;
    if (b.recEvery <= 0) return RP_OK;
    std::vector<int> clk(b.numRods);
    RP_HIP(hipMemcpyAsync(clk.data(), b.clock, clk.size() * 4, hipMemcpyDeviceToHost, b.stream));
    RP_HIP(hipStreamSynchronize(b.stream));
    int n = b.recMaxFrames;
    for (int r = 0; r < b.numRods; ++r) {
        const int k = clk[r] - b.recOrigin[r];
        n = std::min(n, k > 0 ? k / b.recEvery : 0);
    }
    *frames = n;
    return RP_OK;
}

int rp_batch_get_frames(void* batch, int first_frame, int num_frames, float* positions) {
    if (!batch || (!positions && num_frames > 0)) return fail(RP_ERR_INVALID, "null argument");
    rp_batch_impl& b = *static_cast<rp_batch_impl*>(batch);
    if (b.recEvery <= 0) return fail(RP_ERR_STATE, "the batch is not recording");
    if (first_frame < 0 || num_frames < 0 || first_frame > b.recMaxFrames - num_frames)
        return fail(RP_ERR_INVALID, "frames [" + std::to_string(first_frame) + ", +" + std::to_string(num_frames) +
                                        ") outside the " + std::to_string(b.recMaxFrames) + " recorded");
    const size_t pitch = 3 * size_t(b.pStart[b.numRods]);
    if (num_frames == 0) return RP_OK;
    RP_HIP(hipMemcpyAsync(positions, static_cast<const float*>(b.frameBuf->p) + size_t(first_frame) * pitch,
                          size_t(num_frames) * pitch * sizeof(float), hipMemcpyDeviceToHost, b.stream));
    RP_HIP(hipStreamSynchronize(b.stream));
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

int rp_batch_counters(void* batch, rp_counters* out) {
    if (!batch || !out) return fail(RP_ERR_INVALID, "null argument");
    rp_batch_impl& b = *static_cast<rp_batch_impl*>(batch);
    std::memset(out, 0, sizeof(*out));
    std::vector<int> ovf(b.numRods);
    RP_HIP(hipStreamSynchronize(b.stream));
    RP_HIP(hipMemcpy(ovf.data(), b.selfOverflow, b.numRods * 4, hipMemcpyDeviceToHost));
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
    out->motion_margin = b.minMargin;
    out->low_margin_steps = b.lowMarginSteps;
    out->scratch_bytes = b.scratch ? b.scratchBytes : 0;
    if (b.strategy == RP_STRATEGY_PLAIN) {
        out->workgroups = b.numRods;
        out->lanes_total = b.numRods * 128;
        out->lds_bytes_per_workgroup = size_t(b.plainLds);
        out->launches_per_step = 1;
    } else if (b.strategy == RP_STRATEGY_GLOBAL) {
        out->workgroups = b.numRods;
        out->lanes_busy = b.numRods * b.globalThreads;
        out->lanes_total = b.numRods * b.globalThreads;
        out->lds_bytes_per_workgroup = size_t(b.globalXqLds);
        out->launches_per_step = 0;
    } else if (b.tiles) {
        const TileConfig& c = tile_config(b.tiles->config);
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

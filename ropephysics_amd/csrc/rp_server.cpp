// Continuous-batching server: simulation requests of any size arrive, are
// admitted while the device memory set aside for them (an HBM budget) has
// room, stay resident on the device between calls, and are stepped together:
// every call of rp_server_step is one TILED launch over the rods of every
// resident request, each rod stopping at its own request's horizon. Finished
// requests are read back and released by the caller, freeing their ranges for
// requests still waiting.
//
// The reference has no counterpart; the arithmetic is the batch's (the same
// records and kernels), so a request's result does not depend on what else
// was resident.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "rp_host.h"

namespace {

using namespace rp;
using namespace rp::host;
using namespace rp::f32;  // the server steps FP32 requests

// First-fit allocator of index ranges in [0, capacity), coalescing on free.
class RangePool {
   public:
    void reset(long long capacity) {
        cap_ = capacity;
        free_.clear();
        if (capacity > 0) free_[0] = capacity;
        used_ = 0;
    }
    // Start of a free range of `n` (n == 0: returns 0 and uses nothing), or -1.
    long long take(long long n) {
        if (n == 0) return 0;
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second < n) continue;
            const long long start = it->first, len = it->second;
            free_.erase(it);
            if (len > n) free_[start + n] = len - n;
            used_ += n;
            return start;
        }
        return -1;
    }
    void give(long long start, long long n) {
        if (n == 0) return;
        used_ -= n;
        auto next = free_.lower_bound(start);
        if (next != free_.end() && start + n == next->first) {
            n += next->second;
            next = free_.erase(next);
        }
        if (next != free_.begin()) {
            auto prev = std::prev(next);
            if (prev->first + prev->second == start) {
                prev->second += n;
                return;
            }
        }
        free_[start] = n;
    }
    long long used() const { return used_; }
    long long capacity() const { return cap_; }

   private:
    std::map<long long, long long> free_;
    long long cap_ = 0, used_ = 0;
};

enum Pool { kParticles, kFrames, kStretch, kBend, kRods, kPrims, kColours, kNumPools };

// Device bytes per element of each pool.
constexpr long long kBytesPer[kNumPools] = {
    12 * 4 + 4,          // particle: x, v, force, kinVel (xyz); invMass
    16 + 12 + 12 + 32,   // frame: q, omega, torque, invInertia4, inertia4
    16 + 16 + 8,         // stretch: sIdx, sPar, segEnd
    16 + 16 + 16,        // bend: bIdx, bPar, bRest4
    // rod: record, per-rod words; tile, slot and step count for each of the two launches
    int(sizeof(RodInfo)) + 6 * 4 + 2 * (int(sizeof(Tile)) + int(sizeof(TileSlot)) + 4),
    int(sizeof(Prim)),   // primitive
    4,                   // colour-start entry
};

struct Request {
    int id = 0;
    long long start[kNumPools] = {};
    long long count[kNumPools] = {};
    int steps = 0;       // horizon
    int stepsDone = 0;   // steps every rod of the request has taken
    long long bytes = 0;
};

struct DevArray {
    void* p = nullptr;
    ~DevArray() {
        if (p) (void)hipFree(p);
    }
};

// Pinned host blocks that uploads are copied through, so a call can queue its
// copies and return: a block is reused once the event recorded after its
// copies has completed, and freed with the server.
class Staging {
   public:
    ~Staging() {
        for (Block& b : blocks_) {
            if (b.done) (void)hipEventSynchronize(b.done);
            if (b.done) (void)hipEventDestroy(b.done);
            if (b.p) (void)hipHostFree(b.p);
        }
    }
    // A block of at least `bytes` whose earlier copies are done; its index in
    // *idx. Null when the pinned allocation fails.
    char* get(size_t bytes, int* idx) {
        for (size_t i = 0; i < blocks_.size(); ++i) {
            Block& b = blocks_[i];
            if (b.bytes < bytes) continue;
            if (b.busy && hipEventQuery(b.done) != hipSuccess) continue;
            b.busy = false;
            *idx = int(i);
            return static_cast<char*>(b.p);
        }
        Block b;
        b.bytes = std::max<size_t>(bytes, size_t(1) << 20);
        if (hipHostMalloc(&b.p, b.bytes, hipHostMallocDefault) != hipSuccess) return nullptr;
        if (hipEventCreateWithFlags(&b.done, hipEventDisableTiming) != hipSuccess) {
            (void)hipHostFree(b.p);
            return nullptr;
        }
        blocks_.push_back(b);
        *idx = int(blocks_.size()) - 1;
        return static_cast<char*>(b.p);
    }
    // The copies out of block `idx` are queued on `stream`.
    hipError_t queued(int idx, hipStream_t stream) {
        blocks_[size_t(idx)].busy = true;
        return hipEventRecord(blocks_[size_t(idx)].done, stream);
    }

   private:
    struct Block {
        void* p = nullptr;
        size_t bytes = 0;
        hipEvent_t done = nullptr;
        bool busy = false;
    };
    std::vector<Block> blocks_;
};

// Host-to-device copies gathered into one pinned block and queued together.
class Upload {
   public:
    template <typename T>
    void add(T* dst, long long at, const T* src, long long count) {
        if (count <= 0) return;
        parts_.push_back({dst + at, src, size_t(count) * sizeof(T), total_});
        total_ += (size_t(count) * sizeof(T) + 15) & ~size_t(15);
    }
    int queue(Staging& staging, hipStream_t stream) {
        if (parts_.empty()) return RP_OK;
        int idx = 0;
        char* block = staging.get(total_, &idx);
        if (!block) return fail(RP_ERR_MEMORY, "pinned staging allocation failed");
        for (const Part& p : parts_) std::memcpy(block + p.off, p.src, p.bytes);
        for (const Part& p : parts_)
            RP_HIP(hipMemcpyAsync(p.dst, block + p.off, p.bytes, hipMemcpyHostToDevice, stream));
        RP_HIP(staging.queued(idx, stream));
        return RP_OK;
    }

   private:
    struct Part {
        void* dst;
        const void* src;
        size_t bytes, off;
    };
    std::vector<Part> parts_;
    size_t total_ = 0;
};

}  // namespace

struct rp_server_impl {
    rp_step_params params{};
    // Rods without self-collision run the lean kernel, the others the
    // self-collision variant, each packed in its own tile shape and launched
    // on its own stream so the two launches share the device.
    // Default shapes, from a sweep of the serving stream's first 40 requests
    // arriving at once (4 GiB budget): lean tp3-tc1-w4 (tp2-tc1-w4 and
    // tp3-tc1-w2 within 1-3%); self-colliding tp1-tc1-w8 in chunked steps
    // (2.82 G segment-substeps/s against 2.77 G for tp2-tc1-w4) and
    // tp2-tc1-w4 in time-sliced ones (2.56 G against 2.39 G), tp3-tc1-w4 well
    // behind in both (2.11 and 1.82 G).
    int configLean = 10;  // tp3-tc1-w4
    int configSelf = 13;  // tp1-tc1-w8; tp2-tc1-w4 when time-sliced
    int stepsPerLaunch = 64;
    // Time-sliced steps: the slice in device wall-clock ticks (0: chunks of
    // stepsPerLaunch), and whether a launch's progress is still to be read.
    long long sliceTicks = 0;
    bool settlePending = false;
    std::vector<int> clockHost;
    LdsLimits lim;
    hipStream_t stream = nullptr;
    hipStream_t selfStream = nullptr;
    hipEvent_t evReady = nullptr, evSelfDone = nullptr;
    RangePool pools[kNumPools];
    std::vector<std::unique_ptr<DevArray>> owned;
    long long deviceBytes = 0;
    DeviceArrays d{};
    int* stepLimit = nullptr;
    Tile* dTiles = nullptr;
    TileSlot* dSlots = nullptr;
    int* dProgress = nullptr;  // per workgroup of each launch (launch pairing)
    std::vector<RodInfo> rods;  // host mirror, indexed by rod slot
    std::vector<int> colStart;  // host mirror of the colour-start table
    std::vector<unsigned char> selfOf;
    Staging staging;
    std::map<int, Request> requests;  // resident, by id
    std::vector<int> finished;        // ids whose horizon is reached, not yet released
    int nextId = 1;
    bool dirty = true;  // membership changed since the last packing
    Packing packLean, packSelf;
    long long launches = 0;
    long long stepsTaken = 0;

    ~rp_server_impl() {
        if (evReady) (void)hipEventDestroy(evReady);
        if (evSelfDone) (void)hipEventDestroy(evSelfDone);
        // The streams are the device's shared pair (server_streams).
    }

    template <typename T>
    int alloc(T*& out, long long count) {
        auto b = std::make_unique<DevArray>();
        const size_t bytes = size_t(std::max<long long>(count, 1)) * sizeof(T);
        const hipError_t e = hipMalloc(&b->p, bytes);
        if (e != hipSuccess)
            return fail(RP_ERR_MEMORY, "hipMalloc of " + std::to_string(bytes) + " bytes failed: " +
                                           hipGetErrorString(e));
        RP_HIP(hipMemset(b->p, 0, bytes));
        deviceBytes += (long long)bytes;
        out = static_cast<T*>(b->p);
        owned.push_back(std::move(b));
        return RP_OK;
    }
};

namespace {

rp_server_impl* S(void* p) { return static_cast<rp_server_impl*>(p); }

// The two launches of a server step must run side by side. Streams created
// with a compute-unit mask get hardware queues of their own, where plain
// streams share the process's few queues once other streams exist, and two
// launches on one queue run one after the other (measured: a time-sliced
// stream's p50 and p99 latency nearly doubled while another server's streams
// existed). Each such stream holds a hardware queue, so every server on a
// device shares one pair, made on the first server's creation with that
// device current and kept until the process ends: servers on one device
// then queue behind each other, and each server's synchronisations also wait
// for the others' queued work. The mask names every compute unit.
int server_streams(hipStream_t* lean, hipStream_t* self) {
    constexpr int kMaxDevices = 64;
    static std::mutex lock;
    static hipStream_t pairs[kMaxDevices][2] = {};
    int dev = 0, cus = 0;
    RP_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) return fail(RP_ERR_DEVICE, "device index out of range");
    std::lock_guard<std::mutex> guard(lock);
    if (!pairs[dev][0]) {
        RP_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
        std::vector<uint32_t> mask(size_t((cus + 31) / 32), 0xffffffffu);
        if (cus % 32) mask.back() = (1u << (cus % 32)) - 1u;
        hipStream_t a = nullptr, b = nullptr;
        RP_HIP(hipExtStreamCreateWithCUMask(&a, uint32_t(mask.size()), mask.data()));
        RP_HIP(hipExtStreamCreateWithCUMask(&b, uint32_t(mask.size()), mask.data()));
        pairs[dev][0] = a;
        pairs[dev][1] = b;
    }
    *lean = pairs[dev][0];
    *self = pairs[dev][1];
    return RP_OK;
}

void release_ranges(rp_server_impl& s, const Request& q) {
    for (int k = 0; k < kNumPools; ++k) s.pools[k].give(q.start[k], q.count[k]);
    for (long long r = q.start[kRods]; r < q.start[kRods] + q.count[kRods]; ++r) s.selfOf[size_t(r)] = 0;
}

// After a time-sliced launch: wait for it, read the rods' step clocks and set
// each running request's progress to its slowest rod's; requests at their
// horizon become finished.
int settle(rp_server_impl& s) {
    if (!s.settlePending) return RP_OK;
    // One copy of the span of rod slots the running requests occupy.
    long long lo = -1, hi = -1;
    for (const auto& kv : s.requests) {
        const Request& q = kv.second;
        if (q.stepsDone >= q.steps || q.count[kRods] == 0) continue;
        lo = lo < 0 ? q.start[kRods] : std::min(lo, q.start[kRods]);
        hi = std::max(hi, q.start[kRods] + q.count[kRods]);
    }
    s.clockHost.resize(size_t(s.pools[kRods].capacity()));
    if (lo >= 0)
        RP_HIP(hipMemcpyAsync(s.clockHost.data() + lo, s.d.clock + lo, size_t(hi - lo) * sizeof(int),
                              hipMemcpyDeviceToHost, s.stream));
    RP_HIP(hipStreamSynchronize(s.stream));
    s.settlePending = false;
    int advance = 0;
    for (auto& kv : s.requests) {
        Request& q = kv.second;
        if (q.stepsDone >= q.steps) continue;
        int done = q.steps;
        for (long long r = q.start[kRods]; r < q.start[kRods] + q.count[kRods]; ++r)
            done = std::min(done, s.clockHost[size_t(r)]);
        advance = std::max(advance, done - q.stepsDone);
        q.stepsDone = done;
        if (q.stepsDone >= q.steps) {
            s.finished.push_back(q.id);
            s.dirty = true;
        }
    }
    s.stepsTaken += advance;
    return RP_OK;
}

}  // namespace

extern "C" {

int rp_server_create(const rp_server_desc* desc, void** out_server) {
    if (!desc || !out_server) return fail(RP_ERR_INVALID, "null argument");
    *out_server = nullptr;
    if (desc->hbm_budget_bytes <= 0) return fail(RP_ERR_INVALID, "hbm_budget_bytes must be positive");
    if (desc->params.dt <= 0.0f || desc->params.substeps < 1 || desc->params.iterations < 1)
        return fail(RP_ERR_INVALID, "dt, substeps and iterations must be positive");
    if (desc->tile_config < 0 || desc->tile_config > num_tile_configs() || desc->self_tile_config < 0 ||
        desc->self_tile_config > num_tile_configs())
        return fail(RP_ERR_INVALID, "unknown tile configuration");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return fail(RP_ERR_DEVICE, "no usable HIP device");
    auto s = std::make_unique<rp_server_impl>();
    s->params = desc->params;
    if (desc->tile_config > 0) s->configLean = s->configSelf = desc->tile_config;
    if (desc->self_tile_config > 0) s->configSelf = desc->self_tile_config;
    if (desc->steps_per_launch > 0) s->stepsPerLaunch = desc->steps_per_launch;
    if (desc->time_slice_us < 0) return fail(RP_ERR_INVALID, "time_slice_us must not be negative");
    if (desc->time_slice_us > 0) {
        int dev = 0, khz = 0;
        RP_HIP(hipGetDevice(&dev));
        RP_HIP(hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, dev));
        if (khz <= 0) return fail(RP_ERR_DEVICE, "the device reports no wall-clock rate");
        s->sliceTicks = std::max(1LL, (long long)desc->time_slice_us * khz / 1000);
        if (desc->tile_config == 0 && desc->self_tile_config == 0) s->configSelf = 9;  // tp2-tc1-w4
    }
    int rc = lds_limits(s->lim);
    if (rc) return rc;

    // Capacities from the budget: per segment about one particle, one frame,
    // one stretch and one bend constraint; rods, primitives and colour
    // entries sized for rods of 16 segments on average.
    const long long perSegment = kBytesPer[kParticles] + kBytesPer[kFrames] + kBytesPer[kStretch] + kBytesPer[kBend];
    const long long perRodShare = (kBytesPer[kRods] + 2 * kBytesPer[kPrims] + 18 * kBytesPer[kColours]) / 16;
    long long segs = desc->hbm_budget_bytes / (perSegment + perRodShare);
    if (segs < 64) return fail(RP_ERR_INVALID, "hbm_budget_bytes below the smallest useful pool");
    segs = std::min<long long>(segs, 1LL << 30);
    long long rods = desc->max_rods > 0 ? desc->max_rods : std::max<long long>(1, segs / 16);
    rods = std::min<long long>(rods, 1LL << 24);
    const long long cap[kNumPools] = {segs, segs, segs, segs, rods, 2 * rods, 18 * rods};
    for (int k = 0; k < kNumPools; ++k) s->pools[k].reset(cap[k]);

    if ((rc = server_streams(&s->stream, &s->selfStream))) return rc;
    RP_HIP(hipEventCreateWithFlags(&s->evReady, hipEventDisableTiming));
    RP_HIP(hipEventCreateWithFlags(&s->evSelfDone, hipEventDisableTiming));
    DeviceArrays& d = s->d;
    float *x, *v, *q, *om, *force, *kin, *torque, *invMass, *cs, *fs;
    float4 *iI4, *I4, *sPar, *bPar, *bRest;
    int4 *sIdx, *bIdx;
    int *segEnd, *col, *ovf, *clock;
    RodInfo* dRods;
    Prim* dPrims;
    const long long P = cap[kParticles], F = cap[kFrames], St = cap[kStretch], B = cap[kBend], R = cap[kRods];
    if ((rc = s->alloc(x, P * 3)) || (rc = s->alloc(v, P * 3)) || (rc = s->alloc(force, P * 3)) ||
        (rc = s->alloc(kin, P * 3)) || (rc = s->alloc(invMass, P)) || (rc = s->alloc(q, F * 4)) ||
        (rc = s->alloc(om, F * 3)) || (rc = s->alloc(torque, F * 3)) || (rc = s->alloc(iI4, F)) ||
        (rc = s->alloc(I4, F)) || (rc = s->alloc(sIdx, St)) || (rc = s->alloc(sPar, St)) ||
        (rc = s->alloc(segEnd, 2 * St)) || (rc = s->alloc(bIdx, B)) || (rc = s->alloc(bPar, B)) ||
        (rc = s->alloc(bRest, B)) || (rc = s->alloc(dRods, R)) || (rc = s->alloc(cs, R)) ||
        (rc = s->alloc(fs, R)) || (rc = s->alloc(ovf, R)) || (rc = s->alloc(clock, R)) ||
        (rc = s->alloc(s->stepLimit, R)) || (rc = s->alloc(s->dTiles, 2 * R)) || (rc = s->alloc(s->dSlots, 2 * R)) ||
        (rc = s->alloc(s->dProgress, 2 * R)) || (rc = s->alloc(dPrims, cap[kPrims])) || (rc = s->alloc(col, cap[kColours])))
        return rc;
    d.x = x;
    d.v = v;
    d.q = q;
    d.omega = om;
    d.force = force;
    d.kinVel = kin;
    d.torque = torque;
    d.invMass = invMass;
    d.invInertia4 = reinterpret_cast<float*>(iI4);
    d.inertia4 = reinterpret_cast<float*>(I4);
    d.sIdx = reinterpret_cast<int*>(sIdx);
    d.sPar = reinterpret_cast<float*>(sPar);
    d.bIdx = reinterpret_cast<int*>(bIdx);
    d.bPar = reinterpret_cast<float*>(bPar);
    d.bRest4 = reinterpret_cast<float*>(bRest);
    d.segEnd = segEnd;
    d.colStart = col;
    d.rods = dRods;
    d.prims = dPrims;
    d.complianceScale = cs;
    d.frictionScale = fs;
    d.selfOverflow = ovf;
    d.clock = clock;
    d.paths = nullptr;
    d.obs = nullptr;
    d.stepLimit = s->stepLimit;
    d.numRods = int(R);
    s->rods.assign(size_t(R), RodInfo{});
    s->colStart.assign(size_t(cap[kColours]), 0);
    s->selfOf.assign(size_t(R), 0);
    *out_server = s.release();
    return RP_OK;
}

void rp_server_destroy(void* server) {
    if (!server) return;
    rp_server_impl* s = S(server);
    if (s->stream) (void)hipStreamSynchronize(s->stream);
    delete s;
}

int rp_server_submit(void* server, const rp_batch_desc* req, const rp_request_state* st, int steps, int* out_id) {
    if (!server || !req || !st || !out_id) return fail(RP_ERR_INVALID, "null argument");
    rp_server_impl& s = *S(server);
    *out_id = 0;
    if (steps < 1) return fail(RP_ERR_INVALID, "a request needs at least one step");
    if (!st->x || !st->q) return fail(RP_ERR_INVALID, "a request needs initial positions and orientations");
    Records rec;
    int rc = build_records(req, rec);
    if (rc) return rc;
    for (int r = 0; r < rec.R; ++r) {
        const RodInfo& ri = rec.rods[r];
        const TileConfig& cfg = tile_config(ri.self ? s.configSelf : s.configLean);
        if (!rec.colourFit[r]) return fail(RP_ERR_CAPACITY, "rod " + std::to_string(r) + " needs more than two colours");
        // The packer's one-workgroup-per-CU budget, not the raw per-workgroup
        // limit: a rod between the two would be admitted but never packed.
        if (rod_tiled_lds(ri) > lds_budget(s.lim, 1))
            return fail(RP_ERR_CAPACITY, "rod " + std::to_string(r) + " needs more LDS than a workgroup has");
        if (group_lanes(ri, rec.colStart, cfg.tp, cfg.tc) > 64 * cfg.waves)
            return fail(RP_ERR_CAPACITY, "rod " + std::to_string(r) + " is too long for the tile shape");
    }
    Request q;
    const long long need[kNumPools] = {rec.P, rec.F, rec.S, rec.B, rec.R, (long long)rec.prims.size(),
                                       (long long)rec.colStart.size()};
    for (int k = 0; k < kNumPools; ++k) {
        if (need[k] > s.pools[k].capacity())
            return fail(RP_ERR_CAPACITY, "request larger than the server's whole budget");
        const long long at = s.pools[k].take(need[k]);
        if (at < 0) {
            for (int j = 0; j < k; ++j) s.pools[j].give(q.start[j], q.count[j]);
            return fail(RP_ERR_CAPACITY, "no room for the request now; release finished requests and retry");
        }
        q.start[k] = at;
        q.count[k] = need[k];
        q.bytes += need[k] * kBytesPer[k];
    }
    // Rod records with offsets into the pools.
    const long long p0 = q.start[kParticles], f0 = q.start[kFrames], s0 = q.start[kStretch], b0 = q.start[kBend];
    const long long r0 = q.start[kRods], prim0 = q.start[kPrims], c0 = q.start[kColours];
    std::vector<RodInfo> recs(rec.rods);
    for (RodInfo& ri : recs) {
        ri.pStart += int(p0);
        ri.fStart += int(f0);
        ri.sStart += int(s0);
        ri.bStart += int(b0);
        ri.primOff += int(prim0);
        ri.colOff += int(c0);
    }
    const size_t P = size_t(rec.P), F = size_t(rec.F), R = size_t(rec.R);
    std::vector<float> ones(R, 1.0f), comp(R, 1.0f), fric(R, 1.0f), zeros3(std::max(P, F) * 3, 0.0f);
    for (size_t r = 0; r < R; ++r) {
        if (st->youngs_scale) {
            if (!(st->youngs_scale[r] > 0.0f)) {
                release_ranges(s, q);
                return fail(RP_ERR_INVALID, "youngs_scale must be positive");
            }
            // Compliance scales as the inverse of Young's modulus, as the batch forms it.
            comp[r] = 1.0f / st->youngs_scale[r];
        }
        if (st->friction_scale) fric[r] = st->friction_scale[r];
    }
    std::vector<int> zerosR(R, 0), limit(R, steps);
    const DeviceArrays& d = s.d;
    // Staged through pinned memory and queued: the request's first launch, on
    // the same stream, starts after them, and submit does not wait.
    Upload up;
    up.add(d.x, p0 * 3, st->x, P * 3);
    up.add(d.v, p0 * 3, st->v ? st->v : zeros3.data(), P * 3);
    up.add(const_cast<float*>(d.force), p0 * 3, st->force ? st->force : zeros3.data(), P * 3);
    up.add(const_cast<float*>(d.kinVel), p0 * 3, st->kin_vel ? st->kin_vel : zeros3.data(), P * 3);
    up.add(const_cast<float*>(d.invMass), p0, rec.invMass.data(), P);
    up.add(d.q, f0 * 4, st->q, F * 4);
    up.add(d.omega, f0 * 3, st->omega ? st->omega : zeros3.data(), F * 3);
    up.add(const_cast<float*>(d.torque), f0 * 3, st->torque ? st->torque : zeros3.data(), F * 3);
    up.add(reinterpret_cast<float4*>(const_cast<float*>(d.invInertia4)), f0, rec.iI4.data(), F);
    up.add(reinterpret_cast<float4*>(const_cast<float*>(d.inertia4)), f0, rec.I4.data(), F);
    up.add(reinterpret_cast<int4*>(const_cast<int*>(d.sIdx)), s0, rec.sIdx.data(), rec.S);
    up.add(reinterpret_cast<float4*>(const_cast<float*>(d.sPar)), s0, rec.sPar.data(), rec.S);
    up.add(const_cast<int*>(d.segEnd), 2 * s0, rec.segEnd.data(), 2LL * rec.S);
    up.add(reinterpret_cast<int4*>(const_cast<int*>(d.bIdx)), b0, rec.bIdx.data(), rec.B);
    up.add(reinterpret_cast<float4*>(const_cast<float*>(d.bPar)), b0, rec.bPar.data(), rec.B);
    up.add(reinterpret_cast<float4*>(const_cast<float*>(d.bRest4)), b0, rec.bRest4.data(), rec.B);
    up.add(const_cast<Prim*>(d.prims), prim0, rec.prims.data(), (long long)rec.prims.size());
    up.add(const_cast<int*>(d.colStart), c0, rec.colStart.data(), (long long)rec.colStart.size());
    up.add(const_cast<RodInfo*>(d.rods), r0, recs.data(), (long long)R);
    up.add(d.complianceScale, r0, comp.data(), (long long)R);
    up.add(d.frictionScale, r0, fric.data(), (long long)R);
    up.add(d.selfOverflow, r0, zerosR.data(), (long long)R);
    up.add(d.clock, r0, zerosR.data(), (long long)R);
    up.add(s.stepLimit, r0, limit.data(), (long long)R);
    if ((rc = up.queue(s.staging, s.stream))) {
        release_ranges(s, q);
        return rc;
    }
    q.id = s.nextId++;
    for (size_t r = 0; r < R; ++r) {
        s.rods[size_t(r0) + r] = recs[r];
        s.selfOf[size_t(r0) + r] = recs[r].self ? (recs[r].selfOrder ? 2 : 1) : 0;
    }
    std::copy(rec.colStart.begin(), rec.colStart.end(), s.colStart.begin() + c0);
    q.steps = steps;
    s.requests[q.id] = q;
    s.dirty = true;
    *out_id = q.id;
    return RP_OK;
}

int rp_server_step(void* server, int* num_finished) {
    if (!server) return fail(RP_ERR_INVALID, "null server");
    rp_server_impl& s = *S(server);
    if (num_finished) *num_finished = 0;
    int rc0 = settle(s);
    if (rc0) return rc0;
    std::vector<int> lean, self;
    bool coloured = false;  // a running rod asks for the coloured self order
    for (const auto& kv : s.requests) {
        const Request& q = kv.second;
        if (q.stepsDone >= q.steps) continue;
        for (long long r = q.start[kRods]; r < q.start[kRods] + q.count[kRods]; ++r) {
            (s.selfOf[size_t(r)] ? self : lean).push_back(int(r));
            coloured = coloured || s.selfOf[size_t(r)] == 2;
        }
    }
    if (lean.empty() && self.empty()) return RP_OK;
    const long long R = s.pools[kRods].capacity();
    if (s.dirty) {
        Packing* packs[2] = {&s.packLean, &s.packSelf};
        Upload up;
        // Two workgroups per compute unit when every rod fits that budget.
        const std::vector<int>* ids[2] = {&lean, &self};
        const int cfgs[2] = {s.configLean, s.configSelf};
        for (int k = 0; k < 2; ++k) {
            packs[k]->tiles.clear();
            packs[k]->slots.clear();
            packs[k]->pairUnits = 0;
            if (ids[k]->empty()) continue;
            int rc = pack_rods(s.rods, *ids[k], s.colStart, cfgs[k], lds_budget(s.lim, 2), *packs[k]);
            if (rc) rc = pack_rods(s.rods, *ids[k], s.colStart, cfgs[k], lds_budget(s.lim, 1), *packs[k]);
            if (rc) return fail(rc, "resident rods do not pack");
            const TileConfig& pc = tile_config(cfgs[k]);
            packs[k]->pairUnits =
                pair_units(int(packs[k]->tiles.size()), s.lim.computeUnits,
                           tiled_resident(pc.tp, pc.tc, k == 0 ? 0 : (coloured ? 2 : 1), 64 * pc.waves,
                                          size_t(packs[k]->maxLds)));
            up.add(s.dTiles + k * R, 0, packs[k]->tiles.data(), (long long)packs[k]->tiles.size());
            up.add(s.dSlots + k * R, 0, packs[k]->slots.data(), (long long)packs[k]->slots.size());
        }
        s.dirty = false;
        // Staged and queued: the launches below follow on the same stream.
        if (int rc = up.queue(s.staging, s.stream)) return rc;
    }
    // Chunked: stepsPerLaunch steps. Time-sliced: up to the furthest horizon
    // (at most 8192 steps), the slice ending the launch.
    int launchSteps = s.stepsPerLaunch;
    if (s.sliceTicks > 0) {
        launchSteps = 1;
        for (const auto& kv : s.requests)
            launchSteps = std::max(launchSteps, kv.second.steps - kv.second.stepsDone);
        launchSteps = std::min(launchSteps, 8192);
    }
    StepArgs a = step_args<float>(s.params, launchSteps);
    a.sliceTicks = s.sliceTicks;
    // Pairing is by workgroup index within one launch: with both launches
    // sharing the device it would pair workgroups of different launches.
    const bool oneLaunch = lean.empty() != self.empty();
    if (!self.empty()) {
        // Uploads and tiles are on the main stream: the second stream starts
        // after them, and the main stream continues after its launch.
        const TileConfig& cfg = tile_config(s.configSelf);
        RP_HIP(hipEventRecord(s.evReady, s.stream));
        RP_HIP(hipStreamWaitEvent(s.selfStream, s.evReady, 0));
        a.progress = s.dProgress + R;
        a.pairUnits = oneLaunch ? s.packSelf.pairUnits : 0;
        RP_HIP(launch_tiled(cfg.tp, cfg.tc, coloured ? 2 : 1, s.d, a, s.dTiles + R, s.dSlots + R,
                            int(s.packSelf.tiles.size()), 64 * cfg.waves, size_t(s.packSelf.maxLds), s.selfStream));
        RP_HIP(hipEventRecord(s.evSelfDone, s.selfStream));
        ++s.launches;
    }
    if (!lean.empty()) {
        const TileConfig& cfg = tile_config(s.configLean);
        a.progress = s.dProgress;
        a.pairUnits = oneLaunch ? s.packLean.pairUnits : 0;
        RP_HIP(launch_tiled(cfg.tp, cfg.tc, 0, s.d, a, s.dTiles, s.dSlots, int(s.packLean.tiles.size()),
                            64 * cfg.waves, size_t(s.packLean.maxLds), s.stream));
        ++s.launches;
    }
    if (!self.empty()) RP_HIP(hipStreamWaitEvent(s.stream, s.evSelfDone, 0));
    if (s.sliceTicks > 0) {
        s.settlePending = true;
        return RP_OK;
    }
    s.stepsTaken += s.stepsPerLaunch;
    int done = 0;
    for (auto& kv : s.requests) {
        Request& q = kv.second;
        if (q.stepsDone >= q.steps) continue;
        q.stepsDone = std::min(q.steps, q.stepsDone + s.stepsPerLaunch);
        if (q.stepsDone >= q.steps) {
            s.finished.push_back(q.id);
            s.dirty = true;
            ++done;
        }
    }
    if (num_finished) *num_finished = done;
    return RP_OK;
}

int rp_server_sync(void* server) {
    if (!server) return fail(RP_ERR_INVALID, "null server");
    RP_HIP(hipStreamSynchronize(S(server)->stream));
    return settle(*S(server));
}

int rp_server_finished(void* server, int* ids, int capacity) {
    if (!server) return fail(RP_ERR_INVALID, "null server");
    rp_server_impl& s = *S(server);
    const int rc = settle(s);
    if (rc) return rc;
    const int n = int(s.finished.size());
    for (int i = 0; ids && i < std::min(n, capacity); ++i) ids[i] = s.finished[size_t(i)];
    return n;
}

int rp_server_request_info(void* server, int id, rp_request_info* out) {
    if (!server || !out) return fail(RP_ERR_INVALID, "null argument");
    rp_server_impl& s = *S(server);
    if (const int rc = settle(s)) return rc;
    auto it = s.requests.find(id);
    if (it == s.requests.end()) return fail(RP_ERR_INVALID, "unknown request " + std::to_string(id));
    const Request& q = it->second;
    out->num_rods = int(q.count[kRods]);
    out->num_particles = int(q.count[kParticles]);
    out->num_frames = int(q.count[kFrames]);
    out->steps = q.steps;
    out->steps_done = q.stepsDone;
    out->device_bytes = q.bytes;
    return RP_OK;
}

int rp_server_result(void* server, int id, float* x, float* q, float* v, float* omega) {
    if (!server) return fail(RP_ERR_INVALID, "null server");
    rp_server_impl& s = *S(server);
    if (const int rc = settle(s)) return rc;
    auto it = s.requests.find(id);
    if (it == s.requests.end()) return fail(RP_ERR_INVALID, "unknown request " + std::to_string(id));
    const Request& r = it->second;
    if (r.stepsDone < r.steps) return fail(RP_ERR_STATE, "request " + std::to_string(id) + " has not finished");
    const size_t p0 = size_t(r.start[kParticles]), f0 = size_t(r.start[kFrames]);
    const size_t P = size_t(r.count[kParticles]), F = size_t(r.count[kFrames]);
    if (x) RP_HIP(hipMemcpyAsync(x, s.d.x + p0 * 3, P * 12, hipMemcpyDeviceToHost, s.stream));
    if (v) RP_HIP(hipMemcpyAsync(v, s.d.v + p0 * 3, P * 12, hipMemcpyDeviceToHost, s.stream));
    if (q) RP_HIP(hipMemcpyAsync(q, s.d.q + f0 * 4, F * 16, hipMemcpyDeviceToHost, s.stream));
    if (omega) RP_HIP(hipMemcpyAsync(omega, s.d.omega + f0 * 3, F * 12, hipMemcpyDeviceToHost, s.stream));
    RP_HIP(hipStreamSynchronize(s.stream));
    return RP_OK;
}

int rp_server_release(void* server, int id) {
    if (!server) return fail(RP_ERR_INVALID, "null server");
    rp_server_impl& s = *S(server);
    if (const int rc = settle(s)) return rc;
    auto it = s.requests.find(id);
    if (it == s.requests.end()) return fail(RP_ERR_INVALID, "unknown request " + std::to_string(id));
    // The ranges may be reused by the next submission, whose uploads queue on
    // the stream after every launch that touches them (the second stream's
    // launch is joined to it), so nothing waits here.
    release_ranges(s, it->second);
    s.finished.erase(std::remove(s.finished.begin(), s.finished.end(), id), s.finished.end());
    if (it->second.stepsDone < it->second.steps) s.dirty = true;
    s.requests.erase(it);
    return RP_OK;
}

int rp_server_get_stats(void* server, rp_server_stats* out) {
    if (!server || !out) return fail(RP_ERR_INVALID, "null argument");
    rp_server_impl& s = *S(server);
    if (const int rc = settle(s)) return rc;
    std::memset(out, 0, sizeof(*out));
    for (const auto& kv : s.requests) {
        const Request& q = kv.second;
        ++out->resident_requests;
        if (q.stepsDone < q.steps) {
            ++out->running_requests;
            out->running_rods += int(q.count[kRods]);
        }
        out->bytes_in_use += q.bytes;
    }
    out->device_bytes = s.deviceBytes;
    out->capacity_particles = s.pools[kParticles].capacity();
    out->capacity_rods = s.pools[kRods].capacity();
    out->used_particles = s.pools[kParticles].used();
    out->used_rods = s.pools[kRods].used();
    out->launches = s.launches;
    out->workgroups = int(s.packLean.tiles.size() + s.packSelf.tiles.size());
    out->tile_config = s.configLean;
    out->steps_per_launch = s.stepsPerLaunch;
    return RP_OK;
}

}  // extern "C"

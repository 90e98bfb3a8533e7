// C++ facade over the C ABI with the method set of the reference's
// crs::gpu::Batch (src/gpu/gpu_solver.h:66-145), so the reference's call sites
// (src/gpu/gpu_cases.cpp, src/gpu/gpu_suite.cpp, src/validation/
// application_cases.cpp, tools/experiments/hang_gpu_probe.cpp) compile against
// it unchanged apart from the namespace.
//
// Header-only and templated on the caller's rod and world types: any RodT with
// the reference Rod's public fields (state.x, state.q, ..., stretch.p0, ...,
// bend.segA, ..., material.radius) and any WorldT with CollisionWorld's fields
// work, including crs::Rod and crs::CollisionWorld themselves. FP64 host values
// are rounded to FP32 at the boundary, as the reference's Batch does
// (src/gpu/gpu_batch.cpp:18-21).
#pragma once

#include <cstddef>
#include <stdexcept>
#include <string>
#include <vector>

#include "../ropephysics.h"

namespace rp {

// kFused / kMultiKernel are the reference's names (src/gpu/gpu_solver.h:50):
// the fused launch maps to the best strategy that fits (AUTO), the multi-kernel
// strategy -- global-memory state for rods too large for one workgroup's
// shared memory, which GpuSim falls back to (src/gpu/gpu_suite.cpp:125-130) --
// to GLOBAL, which takes rods of any length.
enum class Strategy {
    kPlain = RP_STRATEGY_PLAIN,
    kTiled = RP_STRATEGY_TILED,
    kAuto = RP_STRATEGY_AUTO,
    kGlobal = RP_STRATEGY_GLOBAL,
    kFused = RP_STRATEGY_AUTO,
    kMultiKernel = RP_STRATEGY_GLOBAL
};

// BatchParams (src/gpu/gpu_solver.h:52-60).
struct BatchParams {
    float dt = 1e-3f;
    int substeps = 8;
    int iterations = 1;
    float gravityX = 0, gravityY = 0, gravityZ = -9.81f;
    float linearDamping = 0;
    float angularDamping = 0;
    int contactInterval = 1;
};

struct Vec3f {
    float x = 0, y = 0, z = 0;
};

inline bool deviceAvailable() {
    rp_device_info info;
    return rp_device_query(&info) == RP_OK;
}

inline std::string deviceName() {
    rp_device_info info;
    if (rp_device_query(&info) != RP_OK) return "unknown";
    return std::string(info.name[0] ? info.name : info.arch) + " (" + info.arch + ", " +
           std::to_string(info.compute_units) + " CUs, " + std::to_string(info.lds_per_workgroup / 1024) +
           " KB LDS/workgroup)";
}

inline std::string diagnostic() {
    rp_device_info info;
    return rp_device_query(&info) == RP_OK ? std::string() : std::string(rp_last_error());
}

inline int maxTiledSegments() {
    rp_device_info info;
    return rp_device_query(&info) == RP_OK ? info.max_tiled_segments : 0;
}

class Batch {
   public:
    Batch() = default;
    ~Batch() { destroy(); }
    Batch(const Batch&) = delete;
    Batch& operator=(const Batch&) = delete;

    // Replicate `prototype` across `numRods`, optionally colliding with
    // `world`. Returns false (and lastError() says why) when the batch cannot
    // run.
    template <typename RodT, typename WorldT>
    bool create(const RodT& prototype, int numRods, Strategy strategy, const WorldT* world) {
        std::vector<const RodT*> rods(numRods, &prototype);
        return createMixed(rods, strategy, world);
    }
    template <typename RodT>
    bool create(const RodT& prototype, int numRods, Strategy strategy, std::nullptr_t) {
        return create(prototype, numRods, strategy);
    }
    template <typename RodT>
    bool create(const RodT& prototype, int numRods, Strategy strategy = Strategy::kAuto) {
        std::vector<const RodT*> rods(numRods, &prototype);
        return createMixed(rods, strategy, static_cast<const NoWorld*>(nullptr));
    }

    // Ragged extension: rods of any shapes in one batch.
    template <typename RodT, typename WorldT>
    bool createMixed(const std::vector<const RodT*>& rods, Strategy strategy, const WorldT* world) {
        destroy();
        const int R = int(rods.size());
        if (R == 0) return false;
        std::vector<int> ps(R + 1, 0), fs(R + 1, 0), ss(R + 1, 0), bs(R + 1, 0);
        for (int r = 0; r < R; ++r) {
            ps[r + 1] = ps[r] + int(rods[r]->state.x.size());
            fs[r + 1] = fs[r] + int(rods[r]->state.q.size());
            ss[r + 1] = ss[r] + int(rods[r]->stretch.p0.size());
            bs[r + 1] = bs[r] + int(rods[r]->bend.segA.size());
        }
        std::vector<float> invMass, invI, I, sRest, sComp, bRest, bComp, bRestD, radius;
        std::vector<int> sP0, sP1, sSeg, bA, bB;
        for (const RodT* rp_ : rods) {
            const auto& s = rp_->state;
            for (auto m : s.invMass) invMass.push_back(float(m));
            for (auto& v : s.invInertia) push3(invI, v);
            for (auto& v : s.inertia) push3(I, v);
            for (size_t k = 0; k < rp_->stretch.p0.size(); ++k) {
                sP0.push_back(rp_->stretch.p0[k]);
                sP1.push_back(rp_->stretch.p1[k]);
                sSeg.push_back(rp_->stretch.seg[k]);
                sRest.push_back(float(rp_->stretch.restLength[k]));
                push3(sComp, rp_->stretch.compliance[k]);
            }
            for (size_t k = 0; k < rp_->bend.segA.size(); ++k) {
                bA.push_back(rp_->bend.segA[k]);
                bB.push_back(rp_->bend.segB[k]);
                bRest.push_back(float(rp_->bend.restLength[k]));
                push3(bComp, rp_->bend.compliance[k]);
                push3(bRestD, rp_->bend.restDarboux[k]);
            }
            radius.push_back(float(rp_->material.radius));
        }
        std::vector<rp_primitive> prims;
        rp_world_desc wd{};
        if (world) fillWorld(*world, prims, wd);
        rp_batch_desc d{};
        d.num_rods = R;
        d.particle_start = ps.data();
        d.frame_start = fs.data();
        d.stretch_start = ss.data();
        d.bend_start = bs.data();
        d.inv_mass = invMass.data();
        d.inv_inertia = invI.data();
        d.inertia = I.data();
        d.stretch_p0 = sP0.data();
        d.stretch_p1 = sP1.data();
        d.stretch_seg = sSeg.data();
        d.stretch_rest = sRest.data();
        d.stretch_compliance = sComp.data();
        d.bend_a = bA.data();
        d.bend_b = bB.data();
        d.bend_rest = bRest.data();
        d.bend_compliance = bComp.data();
        d.bend_rest_darboux = bRestD.data();
        d.radius = radius.data();
        d.num_worlds = world ? 1 : 0;
        d.worlds = world ? &wd : nullptr;
        d.world_of_rod = nullptr;
        d.strategy = int(strategy);
        d.tile_config = 0;
        if (rp_batch_create(&d, &h_) != RP_OK) {
            h_ = nullptr;
            return false;
        }
        numRods_ = R;
        pStart_ = ps;
        fStart_ = fs;
        for (int r = 0; r < R; ++r) uploadRod(r, *rods[r]);
        return true;
    }

    void destroy() {
        if (h_) rp_batch_destroy(h_);
        h_ = nullptr;
        numRods_ = 0;
    }

    void step(const BatchParams& p) {
        rp_step_params sp{p.dt, p.substeps, p.iterations, {p.gravityX, p.gravityY, p.gravityZ},
                          p.linearDamping, p.angularDamping, p.contactInterval};
        check(rp_batch_step(h_, &sp, 1));
    }
    void synchronize() const { check(rp_batch_sync(h_)); }

    // Copy one rod's x, v, q, omega back into a host rod of the same topology.
    template <typename RodT>
    void download(int rodIndex, RodT& out) const {
        const int nP = pStart_[rodIndex + 1] - pStart_[rodIndex], nF = fStart_[rodIndex + 1] - fStart_[rodIndex];
        std::vector<float> x(3 * nP), v(3 * nP), q(4 * nF), w(3 * nF);
        check(rp_batch_get_state(h_, rodIndex, 1, x.data(), v.data(), q.data(), w.data()));
        for (int i = 0; i < nP; ++i) {
            set3(out.state.x[i], &x[3 * i]);
            set3(out.state.v[i], &v[3 * i]);
        }
        for (int j = 0; j < nF; ++j) {
            out.state.q[j].w = q[4 * j];
            out.state.q[j].x = q[4 * j + 1];
            out.state.q[j].y = q[4 * j + 2];
            out.state.q[j].z = q[4 * j + 3];
            set3(out.state.omega[j], &w[3 * j]);
        }
    }
    // Every rod's particle positions, rod-major, in one transfer.
    void downloadPositions(std::vector<Vec3f>& out) const {
        out.resize(size_t(pStart_[numRods_]));
        check(rp_batch_get_state(h_, 0, numRods_, &out[0].x, nullptr, nullptr, nullptr));
    }
    template <typename RodT>
    void upload(const RodT& in) {
        for (int r = 0; r < numRods_; ++r) uploadRod(r, in);
    }
    template <typename RodT>
    void uploadRod(int rodIndex, const RodT& src) {
        std::vector<float> x, v, q, w;
        for (auto& p : src.state.x) push3(x, p);
        for (auto& p : src.state.v) push3(v, p);
        for (auto& r : src.state.q) push4(q, r);
        for (auto& o : src.state.omega) push3(w, o);
        check(rp_batch_set_state(h_, rodIndex, 1, x.data(), v.data(), q.data(), w.data()));
        setLoads(rodIndex, src);
    }
    // extForce, kinematicVelocity, extTorque and the orientation of every
    // fixed (zero-inertia) frame.
    template <typename RodT>
    void setLoads(int rodIndex, const RodT& src) {
        std::vector<float> f, k, t, q;
        for (auto& p : src.state.extForce) push3(f, p);
        for (auto& p : src.state.kinematicVelocity) push3(k, p);
        for (auto& p : src.state.extTorque) push3(t, p);
        for (auto& r : src.state.q) push4(q, r);
        check(rp_batch_set_loads(h_, rodIndex, 1, f.data(), k.data(), t.data(), q.data()));
    }
    void setKinematicVelocities(const std::vector<Vec3f>& rodMajor) {
        if (rodMajor.size() != size_t(pStart_[numRods_]))
            throw std::invalid_argument("setKinematicVelocities: one velocity per particle of every rod");
        check(rp_batch_set_loads(h_, 0, numRods_, nullptr, &rodMajor[0].x, nullptr, nullptr));
    }
    void setMaterialScales(const std::vector<float>& youngsScale, const std::vector<float>& frictionScale) {
        if ((!youngsScale.empty() && youngsScale.size() != size_t(numRods_)) ||
            (!frictionScale.empty() && frictionScale.size() != size_t(numRods_)))
            throw std::invalid_argument("setMaterialScales: one value per rod");
        check(rp_batch_set_scales(h_, youngsScale.empty() ? nullptr : youngsScale.data(),
                                  frictionScale.empty() ? nullptr : frictionScale.data()));
    }

    int numRods() const { return numRods_; }
    long long selfContactOverflow() const {
        rp_counters c;
        check(rp_batch_counters(h_, &c));
        return c.self_contact_overflow;
    }
    std::size_t sharedBytesPerBlock() const {
        rp_counters c;
        check(rp_batch_counters(h_, &c));
        return c.lds_bytes_per_workgroup;
    }
    static std::string lastError() { return rp_last_error(); }
    void* handle() const { return h_; }

   private:
    struct NoWorld {
        std::vector<int> primitives;
    };

    template <typename V>
    static void push3(std::vector<float>& out, const V& v) {
        out.push_back(float(v.x));
        out.push_back(float(v.y));
        out.push_back(float(v.z));
    }
    template <typename Q>
    static void push4(std::vector<float>& out, const Q& q) {
        out.push_back(float(q.w));
        out.push_back(float(q.x));
        out.push_back(float(q.y));
        out.push_back(float(q.z));
    }
    template <typename V>
    static void set3(V& v, const float* p) {
        v.x = p[0];
        v.y = p[1];
        v.z = p[2];
    }
    static void fillWorld(const NoWorld&, std::vector<rp_primitive>&, rp_world_desc&) {}
    template <typename WorldT>
    static void fillWorld(const WorldT& w, std::vector<rp_primitive>& prims, rp_world_desc& wd) {
        for (const auto& p : w.primitives) {
            rp_primitive q{};
            q.type = int(p.type);
            const decltype(p.a)* vs[4] = {&p.a, &p.b, &p.normal, &p.halfExtents};
            float* dst[4] = {q.a, q.b, q.normal, q.half_extents};
            for (int i = 0; i < 4; ++i) {
                dst[i][0] = float(vs[i]->x);
                dst[i][1] = float(vs[i]->y);
                dst[i][2] = float(vs[i]->z);
            }
            q.rotation[0] = float(p.rotation.w);
            q.rotation[1] = float(p.rotation.x);
            q.rotation[2] = float(p.rotation.y);
            q.rotation[3] = float(p.rotation.z);
            q.radius = float(p.radius);
            q.friction = float(p.friction);
            prims.push_back(q);
        }
        wd.num_prims = int(prims.size());
        wd.prims = prims.data();
        wd.contact_margin_radii = float(w.contactMarginRadii);
        wd.self_collision = w.selfCollision ? 1 : 0;
        wd.self_friction = float(w.selfFriction);
        wd.self_skip = w.selfSkip;
        wd.hash_table_size = w.hashTableSize;
        wd.self_capacity = 0;
        wd.self_capacity_per_segment = 0.0f;
    }
    static void check(int rc) {
        if (rc != RP_OK) throw std::runtime_error(std::string("ropephysics: ") + rp_last_error());
    }

    void* h_ = nullptr;
    int numRods_ = 0;
    std::vector<int> pStart_, fStart_;
};

}  // namespace rp

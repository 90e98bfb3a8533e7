// Batched XPBD Cosserat-rod stepper kernels for gfx950 (CDNA4).
//
// Two launch shapes over the same per-constraint arithmetic (rp_math.h):
//
//  k_plain_step   one rod per 128-lane workgroup, whole state staged in LDS,
//                 one launch per step, state re-read from and written back to
//                 HBM every step. This is the reference's fused launch shape
//                 (src/gpu/gpu_solver.cu:531-681) and serves as the baseline.
//
//  k_tiled_step   rods of any length packed into lane groups of 64-lane
//                 wavefronts. Positions and orientations, which constraints
//                 share across lanes, live in LDS; everything a lane owns
//                 (velocities, substep-start state, multipliers, per-
//                 constraint constants) lives in its VGPRs for the whole
//                 launch, which runs many steps. HBM is touched once per
//                 launch. A lane group walks each colour in trips of its
//                 own width, so rods of uneven segment counts share one
//                 launch without padding to the longest.
#include <hip/hip_runtime.h>

#include <mutex>

#include "rp_layout.h"
#include "rp_math.h"
#include "rp_phases.h"

namespace rp {
namespace RP_PREC {

// Bytes of a real and of a 4-vector of reals in LDS.
constexpr int kR = int(sizeof(real));
constexpr int kR4 = int(sizeof(real4));

// Where the rod's positions go after the step that brought its clock to
// `clk`: its particles' slot in the frame this step completes, or null when
// the step completes no frame (or the rod is not recording, or its frames are
// all taken). The rod's particle start is read here rather than kept in a
// register across the step loop.
__device__ inline float* frame_slot(const DeviceArrays& d, int rod, int clk) {
    if (!d.rec) return nullptr;
    const RecRec rr = d.rec[rod];
    const int k = clk - rr.origin;
    if (rr.every <= 0 || k <= 0 || k % rr.every != 0 || k / rr.every > rr.maxFrames) return nullptr;
    return d.frames + (long long)(k / rr.every - 1) * d.framePitch + 3LL * d.rods[rod].pStart;
}

// One particle's position into a frame, rounded to FP32.
__device__ inline void frame_store(float* slot, int i, const v3& x) {
    slot[3 * i + 0] = float(x.x);
    slot[3 * i + 1] = float(x.y);
    slot[3 * i + 2] = float(x.z);
}

// ------------------------------------------------------------------ PLAIN

// LDS carve-up of one rod for the PLAIN kernel, in bytes.
struct PlainLayout {
    int x, xp, v, f, kv, q, qp, om, tq, iI, I, lamS, lamB, cont, prims, self, total;
};

__host__ __device__ inline PlainLayout plain_layout(int nP, int nF, int nS, int nB, int nPrim, int self, int table,
                                                    int cap) {
    PlainLayout L;
    int o = 0;
    L.x = o;
    o += nP * kR4;
    L.xp = o;
    o += nP * kR4;
    L.v = o;
    o += nP * kR4;
    L.f = o;
    o += nP * kR4;
    L.kv = o;
    o += nP * kR4;
    L.q = o;
    o += nF * kR4;
    L.qp = o;
    o += nF * kR4;
    L.om = o;
    o += nF * kR4;
    L.tq = o;
    o += nF * kR4;
    L.iI = o;
    o += nF * kR4;
    L.I = o;
    o += nF * kR4;
    L.lamS = o;
    o += nS * kR4;
    L.lamB = o;
    o += nB * kR4;
    L.cont = o;
    o += nP * nPrim * kContactFloats * kR;
    L.prims = o;
    o += nPrim * kPrimFloats * kR;
    L.self = o;
    if (self) o += self_lds_bytes(nS, table, cap);
    L.total = (o + kR4 - 1) & ~(kR4 - 1);
    return L;
}

// The whole step of rod blockIdx.x by one workgroup, its working arrays
// carved from `smem` as PlainLayout says: LDS for the PLAIN kernel, the rod's
// own HBM region for the GLOBAL one. With kXqLds the positions and
// orientations, which every phase reads and the colour sweeps write, live in
// LDS at `xq` instead (positions, then orientations). Every lane strides the
// rod's elements by the workgroup width, so any width works; the phases are
// separated by workgroup barriers, which order global memory within the
// workgroup as they order LDS.
template <bool kXqLds = false>
__device__ __forceinline__ void plain_rod_steps(const DeviceArrays& d, const StepArgs& sa, char* smem,
                                                char* xq = nullptr) {
    const int r = blockIdx.x;
    if (d.stepLimit && d.clock[r] >= d.stepLimit[r]) return;
    const RodInfo ri = d.rods[r];
    const int tid = threadIdx.x, nt = blockDim.x;
    const int nP = ri.nP, nF = ri.nF, nS = ri.nS, nB = ri.nB, nPrim = ri.nPrim;
    const PlainLayout L = plain_layout(nP, nF, nS, nB, nPrim, ri.self, ri.table, ri.selfCap);
    real4* x4 = reinterpret_cast<real4*>(kXqLds ? xq : smem + L.x);
    real4* xp4 = reinterpret_cast<real4*>(smem + L.xp);
    real4* v4 = reinterpret_cast<real4*>(smem + L.v);
    real4* f4 = reinterpret_cast<real4*>(smem + L.f);
    real4* kv4 = reinterpret_cast<real4*>(smem + L.kv);
    real4* q4s = reinterpret_cast<real4*>(kXqLds ? xq + nP * kR4 : smem + L.q);
    real4* qp4 = reinterpret_cast<real4*>(smem + L.qp);
    real4* om4 = reinterpret_cast<real4*>(smem + L.om);
    real4* tq4 = reinterpret_cast<real4*>(smem + L.tq);
    real4* iI4 = reinterpret_cast<real4*>(smem + L.iI);
    real4* I4 = reinterpret_cast<real4*>(smem + L.I);
    real4* lamS = reinterpret_cast<real4*>(smem + L.lamS);
    real4* lamB = reinterpret_cast<real4*>(smem + L.lamB);
    real4* cont4 = reinterpret_cast<real4*>(smem + L.cont);
    real* prims = reinterpret_cast<real*>(smem + L.prims);
    SelfLds sl;
    if (ri.self) sl = self_carve(smem + L.self, nS, ri.table, ri.selfCap);

    const int p0 = ri.pStart, f0 = ri.fStart;
    for (int i = tid; i < nP; i += nt) {
        x4[i] = vf4(ld3(d.x, p0 + i), d.invMass[p0 + i]);
        v4[i] = vf4(ld3(d.v, p0 + i), real(0.0));
        f4[i] = vf4(ld3(d.force, p0 + i), real(0.0));
        kv4[i] = vf4(ld3(d.kinVel, p0 + i), real(0.0));
    }
    for (int j = tid; j < nF; j += nt) {
        q4s[j] = qf4(ld4q(d.q, f0 + j));
        om4[j] = vf4(ld3(d.omega, f0 + j), real(0.0));
        tq4[j] = vf4(ld3(d.torque, f0 + j), real(0.0));
        iI4[j] = reinterpret_cast<const real4*>(d.invInertia4)[f0 + j];
        I4[j] = reinterpret_cast<const real4*>(d.inertia4)[f0 + j];
    }
    for (int k = tid; k < nPrim; k += nt) stage_prim(prims + k * kPrimFloats, d.prims[ri.primOff + k], ri.radius, ri.margin);
    if (ri.self && tid == 0) {
        sl.misc[0] = 0;
        sl.misc[1] = 0;
    }
    const real cscale = d.complianceScale[r];
    const real fscale = d.frictionScale[r];
    const int* colS = d.colStart + ri.colOff;
    const int* colB = colS + ri.nSCol + 1;
    const int4* sIdx = reinterpret_cast<const int4*>(d.sIdx) + ri.sStart;
    const real4* sPar = reinterpret_cast<const real4*>(d.sPar) + ri.sStart;
    const int4* bIdx = reinterpret_cast<const int4*>(d.bIdx) + ri.bStart;
    const real4* bPar = reinterpret_cast<const real4*>(d.bPar) + ri.bStart;
    const real4* bRst = reinterpret_cast<const real4*>(d.bRest4) + ri.bStart;
    const int* segEnd = d.segEnd + 2 * ri.sStart;
    const v3 g = mk(sa.gx, sa.gy, sa.gz);
    const real h = sa.h;
    int clk = d.clock[r];
    const int drivenP = d.paths ? d.paths[r].particle : -1;
    const int obsP = d.obs ? d.obs[r].particle : -1;
    double maxRise = obsP >= 0 ? d.obs[r].maxRise : 0.0;
    __syncthreads();

    for (int step = 0; step < sa.steps; ++step) {
        if (drivenP >= 0 && drivenP % nt == tid) kv4[drivenP] = vf4(path_velocity(d.paths[r], clk), real(0.0));
        for (int sub = 0; sub < sa.substeps; ++sub) {
            const bool regen = sub % sa.contactInterval == 0;
            for (int k = tid; k < nS; k += nt) lamS[k] = make_real4(real(0.), real(0.), real(0.), real(0.));
            for (int k = tid; k < nB; k += nt) lamB[k] = make_real4(real(0.), real(0.), real(0.), real(0.));
            if (nPrim > 0)
                for (int i = tid; i < nP; i += nt)
                    world_contacts_begin(cont4 + i, nP, prims, nPrim, f4v(x4[i]), x4[i].w,
                                         ri.radius, ri.margin, fscale, regen);
            if (ri.self) {
                if (regen) {
                    if (tid == 0) self_build_hash(sl, x4, segEnd, nS, ri.table, ri.radius);
                    __syncthreads();
                    for (int j = tid; j < nS; j += nt)
                        self_query(sl, x4, segEnd, nS, ri.table, ri.selfGap, ri.selfCap, ri.radius,
                                   ri.selfFriction * fscale, j, false);
                    __syncthreads();
                    if (tid == 0) self_allocate(sl, nS, ri.selfCap);
                    __syncthreads();
                    for (int j = tid; j < nS; j += nt)
                        self_query(sl, x4, segEnd, nS, ri.table, ri.selfGap, ri.selfCap, ri.radius,
                                   ri.selfFriction * fscale, j, true);
                } else if (tid == 0) {
                    self_reset(sl);
                }
                __syncthreads();
                if (regen && ri.selfOrder) {
                    if (tid == 0) self_colour(sl);
                    __syncthreads();
                }
            }
            for (int i = tid; i < nP; i += nt) {
                v3 x = f4v(x4[i]), v = f4v(v4[i]), xp;
                const real w = x4[i].w;
                predict_particle(x, v, xp, w, f4v(f4[i]), f4v(kv4[i]), h, g);
                x4[i] = vf4(x, w);
                v4[i] = vf4(v, real(0.0));
                xp4[i] = vf4(xp, w);
            }
            for (int j = tid; j < nF; j += nt) {
                q4 q = f4q(q4s[j]), qp;
                v3 om = f4v(om4[j]);
                predict_frame(q, om, qp, f4v(iI4[j]), f4v(I4[j]), f4v(tq4[j]), h);
                q4s[j] = qf4(q);
                qp4[j] = qf4(qp);
                om4[j] = vf4(om, real(0.0));
            }
            __syncthreads();

            for (int it = 0; it < sa.iterations; ++it) {
                for (int c = 0; c < ri.nSCol; ++c) {
                    for (int k = colS[c] + tid; k < colS[c + 1]; k += nt) {
                        const int4 id = sIdx[k];
                        const real4 par = sPar[k];
                        v3 xa = f4v(x4[id.x]), xb = f4v(x4[id.y]);
                        const real wa = x4[id.x].w, wb = x4[id.y].w;
                        q4 q = f4q(q4s[id.z]);
                        v3 lam = f4v(lamS[k]);
                        if (project_stretch(xa, xb, q, lam, wa, wb, f4v(iI4[id.z]), par.x, mk(par.y, par.z, par.w),
                                            sa.invH2, cscale)) {
                            x4[id.x] = vf4(xa, wa);
                            x4[id.y] = vf4(xb, wb);
                            q4s[id.z] = qf4(q);
                            lamS[k] = vf4(lam, real(0.0));
                        }
                    }
                    __syncthreads();
                }
                for (int c = 0; c < ri.nBCol; ++c) {
                    for (int k = colB[c] + tid; k < colB[c + 1]; k += nt) {
                        const int4 id = bIdx[k];
                        const real4 par = bPar[k];
                        q4 qa = f4q(q4s[id.x]), qb = f4q(q4s[id.y]);
                        v3 lam = f4v(lamB[k]);
                        if (project_bend(qa, qb, lam, f4v(iI4[id.x]), f4v(iI4[id.y]), par.x, mk(par.y, par.z, par.w),
                                         f4v(bRst[k]), sa.invH2, cscale)) {
                            q4s[id.x] = qf4(qa);
                            q4s[id.y] = qf4(qb);
                            lamB[k] = vf4(lam, real(0.0));
                        }
                    }
                    __syncthreads();
                }
                if (nPrim > 0) {
                    for (int i = tid; i < nP; i += nt) {
                        v3 x = f4v(x4[i]);
                        const real w = x4[i].w;
                        if (w <= real(0.0)) continue;
                        world_contacts_project(cont4 + i, nP, nPrim, x, f4v(xp4[i]), w);
                        x4[i] = vf4(x, w);
                    }
                    __syncthreads();
                }
                if (ri.self) {
                    if (tid == 0) {
                        if (ri.selfOrder)
                            self_project_coloured_serial(sl, x4, xp4);
                        else
                            self_project(sl, x4, xp4);
                    }
                    __syncthreads();
                }
            }

            for (int i = tid; i < nP; i += nt) {
                const v3 v = finish_particle(f4v(x4[i]), f4v(xp4[i]), x4[i].w, f4v(kv4[i]), h, sa.linDecay);
                v4[i] = vf4(v, real(0.0));
            }
            for (int j = tid; j < nF; j += nt) {
                const v3 om = finish_frame(f4q(q4s[j]), f4q(qp4[j]), f4v(iI4[j]), h, sa.angDecay);
                om4[j] = vf4(om, real(0.0));
            }
            __syncthreads();
        }
        ++clk;
        if (obsP >= 0 && obsP % nt == tid && clk % d.obs[r].every == 0) {
            const double rise = double(x4[obsP].z) - d.obs[r].z0;
            maxRise = maxRise < rise ? rise : maxRise;
        }
        if (float* slot = frame_slot(d, r, clk))
            for (int i = tid; i < nP; i += nt) frame_store(slot, i, f4v(x4[i]));
    }
    if (tid == 0) d.clock[r] = clk;
    if (obsP >= 0 && obsP % nt == tid) d.obs[r].maxRise = maxRise;

    for (int i = tid; i < nP; i += nt) {
        st3(d.x, p0 + i, f4v(x4[i]));
        st3(d.v, p0 + i, f4v(v4[i]));
    }
    for (int j = tid; j < nF; j += nt) {
        st4q(d.q, f0 + j, f4q(q4s[j]));
        st3(d.omega, f0 + j, f4v(om4[j]));
    }
    if (ri.self && tid == 0) d.selfOverflow[r] += sl.misc[1];
}

__global__ __launch_bounds__(128) void k_plain_step(DeviceArrays d, StepArgs sa) {
    extern __shared__ real4 smem4[];
    if (int(blockIdx.x) >= d.numRods) return;
    plain_rod_steps(d, sa, reinterpret_cast<char*>(smem4));
}

// ------------------------------------------------------------------ GLOBAL

// Lanes of a GLOBAL workgroup at most (two waves per SIMD).
constexpr int kGlobalMaxThreads = 512;

// The PLAIN step with the rod's working arrays in HBM (rod r's region starts
// at scratch + scratchOff[r]), for rods whose arrays do not fit a workgroup's
// LDS: the reference's multi-kernel strategy keeps state in global memory for
// the same reason (src/gpu/gpu_solver.cu:417-529,691-730). Here one
// workgroup still steps one rod through every phase, with barriers where the
// reference launches a kernel per phase and colour, and a launch runs many
// steps. With XQ the rod's positions and orientations are held in LDS for
// the launch (the launch's dynamic LDS covers every rod's), and only the
// other arrays are read from and written to HBM in every phase.
template <bool XQ>
__global__ __launch_bounds__(kGlobalMaxThreads) void k_global_step(DeviceArrays d, StepArgs sa) {
    extern __shared__ real4 xq4[];
    if (int(blockIdx.x) >= d.numRods) return;
    plain_rod_steps<XQ>(d, sa, d.scratch + d.scratchOff[blockIdx.x], reinterpret_cast<char*>(xq4));
}

// ------------------------------------------------------------------ TILED

// LDS carve-up of one rod slot for the TILED kernel, in bytes.
struct TiledLayout {
    int x, q, f, kv, tq, cont, prims, xp, self, total;
};

__host__ __device__ inline TiledLayout tiled_layout(int nP, int nF, int nS, int nPrim, int self, int table, int cap) {
    TiledLayout L;
    int o = 0;
    L.x = o;
    o += nP * kR4;
    L.q = o;
    o += nF * kR4;
    L.f = o;
    o += nP * kR4;
    L.kv = o;
    o += nP * kR4;
    L.tq = o;
    o += nF * kR4;
    L.cont = o;
    o += nP * nPrim * kContactFloats * kR;
    L.prims = o;
    o += nPrim * kPrimFloats * kR;
    L.xp = o;
    if (self) o += nP * kR4;
    L.self = o;
    if (self) o += self_lds_bytes(nS, table, cap);
    L.total = (o + kR4 - 1) & ~(kR4 - 1);
    return L;
}

// TP: trips over particles and frames; TC: trips per colour per family.
// A lane group of G lanes covers up to TP*G particles and frames and TC*G
// constraints of each colour; the host sizes G per rod to satisfy that.
// SELF: 0, compiled without rod self-collision (batches without it run this
// variant, whose smaller register footprint keeps two waves per SIMD); 1, with
// self-collision in the reference's list order; 2, also with the coloured
// order for the rods that ask for it (batches holding such a rod run this
// variant, so the others carry none of its code).
//
// In FP32, shapes with TP <= 2 are held to 256 registers, so two waves fit
// on a SIMD (the self-collision variant spills a few bytes to do so), and TP
// = 1 also runs in 512-lane workgroups, for rods of up to 511 particles at
// one per lane. In FP64, where every real takes two registers, workgroups
// have at most 256 lanes and a wave may use the whole register file.
constexpr int tiled_max_lanes(int tp) { return (!RP_FP64 && tp == 1) ? 512 : 256; }
constexpr int tiled_waves_per_simd(int tp) { return (!RP_FP64 && tp <= 2) ? 2 : 1; }

// Issue priority of a workgroup that shares its compute unit with another of
// the same launch (b and b +- pairUnits). The SIMDs issue the older
// workgroup's waves first, so without this the younger one falls behind and
// finishes last, alone on the unit. At each of a launch's first
// kPriorityEvery steps, and every kPriorityEvery steps after, each workgroup
// publishes its step count and reads its partner's; the one behind raises its
// waves' priority until the next exchange, and at equal counts the two take
// turns on the wall clock. Only the issue order changes, never what a lane
// computes, so a stale count (the partner's previous launch, or a value not
// yet visible) costs scheduling quality and nothing else. The counts move as
// relaxed atomics, which reach the memory side: one 32-byte access per
// workgroup and per wave at each exchange, about 10 MB per 6000-step launch
// of 490 workgroups (exchanging at every step moved 400 MB for 1% less
// time).
constexpr int kPriorityEvery = 64;

__device__ __forceinline__ void prioritise_lagging(const StepArgs& sa, int step) {
    const int b = blockIdx.x;
    const int partner = b < sa.pairUnits ? b + sa.pairUnits : b - sa.pairUnits;
    if (threadIdx.x == 0)
        (void)__hip_atomic_exchange(sa.progress + b, step, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    bool behind = false;
    if (partner < int(gridDim.x)) {
        int other = 0;
        if ((threadIdx.x & 63) == 0)
            other = __hip_atomic_fetch_add(sa.progress + partner, 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        other = __builtin_amdgcn_readfirstlane(other);
        behind = other > step || (other == step && (((wall_clock64() >> 14) ^ unsigned(b / sa.pairUnits)) & 1u));
    }
    if (behind)
        __builtin_amdgcn_s_setprio(1);
    else
        __builtin_amdgcn_s_setprio(0);
}

template <int TP, int TC, int SELF>
__global__ __launch_bounds__(tiled_max_lanes(TP)) __attribute__((amdgpu_waves_per_eu(tiled_waves_per_simd(TP)))) void
k_tiled_step(DeviceArrays d, StepArgs sa, const Tile* tiles, const TileSlot* slots) {
    extern __shared__ real4 smem4[];
    char* smem = reinterpret_cast<char*>(smem4);
    const Tile tile = tiles[blockIdx.x];
    const int lane = threadIdx.x;

    // Find this lane's rod slot.
    int rod = -1, l = 0, G = 1, ldsOff = 0;
    for (int s = 0; s < tile.numSlots; ++s) {
        const TileSlot ts = slots[tile.firstSlot + s];
        if (lane >= ts.lane0 && lane < ts.lane0 + ts.lanes) {
            rod = ts.rod;
            l = lane - ts.lane0;
            G = ts.lanes;
            ldsOff = ts.lds;
        }
    }
    const bool active = rod >= 0;
    RodInfo ri;
    if (active) {
        ri = d.rods[rod];
    } else {
        ri.nP = ri.nF = ri.nS = ri.nB = 0;
        ri.nPrim = 0;
        ri.self = 0;
        ri.table = 1;
        ri.selfCap = 0;
        ri.pStart = ri.fStart = ri.sStart = ri.bStart = 0;
        ri.colOff = 0;
        ri.radius = ri.margin = ri.selfFriction = real(0.0);
        ri.selfGap = 1;
        ri.primOff = 0;
    }
    const int nP = ri.nP, nF = ri.nF, nS = ri.nS, nPrim = ri.nPrim;
    const TiledLayout L = tiled_layout(nP, nF, nS, nPrim, ri.self, ri.table, ri.selfCap);
    char* base = smem + ldsOff;
    // Positions with inverse mass. A self-colliding rod keeps two buffers and
    // alternates them: each substep predicts from one (which then holds the
    // substep-start positions for the rest of the substep) into the other.
    // Its broadphase can then read start positions before and after predict
    // with no barrier between. Other rods update one buffer in place.
    real4* const x4a = reinterpret_cast<real4*>(base + L.x);
    real4* const x4b = reinterpret_cast<real4*>(base + L.xp);
    real4* q4s = reinterpret_cast<real4*>(base + L.q);
    real4* f4 = reinterpret_cast<real4*>(base + L.f);
    real4* kv4 = reinterpret_cast<real4*>(base + L.kv);
    real4* tq4 = reinterpret_cast<real4*>(base + L.tq);
    real4* cont4 = reinterpret_cast<real4*>(base + L.cont);
    real* prims = reinterpret_cast<real*>(base + L.prims);
    SelfLds sl;
    if (SELF && ri.self) sl = self_carve(base + L.self, nS, ri.table, ri.selfCap);
    // Any lane group in the workgroup may run self-collision; the barriers
    // around it must be reached by every lane of the workgroup.
    const bool anySelf = SELF && __syncthreads_or(ri.self) != 0;
    const bool anyColoured = SELF == 2 && anySelf && __syncthreads_or(ri.self && ri.selfOrder) != 0;
    const bool anyPrim = __syncthreads_or(nPrim > 0) != 0;

    const int p0 = ri.pStart, f0 = ri.fStart;
    // ---- lane-owned state and constants, loaded once per launch
    v3 vel[TP], xpv[TP];
    real wP[TP];
    v3 om[TP], iIF[TP], IF[TP];
    q4 qpv[TP];
#pragma unroll
    for (int t = 0; t < TP; ++t) {
        const int i = l + G * t;
        if (active && i < nP) {
            vel[t] = ld3(d.v, p0 + i);
            wP[t] = d.invMass[p0 + i];
            x4a[i] = vf4(ld3(d.x, p0 + i), wP[t]);
            f4[i] = vf4(ld3(d.force, p0 + i), real(0.0));
            kv4[i] = vf4(ld3(d.kinVel, p0 + i), real(0.0));
        }
        const int j = l + G * t;
        if (active && j < nF) {
            om[t] = ld3(d.omega, f0 + j);
            const real4 a = reinterpret_cast<const real4*>(d.invInertia4)[f0 + j];
            const real4 b = reinterpret_cast<const real4*>(d.inertia4)[f0 + j];
            iIF[t] = f4v(a);
            IF[t] = f4v(b);
            q4s[j] = qf4(ld4q(d.q, f0 + j));
            tq4[j] = vf4(ld3(d.torque, f0 + j), real(0.0));
        }
    }
    for (int k = l; active && k < nPrim; k += G) stage_prim(prims + k * kPrimFloats, d.prims[ri.primOff + k], ri.radius, ri.margin);

    // Stretch constraints: colour c, trip t -> position colS[c] + l + G t.
    // Each constraint's fixed quotients and scaled compliance are formed here
    // once (StretchConst, BendConst); a constraint whose participants are all
    // fixed is off.
    int sA[2][TC], sB[2][TC], sF[2][TC];
    StretchConst sK[2][TC];
    v3 sAlpha[2][TC], sIIv[2][TC], sLam[2][TC];
    bool sOn[2][TC], sRot[2][TC];
    int bA_[2][TC], bB_[2][TC];
    BendConst bK[2][TC];
    v3 bAlpha[2][TC], bRD[2][TC], bIa[2][TC], bIb[2][TC], bLam[2][TC];
    bool bOn[2][TC], bRotA[2][TC], bRotB[2][TC];
    const int* colS = d.colStart + ri.colOff;
    const int* colB = colS + ri.nSCol + 1;
    const real cscale = active ? d.complianceScale[rod] : real(1.0);
#pragma unroll
    for (int c = 0; c < 2; ++c)
#pragma unroll
        for (int t = 0; t < TC; ++t) {
            sOn[c][t] = false;
            bOn[c][t] = false;
            if (active && c < ri.nSCol) {
                const int k = colS[c] + l + G * t;
                if (k < colS[c + 1]) {
                    const int4 id = reinterpret_cast<const int4*>(d.sIdx)[ri.sStart + k];
                    const real4 par = reinterpret_cast<const real4*>(d.sPar)[ri.sStart + k];
                    const real w0 = d.invMass[p0 + id.x], w1 = d.invMass[p0 + id.y];
                    sA[c][t] = id.x;
                    sB[c][t] = id.y;
                    sF[c][t] = id.z;
                    sIIv[c][t] = f4v(reinterpret_cast<const real4*>(d.invInertia4)[f0 + id.z]);
                    sOn[c][t] = !stretch_fixed(w0, w1, sIIv[c][t]);
                    sRot[c][t] = norm2(sIIv[c][t]) > real(0.0);
                    sK[c][t] = stretch_const(w0, w1, par.x);
                    sAlpha[c][t] = scaled_compliance(mk(par.y, par.z, par.w), sa.invH2, cscale);
                }
            }
            if (active && c < ri.nBCol) {
                const int k = colB[c] + l + G * t;
                if (k < colB[c + 1]) {
                    const int4 id = reinterpret_cast<const int4*>(d.bIdx)[ri.bStart + k];
                    const real4 par = reinterpret_cast<const real4*>(d.bPar)[ri.bStart + k];
                    bA_[c][t] = id.x;
                    bB_[c][t] = id.y;
                    bRD[c][t] = f4v(reinterpret_cast<const real4*>(d.bRest4)[ri.bStart + k]);
                    bIa[c][t] = f4v(reinterpret_cast<const real4*>(d.invInertia4)[f0 + id.x]);
                    bIb[c][t] = f4v(reinterpret_cast<const real4*>(d.invInertia4)[f0 + id.y]);
                    bOn[c][t] = !bend_fixed(bIa[c][t], bIb[c][t]);
                    bRotA[c][t] = norm2(bIa[c][t]) > real(0.0);
                    bRotB[c][t] = norm2(bIb[c][t]) > real(0.0);
                    bK[c][t] = bend_const(par.x);
                    bAlpha[c][t] = scaled_compliance(mk(par.y, par.z, par.w), sa.invH2, cscale);
                }
            }
        }
    const real fscale = active ? d.frictionScale[rod] : real(1.0);
    // Null for a rod whose segment k joins particles k and k + 1 (seg_a, seg_b).
    const int* segEnd = (ri.selfChain & 2) ? nullptr : d.segEnd + 2 * ri.sStart;
    if (SELF && ri.self && l == 0) {
        sl.misc[0] = 0;
        sl.misc[1] = 0;
        self_regen_reset(sl);
        self_list_reset(sl, nS);
        self_candidates_reset(sl);
    }
    // How far this lane's particles moved since the rod's candidate lists were
    // last rebuilt (a bound: per substep the largest displacement, summed),
    // and whether that bound has invalidated the lists (self_lists_stale).
    real moved = real(0.0);
    bool stale = false;
    // Which position buffer holds the current positions (double-buffered
    // rods); it flips at every substep the rod takes.
    bool flip = false;
    const bool dbl = SELF && ri.self;
    const v3 g = mk(sa.gx, sa.gy, sa.gz);
    const real h = sa.h;
    int clk = active ? d.clock[rod] : 0;
    // Whether some rod of the workgroup had self contacts at the last
    // regeneration (the same on every lane).
    bool wgContacts = false;
    // Per particle of the lane, its world contact slots that hold a contact
    // (bit k for primitive k), set at every regeneration.
    unsigned wAct[TP];
#pragma unroll
    for (int t = 0; t < TP; ++t) wAct[t] = 0u;
    // One iteration with contacts regenerated at every substep: the world
    // contacts are formed and projected in the sweep, with no slots.
    const bool fusedWorld = sa.iterations == 1 && sa.contactInterval == 1;
    // A rod whose clock has reached its step limit stops changing; its lanes
    // still take part in every barrier.
    const int limit = (active && d.stepLimit) ? d.stepLimit[rod] : 0x7fffffff;
    const int drivenP = (active && d.paths) ? d.paths[rod].particle : -1;
    const bool driver = drivenP >= 0 && drivenP % G == l;
    const int obsP = (active && d.obs) ? d.obs[rod].particle : -1;
    const bool observer = obsP >= 0 && obsP % G == l;
    double maxRise = observer ? d.obs[rod].maxRise : 0.0;
    const long long sliceStart = sa.sliceTicks > 0 ? (long long)wall_clock64() : 0;
    __syncthreads();

    for (int step = 0; step < sa.steps; ++step) {
        // Time-sliced launches: the workgroup decides together, on thread 0's
        // clock, whether to start another step, and stops once none of its
        // rods has steps left.
        if (sa.sliceTicks > 0 && step > 0 &&
            (__syncthreads_or(threadIdx.x == 0 && (long long)wall_clock64() - sliceStart >= sa.sliceTicks) ||
             !__syncthreads_or(active && clk < limit)))
            break;
        // Addresses that depend on the rod (its path and observer records) are
        // formed where they are used, not held in registers across the loop.
        asm volatile("" : "+v"(rod));
        const bool live = active && clk < limit;
        if (sa.pairUnits > 0 && (step < kPriorityEvery || step % kPriorityEvery == 0)) prioritise_lagging(sa, step);
        if (driver && live) kv4[drivenP] = vf4(path_velocity(d.paths[rod], clk), real(0.0));
        for (int sub = 0; sub < sa.substeps; ++sub) {
            const bool regen = sub % sa.contactInterval == 0;
            // xs: positions at the start of the substep; x4: the positions the
            // substep moves (the same buffer for single-buffered rods).
            real4* const xs = dbl && flip ? x4b : x4a;
            real4* const x4 = dbl ? (flip ? x4a : x4b) : x4a;
#pragma unroll
            for (int c = 0; c < 2; ++c)
#pragma unroll
                for (int t = 0; t < TC; ++t) {
                    sLam[c][t] = mk(real(0.), real(0.), real(0.));
                    bLam[c][t] = mk(real(0.), real(0.), real(0.));
                }
            // Self contacts: each lane first checks, with no barrier, whether its
            // segments need the full regeneration (self_quick_check); the vote
            // at the barrier after predict decides whether the workgroup runs
            // it, on the substep-start positions. The check comes before the
            // world contacts, so its LDS reads do not wait behind their slot
            // writes.
            bool needFull = false;
            if (anySelf) {
                if (regen)
                    needFull = self_quick_check<TP>(sl, xs, segEnd, nS, ri.selfGap, ri.radius, ri.self != 0 && live,
                                                    stale, l, G);
                else if (ri.self && l == 0 && live)
                    self_reset(sl);
            }
            if (anyPrim && nPrim > 0 && !fusedWorld) {
#pragma unroll
                for (int t = 0; t < TP; ++t) {
                    const int i = l + G * t;
                    if (live && i < nP)
                        wAct[t] = world_contacts_begin_masked(cont4 + i, nP, prims, nPrim, f4v(xs[i]), wP[t],
                                                              ri.radius, ri.margin, fscale, regen, wAct[t]);
                }
            }
            // ---- predict
#pragma unroll
            for (int t = 0; t < TP; ++t) {
                const int i = l + G * t;
                if (live && i < nP) {
                    v3 x = f4v(xs[i]);
                    predict_particle(x, vel[t], xpv[t], wP[t], f4v(f4[i]), f4v(kv4[i]), h, g);
                    x4[i] = vf4(x, wP[t]);
                }
                const int j = l + G * t;
                if (live && j < nF) {
                    q4 q = f4q(q4s[j]);
                    predict_frame(q, om[t], qpv[t], iIF[t], IF[t], f4v(tq4[j]), h);
                    q4s[j] = qf4(q);
                }
            }
            if (anySelf && regen) {
                if (__syncthreads_or(needFull)) {
                    bool rebuilt = false;
                    wgContacts = self_regen_group<TP>(sl, xs, segEnd, nS, ri.table, ri.selfGap, ri.selfCap,
                                                      ri.radius, ri.selfFriction * fscale, (ri.selfChain & 1) != 0,
                                                      ri.self != 0 && live, l, G, rebuilt);
                    if (rebuilt) {
                        moved = real(0.0);
                        stale = false;
                    }
                } else {
                    wgContacts = false;
                }
                // The coloured order colours the contacts the fill wrote.
                if (anyColoured && wgContacts) {
                    __syncthreads();
                    if (ri.self && ri.selfOrder && live && l == 0) self_colour(sl);
                    __syncthreads();
                }
            } else {
                __syncthreads();
            }

            for (int it = 0; it < sa.iterations; ++it) {
#pragma unroll
                for (int c = 0; c < 2; ++c) {
#pragma unroll
                    for (int t = 0; t < TC; ++t) {
                        if (live && sOn[c][t]) {
                            const real4 ax = x4[sA[c][t]], bx = x4[sB[c][t]];
                            v3 xa = f4v(ax), xb = f4v(bx);
                            q4 q = f4q(q4s[sF[c][t]]);
                            if (project_stretch_k(xa, xb, q, sLam[c][t], sK[c][t], sIIv[c][t], sRot[c][t],
                                                  sAlpha[c][t])) {
                                x4[sA[c][t]] = vf4(xa, ax.w);
                                x4[sB[c][t]] = vf4(xb, bx.w);
                                q4s[sF[c][t]] = qf4(q);
                            }
                        }
                    }
                    __syncthreads();
                }
#pragma unroll
                for (int c = 0; c < 2; ++c) {
#pragma unroll
                    for (int t = 0; t < TC; ++t) {
                        if (live && bOn[c][t]) {
                            q4 qa = f4q(q4s[bA_[c][t]]), qb = f4q(q4s[bB_[c][t]]);
                            if (project_bend_k(qa, qb, bLam[c][t], bIa[c][t], bIb[c][t], bK[c][t], bAlpha[c][t],
                                               bRD[c][t], bRotA[c][t], bRotB[c][t])) {
                                q4s[bA_[c][t]] = qf4(qa);
                                q4s[bB_[c][t]] = qf4(qb);
                            }
                        }
                    }
                    // After the last bend colour the next readers of the
                    // orientations are finish, a self projection or the next
                    // sweep. When the contact sweep, which reads and writes
                    // only positions and contact slots, ends with a barrier,
                    // that barrier orders them.
                    if (c == 0 || !(anyPrim && (anySelf || it + 1 < sa.iterations))) __syncthreads();
                }
                if (anyPrim) {
                    if (nPrim > 0) {
#pragma unroll
                        for (int t = 0; t < TP; ++t) {
                            const int i = l + G * t;
                            if (live && i < nP && wP[t] > real(0.0)) {
                                v3 x = f4v(x4[i]);
                                if (fusedWorld)
                                    world_contacts_fused(prims, nPrim, x, xpv[t], wP[t], ri.radius, ri.margin, fscale);
                                else
                                    world_contacts_project_masked(cont4 + i, nP, wAct[t], x, xpv[t], wP[t]);
                                x4[i] = vf4(x, wP[t]);
                            }
                        }
                    }
                    // Each lane moved only its own particles. Other lanes read
                    // them next in the self-contact projection, the next sweep
                    // or the next regeneration; after the last sweep of a
                    // workgroup without self-collision, only the same lanes read
                    // them (finish, then predict, which ends with a barrier).
                    if (anySelf || it + 1 < sa.iterations) __syncthreads();
                }
                // Self contacts exist in the workgroup only when the last
                // regeneration's vote said so; otherwise nothing is projected.
                if (anySelf && wgContacts) {
                    if (ri.self && !ri.selfOrder && l == 0 && live) self_project(sl, x4, xs);
                    if (anyColoured)
                        self_project_coloured(sl, x4, xs, ri.self && ri.selfOrder && live, l, G);
                    __syncthreads();
                }
            }

            // ---- finish: velocities from positions; nothing shared is written
            real moved2 = real(0.0);
#pragma unroll
            for (int t = 0; t < TP; ++t) {
                const int i = l + G * t;
                if (live && i < nP) {
                    const v3 x = f4v(x4[i]);
                    vel[t] = finish_particle(x, xpv[t], wP[t], f4v(kv4[i]), h, sa.linDecay);
                    if (SELF) moved2 = r_max(moved2, norm2(mk(x.x - xpv[t].x, x.y - xpv[t].y, x.z - xpv[t].z)));
                }
                const int j = l + G * t;
                if (live && j < nF) om[t] = finish_frame(f4q(q4s[j]), qpv[t], iIF[t], h, sa.angDecay);
            }
            // Rounded up past the rounding of the distance.
            if (SELF) moved += r_sqrt(moved2) * real(1.000001);
            if (SELF && ri.self && live && !stale && self_lists_stale(sl, moved, ri.radius)) {
                stale = true;
                sl.rmisc[7] = 0;
            }
            if (live) flip = !flip;
        }
        // The rod's current positions.
        real4* const x4 = dbl && flip ? x4b : x4a;
        if (!live) continue;
        ++clk;
        if (observer && clk % d.obs[rod].every == 0) {
            const double rise = double(x4[obsP].z) - d.obs[rod].z0;
            maxRise = maxRise < rise ? rise : maxRise;
        }
        if (float* slot = frame_slot(d, rod, clk)) {
#pragma unroll
            for (int t = 0; t < TP; ++t) {
                const int i = l + G * t;
                if (i < nP) frame_store(slot, i, f4v(x4[i]));
            }
        }
    }

    __syncthreads();
    // The write-back addresses are formed here, after the loop.
    int p0e = p0, f0e = f0;
    asm volatile("" : "+v"(p0e), "+v"(f0e), "+v"(rod));
    const real4* const x4f = dbl && flip ? x4b : x4a;
#pragma unroll
    for (int t = 0; t < TP; ++t) {
        const int i = l + G * t;
        if (active && i < nP) {
            st3(d.x, p0e + i, f4v(x4f[i]));
            st3(d.v, p0e + i, vel[t]);
        }
        const int j = l + G * t;
        if (active && j < nF) {
            st4q(d.q, f0e + j, f4q(q4s[j]));
            st3(d.omega, f0e + j, om[t]);
        }
    }
    if (SELF && ri.self && l == 0) d.selfOverflow[rod] += sl.misc[1];
    if (active && l == 0) d.clock[rod] = clk;
    if (observer) d.obs[rod].maxRise = maxRise;
}

// Copy the orientation of zero-inertia frames from `src` (frames * 4, packed
// for frames [f0, f0 + count)).
__global__ void k_apply_fixed_frames(real* q, const real* src, const real* invInertia4, int f0, int count) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= count) return;
    const real4 iI = reinterpret_cast<const real4*>(invInertia4)[f0 + j];
    if (norm2(mk(iI.x, iI.y, iI.z)) == real(0.0))
        for (int c = 0; c < 4; ++c) q[4 * (f0 + j) + c] = src[4 * j + c];
}

}  // namespace RP_PREC
}  // namespace rp

// ------------------------------------------------------------------ host-side launch wrappers

namespace rp {
namespace RP_PREC {

int plain_lds_bytes(int nP, int nF, int nS, int nB, int nPrim, int self, int table, int cap) {
    return plain_layout(nP, nF, nS, nB, nPrim, self, table, cap).total;
}

int tiled_lds_bytes(int nP, int nF, int nS, int nPrim, int self, int table, int cap) {
    return tiled_layout(nP, nF, nS, nPrim, self, table, cap).total;
}

hipError_t launch_plain(const DeviceArrays& d, const StepArgs& sa, size_t lds, hipStream_t s) {
    hipLaunchKernelGGL(k_plain_step, dim3(d.numRods), dim3(128), lds, s, d, sa);
    return hipGetLastError();
}

int global_max_threads() { return kGlobalMaxThreads; }

hipError_t launch_global(const DeviceArrays& d, const StepArgs& sa, int threads, size_t xqLds, hipStream_t s) {
    if (threads < 64 || threads % 64 != 0 || threads > kGlobalMaxThreads || !d.scratch || !d.scratchOff)
        return hipErrorInvalidValue;
    if (xqLds > 0)
        hipLaunchKernelGGL(k_global_step<true>, dim3(d.numRods), dim3(threads), xqLds, s, d, sa);
    else
        hipLaunchKernelGGL(k_global_step<false>, dim3(d.numRods), dim3(threads), 0, s, d, sa);
    return hipGetLastError();
}

template <int TP, int TC>
static hipError_t launch_tiled_t(int self, const DeviceArrays& d, const StepArgs& sa, const Tile* tiles,
                                 const TileSlot* slots, int numTiles, int threads, size_t lds, hipStream_t s) {
    if (self == 2)
        hipLaunchKernelGGL((k_tiled_step<TP, TC, 2>), dim3(numTiles), dim3(threads), lds, s, d, sa, tiles, slots);
    else if (self == 1)
        hipLaunchKernelGGL((k_tiled_step<TP, TC, 1>), dim3(numTiles), dim3(threads), lds, s, d, sa, tiles, slots);
    else if (self == 0)
        hipLaunchKernelGGL((k_tiled_step<TP, TC, 0>), dim3(numTiles), dim3(threads), lds, s, d, sa, tiles, slots);
    else
        return hipErrorInvalidValue;
    return hipGetLastError();
}

// Raise a kernel's dynamic LDS limit to `bytes` less its static LDS (the
// wave-vote barriers keep a little); lowers `*granted` to what it got.
static hipError_t allow_lds(const void* kernel, int bytes, int* granted) {
    hipFuncAttributes a;
    hipError_t e = hipFuncGetAttributes(&a, kernel);
    if (e != hipSuccess) return e;
    const int dyn = bytes - int(a.sharedSizeBytes);
    e = hipFuncSetAttribute(kernel, hipFuncAttributeMaxDynamicSharedMemorySize, dyn);
    if (e == hipSuccess && dyn < *granted) *granted = dyn;
    return e;
}

template <int TP, int TC>
static hipError_t allow_tiled_lds(int bytes, int* granted) {
    hipError_t e = allow_lds(reinterpret_cast<const void*>(&k_tiled_step<TP, TC, 2>), bytes, granted);
    if (e == hipSuccess) e = allow_lds(reinterpret_cast<const void*>(&k_tiled_step<TP, TC, 1>), bytes, granted);
    if (e != hipSuccess) return e;
    return allow_lds(reinterpret_cast<const void*>(&k_tiled_step<TP, TC, 0>), bytes, granted);
}

// Let every stepping kernel take up to `bytes` of LDS per workgroup, static
// and dynamic together; `*granted` receives the dynamic bytes every one of
// them may use.
hipError_t allow_dynamic_lds(int bytes, int* granted) {
    *granted = bytes;
    hipError_t e = allow_lds(reinterpret_cast<const void*>(&k_plain_step), bytes, granted);
    if (e == hipSuccess) e = allow_lds(reinterpret_cast<const void*>(&k_global_step<true>), bytes, granted);
    if (e == hipSuccess) e = allow_tiled_lds<1, 1>(bytes, granted);
    if (e == hipSuccess) e = allow_tiled_lds<2, 1>(bytes, granted);
    if (e == hipSuccess) e = allow_tiled_lds<3, 1>(bytes, granted);
#if !RP_FP64
    if (e == hipSuccess) e = allow_tiled_lds<3, 2>(bytes, granted);
    if (e == hipSuccess) e = allow_tiled_lds<5, 2>(bytes, granted);
#endif
    return e;
}

// Trip shapes compiled into the library; (TP, TC) pairs. In FP64 the
// two-trip shapes are left out: their state does not fit the register file.
hipError_t launch_tiled(int tp, int tc, int self, const DeviceArrays& d, const StepArgs& sa, const Tile* tiles,
                        const TileSlot* slots, int numTiles, int threads, size_t lds, hipStream_t s) {
    if (threads < 64 || threads % 64 != 0 || threads > tiled_max_lanes(tp)) return hipErrorInvalidValue;
    if (tp == 1 && tc == 1) return launch_tiled_t<1, 1>(self, d, sa, tiles, slots, numTiles, threads, lds, s);
    if (tp == 2 && tc == 1) return launch_tiled_t<2, 1>(self, d, sa, tiles, slots, numTiles, threads, lds, s);
    if (tp == 3 && tc == 1) return launch_tiled_t<3, 1>(self, d, sa, tiles, slots, numTiles, threads, lds, s);
#if !RP_FP64
    if (tp == 3 && tc == 2) return launch_tiled_t<3, 2>(self, d, sa, tiles, slots, numTiles, threads, lds, s);
    if (tp == 5 && tc == 2) return launch_tiled_t<5, 2>(self, d, sa, tiles, slots, numTiles, threads, lds, s);
#endif
    return hipErrorInvalidValue;
}

bool tiled_shape_available(int tp, int tc, int threads) {
    if (threads < 64 || threads % 64 != 0 || threads > tiled_max_lanes(tp)) return false;
    if (tc == 1) return tp >= 1 && tp <= 3;
    return !RP_FP64 && tc == 2 && (tp == 3 || tp == 5);
}

template <int TP, int TC>
static int tiled_resident_t(int self, int threads, size_t lds) {
    const void* k = self == 2   ? reinterpret_cast<const void*>(&k_tiled_step<TP, TC, 2>)
                    : self == 1 ? reinterpret_cast<const void*>(&k_tiled_step<TP, TC, 1>)
                                : reinterpret_cast<const void*>(&k_tiled_step<TP, TC, 0>);
    int n = 0;
    return hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k, threads, lds) == hipSuccess ? n : 0;
}

// Workgroups of a TILED launch that fit on one compute unit at once, as its
// registers and `lds` bytes of dynamic LDS allow; 0 when unknown.
int tiled_resident(int tp, int tc, int self, int threads, size_t lds) {
    if (!tiled_shape_available(tp, tc, threads) || self < 0 || self > 2) return 0;
    if (tp == 1 && tc == 1) return tiled_resident_t<1, 1>(self, threads, lds);
    if (tp == 2 && tc == 1) return tiled_resident_t<2, 1>(self, threads, lds);
    if (tp == 3 && tc == 1) return tiled_resident_t<3, 1>(self, threads, lds);
#if !RP_FP64
    if (tp == 3 && tc == 2) return tiled_resident_t<3, 2>(self, threads, lds);
    if (tp == 5 && tc == 2) return tiled_resident_t<5, 2>(self, threads, lds);
#endif
    return 0;
}

#if !RP_FP64
// The working arrays of GLOBAL steps taken through rp_batch_step_noalloc, for
// callers that cannot allocate device memory: a fixed region of the code
// object, one copy per device, held until the process ends. Its name is a
// host-side stand-in; host code uses the device address noalloc_scratch gives.
constexpr size_t kNoAllocScratchBytes = size_t(32) << 20;
static __device__ __align__(256) char scratch[kNoAllocScratchBytes];

// Device address of the current device's `scratch`, looked up with
// hipGetSymbolAddress on the first call made with that device current and
// kept per device. Allocates nothing.
constexpr int kNoAllocMaxDevices = 64;

hipError_t noalloc_scratch(char** out, size_t* bytes) {
    static std::mutex lock;
    static char* addr[kNoAllocMaxDevices] = {};
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    if (dev < 0 || dev >= kNoAllocMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> guard(lock);
    if (!addr[dev]) {
        void* p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        addr[dev] = static_cast<char*>(p);
    }
    *out = addr[dev];
    *bytes = kNoAllocScratchBytes;
    return hipSuccess;
}
#endif

hipError_t launch_fixed_frames(real* q, const real* src, const real* invInertia4, int f0, int count,
                               hipStream_t s) {
    if (count <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_apply_fixed_frames, dim3((count + 255) / 256), dim3(256), 0, s, q, src, invInertia4, f0,
                       count);
    return hipGetLastError();
}

}  // namespace RP_PREC
}  // namespace rp

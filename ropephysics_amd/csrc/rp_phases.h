// Substep phases shared by the PLAIN and TILED kernels: predict, finish and
// the rod's self-collision (broadphase, pool allocation, sequential
// projection). Arithmetic order follows the reference CPU substep
// (src/core/solver.cpp:218-267) and contact generation
// (src/core/collision.cpp:10-152); see rp_math.h.
#pragma once

#include "rp_layout.h"
#include "rp_math.h"

namespace rp {
namespace RP_PREC {

__device__ __forceinline__ v3 f4v(real4 a) { return mk(a.x, a.y, a.z); }
__device__ __forceinline__ q4 f4q(real4 a) { return q4{a.x, a.y, a.z, a.w}; }
__device__ __forceinline__ real4 vf4(v3 a, real w) { return make_real4(a.x, a.y, a.z, w); }
__device__ __forceinline__ real4 qf4(q4 a) { return make_real4(a.w, a.x, a.y, a.z); }
// The particles at the ends of a rod's segment k (its stretch constraint k):
// from the rod's segEnd table, or, for a rod whose table is null, k and k + 1
// (constraint k joins particles k and k + 1, the usual chain).
__device__ __forceinline__ int seg_a(const int* segEnd, int k) { return segEnd ? segEnd[2 * k] : k; }
__device__ __forceinline__ int seg_b(const int* segEnd, int k) { return segEnd ? segEnd[2 * k + 1] : k + 1; }
__device__ __forceinline__ v3 ld3(const real* p, int i) { return mk(p[3 * i], p[3 * i + 1], p[3 * i + 2]); }
__device__ __forceinline__ void st3(real* p, int i, v3 a) {
    p[3 * i] = a.x;
    p[3 * i + 1] = a.y;
    p[3 * i + 2] = a.z;
}
__device__ __forceinline__ q4 ld4q(const real* p, int i) {
    return q4{p[4 * i], p[4 * i + 1], p[4 * i + 2], p[4 * i + 3]};
}
__device__ __forceinline__ void st4q(real* p, int i, q4 a) {
    p[4 * i] = a.w;
    p[4 * i + 1] = a.x;
    p[4 * i + 2] = a.y;
    p[4 * i + 3] = a.z;
}

// Symplectic-Euler prediction of one particle. x is updated in place and the
// substep-start position returned through xPrev.
__device__ __forceinline__ void predict_particle(v3& x, v3& v, v3& xPrev, real w, v3 force, v3 kinVel, real h,
                                                 v3 g) {
    xPrev = x;
    if (w == real(0.0)) {
        x = add(x, scale(kinVel, h));
        return;
    }
    v = add(v, scale(add(g, scale(force, w)), h));
    x = add(x, scale(v, h));
}

// Gyroscopic angular-velocity update and quaternion integration of one frame.
__device__ __forceinline__ void predict_frame(q4& q, v3& omega, q4& qPrev, v3 iI, v3 I, v3 torque, real h) {
    qPrev = q;
    if (norm2(iI) == real(0.0)) return;
    const v3 w = omega;
    const v3 tau = rotate_inv(q, torque);
    omega = add(w, scale(cwise(iI, sub(tau, cross(w, cwise(I, w)))), h));
    q = qnormalize(qadd(q, qscale(qmul(q, mkq(real(0.0), omega)), real(0.5) * h)));
}

__device__ __forceinline__ v3 finish_particle(v3 x, v3 xPrev, real w, v3 kinVel, real h, real decay) {
    if (w == real(0.0)) return kinVel;
    return scale(divs(sub(x, xPrev), h), decay);
}

__device__ __forceinline__ v3 finish_frame(q4 q, q4 qPrev, v3 iI, real h, real decay) {
    if (norm2(iI) == real(0.0)) return mk(real(0.0), real(0.0), real(0.0));
    const q4 dq = hemi_identity(qmul(conj(qPrev), q));
    return scale(scale(im(dq), real(2.0) / h), decay);
}

__device__ __forceinline__ prim load_prim(const real* s) {
    prim p;
    p.type = slot_int(s[0]);
    p.a = mk(s[1], s[2], s[3]);
    p.b = mk(s[4], s[5], s[6]);
    p.normal = mk(s[7], s[8], s[9]);
    p.half = mk(s[10], s[11], s[12]);
    p.rot = q4{s[13], s[14], s[15], s[16]};
    p.radius = s[17];
    p.friction = s[18];
    return p;
}
__device__ __forceinline__ void stage_prim(real* s, const Prim& g, real radius, real margin) {
    s[0] = int_slot(g.type);
    for (int c = 0; c < 3; ++c) {
        s[1 + c] = g.a[c];
        s[4 + c] = g.b[c];
        s[7 + c] = g.normal[c];
        s[10 + c] = g.half[c];
    }
    for (int c = 0; c < 4; ++c) s[13 + c] = g.rot[c];
    s[17] = g.radius;
    s[18] = g.friction;
    // Bounding radius about the centre (a, or the capsule's mid-axis): every
    // point of the shape lies within it; -1 for planes.
    real bound = real(-1.0);
    if (g.type == kSphere) {
        bound = g.radius;
    } else if (g.type == kCapsule) {
        const v3 ab = mk(g.b[0] - g.a[0], g.b[1] - g.a[1], g.b[2] - g.a[2]);
        bound = real(0.5) * norm(ab) + g.radius;
    } else if (g.type == kBox) {
        bound = norm(mk(g.half[0], g.half[1], g.half[2]));
    }
    s[19] = bound;
    // The rod's far-test record (prim_far_staged): the bounding centre and
    // the part of the cull radius that does not depend on the particle,
    // formed by prim_far's operations. A shape with no bound has -inf there
    // and its normal in place of the centre.
    v3 c = mk(s[1], s[2], s[3]);
    if (g.type == kCapsule) c = scale(add(c, mk(s[4], s[5], s[6])), real(0.5));
    if (bound < real(0.0)) c = mk(s[7], s[8], s[9]);
    s[20] = c.x;
    s[21] = c.y;
    s[22] = c.z;
    s[23] = bound < real(0.0) ? -r_inf() : real(1.001) * (bound + radius + r_max(margin, real(0.0)));
}

// A particle farther from a convex primitive's bounding sphere than radius +
// margin (with a 1e-3 relative and 1e-6 * coordinate-scale absolute allowance,
// far above the rounding of the signed distance) has no contact with it: the
// signed distance of a convex shape is at least the distance to its centre
// less the bounding radius.
__device__ __forceinline__ bool prim_far(const real* s, v3 x, real radius, real margin) {
    const real bound = s[19];
    if (bound < real(0.0)) return false;
    const int type = slot_int(s[0]);
    v3 c = mk(s[1], s[2], s[3]);
    if (type == kCapsule) c = scale(add(c, mk(s[4], s[5], s[6])), real(0.5));
    const real lim = real(1.001) * (bound + radius + r_max(margin, real(0.0))) +
                     real(1e-6) * (real(1.0) + r_max(abs_max3v(x), abs_max3v(c)));
    return norm2(sub(x, c)) > lim * lim;
}

// prim_far for the particles of the rod whose radius and margin the record
// was staged with, from the record's far-test word f = s[20..23]: prim_far's
// remaining operations on the same operands.
__device__ __forceinline__ bool prim_far_staged(real4 f, v3 x) {
    if (f.w == -r_inf()) return false;
    const v3 c = mk(f.x, f.y, f.z);
    const real lim = f.w + real(1e-6) * (real(1.0) + r_max(abs_max3v(x), abs_max3v(c)));
    return norm2(sub(x, c)) > lim * lim;
}

// World contacts of one particle: its slot for primitive k is two real4
// words, c4[2k * stride] = (normal, offset) and c4[(2k + 1) * stride] =
// (friction, lamN, applied, active), `stride` being the rod's particle count:
// lanes holding consecutive particles touch consecutive 16-byte words.
// Regenerate (or just reset) the particle's slots.
__device__ __forceinline__ void world_contacts_begin(real4* c4, int stride, const real* prims, int nPrim, v3 x,
                                                     real w, real radius, real margin, real fscale, bool regen) {
    for (int k = 0; k < nPrim; ++k) {
        real4* A = c4 + 2 * k * stride;
        real4* B = A + stride;
        if (!regen) {
            real4 b = *B;
            b.y = real(0.0);  // lamN
            b.z = real(0.0);  // applied
            *B = b;
            continue;
        }
        wcontact c;
        if (prim_far(prims + k * kPrimFloats, x, radius, margin))
            c.active = 0;
        else
            world_contact_generate(c, load_prim(prims + k * kPrimFloats), x, w, radius, margin, fscale);
        if (c.active) {
            *A = make_real4(c.n.x, c.n.y, c.n.z, c.offset);
            *B = make_real4(c.friction, real(0.0), real(0.0), int_slot(1));
        } else {
            // An inactive slot's other fields are never read.
            *B = make_real4(real(0.0), real(0.0), real(0.0), int_slot(0));
        }
    }
}

__device__ __forceinline__ void world_contacts_project(real4* c4, int stride, int nPrim, v3& x, v3 xPrev, real w) {
    for (int k = 0; k < nPrim; ++k) {
        const real4* A = c4 + 2 * k * stride;
        real4* B = c4 + (2 * k + 1) * stride;
        real4 b = *B;
        if (!slot_int(b.w)) continue;
        const real4 a = *A;
        wcontact c;
        c.active = 1;
        c.n = mk(a.x, a.y, a.z);
        c.offset = a.w;
        c.friction = b.x;
        c.lamN = b.y;
        c.applied = b.z;
        world_contact_project(c, x, xPrev, w);
        b.y = c.lamN;
        b.z = c.applied;
        *B = b;
    }
}

// The same two passes with the particle's active slots kept by the caller in
// a bit mask (bit k: slot k holds a contact; nPrim <= 32): a regeneration
// returns the new mask and writes only the slots that hold a contact, and the
// reset and the projection touch only those. An inactive slot's words are
// never read, so the arithmetic, the order over k and every result are
// world_contacts_begin's and world_contacts_project's.
static_assert(kMaxPrims <= 32, "one mask bit per primitive");
// world_contact_generate for the staged record sk of a rod's primitive,
// after its far test: c.active = 0 for a primitive that is far. A plane
// reads its type and point in one word, its normal from the far-test word and
// its friction: the fields the plane case of world_contact_generate uses.
__device__ __forceinline__ void world_contact_staged(wcontact& c, const real* sk, v3 x, real w, real radius,
                                                     real margin, real fscale) {
    const real4 f = *reinterpret_cast<const real4*>(sk + 20);
    c.active = 0;
    if (f.w == -r_inf()) {
        const real4 ta = *reinterpret_cast<const real4*>(sk);
        if (slot_int(ta.x) == kPlane) {
            prim pr;
            pr.type = kPlane;
            pr.a = mk(ta.y, ta.z, ta.w);
            pr.normal = mk(f.x, f.y, f.z);
            pr.friction = sk[18];
            world_contact_generate(c, pr, x, w, radius, margin, fscale);
        } else {
            world_contact_generate(c, load_prim(sk), x, w, radius, margin, fscale);
        }
    } else if (!prim_far_staged(f, x)) {
        world_contact_generate(c, load_prim(sk), x, w, radius, margin, fscale);
    }
}

// The world contacts of one particle in a substep of one iteration whose
// contacts are regenerated: each slot would be regenerated from the
// substep-start position xPrev and projected once with lamN = applied = 0,
// so the contact is formed here and projected at once, in the same order
// over k, and nothing is kept.
__device__ __forceinline__ void world_contacts_fused(const real* prims, int nPrim, v3& x, v3 xPrev, real w,
                                                     real radius, real margin, real fscale) {
    for (int k = 0; k < nPrim; ++k) {
        wcontact c;
        world_contact_staged(c, prims + k * kPrimFloats, xPrev, w, radius, margin, fscale);
        if (c.active) {
            c.lamN = real(0.0);
            c.applied = real(0.0);
            world_contact_project(c, x, xPrev, w);
        }
    }
}

__device__ __forceinline__ unsigned world_contacts_begin_masked(real4* c4, int stride, const real* prims, int nPrim,
                                                                v3 x, real w, real radius, real margin, real fscale,
                                                                bool regen, unsigned act) {
    if (!regen) {
        for (unsigned m = act; m != 0u; m &= m - 1u) {
            real4* B = c4 + (2 * __builtin_ctz(m) + 1) * stride;
            real4 b = *B;
            b.y = real(0.0);  // lamN
            b.z = real(0.0);  // applied
            *B = b;
        }
        return act;
    }
    unsigned out = 0u;
    for (int k = 0; k < nPrim; ++k) {
        wcontact c;
        world_contact_staged(c, prims + k * kPrimFloats, x, w, radius, margin, fscale);
        if (c.active) {
            real4* A = c4 + 2 * k * stride;
            A[0] = make_real4(c.n.x, c.n.y, c.n.z, c.offset);
            A[stride] = make_real4(c.friction, real(0.0), real(0.0), int_slot(1));
            out |= 1u << k;
        }
    }
    return out;
}

__device__ __forceinline__ void world_contacts_project_masked(real4* c4, int stride, unsigned act, v3& x, v3 xPrev,
                                                              real w) {
    for (; act != 0u; act &= act - 1u) {
        const int k = __builtin_ctz(act);
        const real4 a = c4[2 * k * stride];
        real4* B = c4 + (2 * k + 1) * stride;
        real4 b = *B;
        wcontact c;
        c.active = 1;
        c.n = mk(a.x, a.y, a.z);
        c.offset = a.w;
        c.friction = b.x;
        c.lamN = b.y;
        c.applied = b.z;
        world_contact_project(c, x, xPrev, w);
        b.y = c.lamN;
        b.z = c.applied;
        *B = b;
    }
}

// ------------------------------------------------------------------ drivers and observers

__device__ inline void path_point(const PathRec& p, double t, double& x, double& y) {
    double fx = p.sx, fy = p.sy, t0 = 0.0;
    for (int k = 0; k < p.nWay; ++k) {
        const double t1 = t0 + p.legs[k];
        if (t < t1) {
            double u = (t - t0) / p.legs[k];
            u = u * u * (3.0 - 2.0 * u);
            x = fx + (p.wx[k] - fx) * u;
            y = fy + (p.wy[k] - fy) * u;
            return;
        }
        fx = p.wx[k];
        fy = p.wy[k];
        t0 = t1;
    }
    x = fx;
    y = fy;
}

// Velocity of the driven particle during the rod's step `step`: constant over
// each control tick, (P(t_{k+1}) - P(t_k)) / period, rounded to the working
// precision.
__device__ inline v3 path_velocity(const PathRec& p, int step) {
    const int tick = step / p.stepsPerControl;
    double x1, y1, x0, y0;
    path_point(p, double(tick + 1) * p.period, x1, y1);
    path_point(p, double(tick) * p.period, x0, y0);
    const double inv = 1.0 / p.period;
    return mk(real((x1 - x0) * inv), real((y1 - y0) * inv), real((p.height - p.height) * inv));
}

// ------------------------------------------------------------------ self-collision
//
// One rod's working set in LDS. Segments are visited in the rod's original
// stretch-constraint order, as the CPU broadphase does.
struct SelfLds {
    real4* cen;     // nS centres and lengths (after a lane-group build: in bucket order)
    int* cellOf;     // nS
    int* sorted;     // nS
    int* segCount;   // nS
    int* segStart;   // nS
    int* cellStart;  // table + 1
    real* pool;     // cap * kSelfFloats
    int* misc;       // [0] live, [1] overflow, [2] long segments, [3] colours in use (coloured order),
                     // [4..7] the first four long segments (nS: none)
    rbits* rmisc;    // [0] cell size, [1] longest segment, [2] largest |coordinate| of a centre, [3] shortest
                     // segment, [4] longest segment that is not long, as bits (non-negative: ordered as
                     // integers); [5] unused; [6] the largest |coordinate| of a centre at the last full
                     // regeneration; [7] 1 while the candidate lists are valid (any lane of the rod
                     // clears it); [8] the length above which a segment is long at the next regeneration
    int* scan;       // kScanWaves, scratch of the group-parallel scans
};

// Wavefronts a lane group may span (512 lanes).
constexpr int kScanWaves = 8;

// misc (8 ints) then rmisc (5 reals), padded so the int arrays after them stay
// 16-byte aligned.
constexpr int kMiscBytes = 112;
constexpr int kLongSlots = 4;

__host__ __device__ inline int self_lds_bytes(int nS, int table, int cap) {
    return nS * int(sizeof(real4)) + 4 * nS * 4 + (table + 1) * 4 + cap * kSelfFloats * int(sizeof(real)) +
           kMiscBytes + kScanWaves * 4;
}

__device__ inline SelfLds self_carve(char* base, int nS, int table, int cap) {
    SelfLds s;
    s.cen = reinterpret_cast<real4*>(base);
    s.pool = reinterpret_cast<real*>(base + nS * int(sizeof(real4)));
    s.misc = reinterpret_cast<int*>(base + nS * int(sizeof(real4)) + cap * kSelfFloats * int(sizeof(real)));
    s.rmisc = reinterpret_cast<rbits*>(s.misc + 8);
    s.cellOf = s.misc + kMiscBytes / 4;
    s.sorted = s.cellOf + nS;
    s.segCount = s.sorted + nS;
    s.segStart = s.segCount + nS;
    s.cellStart = s.segStart + nS;
    s.scan = s.cellStart + table + 1;
    return s;
}

// Centres, cell size and the counting sort into buckets; one lane.
__device__ inline void self_build_hash(const SelfLds& s, const real4* x4, const int* segEnd, int nS, int table,
                                       real radius) {
    real maxSeg = real(0.0);
    for (int j = 0; j < nS; ++j) {
        const v3 a = f4v(x4[seg_a(segEnd, j)]), b = f4v(x4[seg_b(segEnd, j)]);
        s.cen[j] = vf4(scale(add(a, b), real(0.5)), real(0.0));
        const real len = norm(sub(b, a));
        maxSeg = maxSeg < len ? len : maxSeg;
    }
    const real cell = maxSeg + real(2.0) * radius;
    s.rmisc[0] = real_bits(cell);
    for (int t = 0; t <= table; ++t) s.cellStart[t] = 0;
    for (int j = 0; j < nS; ++j) {
        const v3 c = f4v(s.cen[j]);
        const int h = int(cell_hash(cell_coord(c.x, cell), cell_coord(c.y, cell), cell_coord(c.z, cell), table));
        s.cellOf[j] = h;
        ++s.cellStart[h + 1];
    }
    for (int t = 0; t < table; ++t) s.cellStart[t + 1] += s.cellStart[t];
    for (int j = 0; j < nS; ++j) s.sorted[s.cellStart[s.cellOf[j]]++] = j;
    for (int t = table; t > 0; --t) s.cellStart[t] = s.cellStart[t - 1];
    s.cellStart[0] = 0;
}

// Narrow phase of the candidate pair (segment j = particles j0, j1 at a0, a1;
// segment k = particles k0, k1): a contact when the closest points are nearer
// than 2r. Each contact takes the next slot from slotBase; with `fill`, a slot
// below the capacity receives the contact record.
__device__ __forceinline__ void self_pair(const SelfLds& s, const real4* x4, int j0, int j1, v3 a0, v3 a1, int k0,
                                          int k1, real radius, real friction, bool fill, int slotBase, int cap,
                                          int& count) {
    const v3 b0 = f4v(x4[k0]), b1 = f4v(x4[k1]);
    real u, v;
    seg_seg(a0, a1, b0, b1, u, v);
    const v3 pa = add(a0, scale(sub(a1, a0), u));
    const v3 pb = add(b0, scale(sub(b1, b0), v));
    const v3 delta = sub(pa, pb);
    const real dist = norm(delta);
    const real sep = dist - real(2.0) * radius;
    if (sep >= real(0.0)) return;
    const int slot = slotBase + count++;
    if (!fill || slot >= cap) return;
    const v3 n = dist > real(1e-12) ? divs(delta, dist) : mk(real(0.0), real(0.0), real(1.0));
    const int idx[4] = {j0, j1, k0, k1};
    const real wt[4] = {real(1.0) - u, u, -(real(1.0) - v), -v};
    real dotSum = real(0.0);
    for (int m = 0; m < 4; ++m) dotSum += wt[m] * dot(f4v(x4[idx[m]]), n);
    real* c = s.pool + slot * kSelfFloats;
    for (int m = 0; m < 4; ++m) {
        c[m] = int_slot(idx[m]);
        c[4 + m] = wt[m];
    }
    c[8] = n.x;
    c[9] = n.y;
    c[10] = n.z;
    c[11] = dotSum - sep;  // offset
    c[12] = friction;
    c[13] = real(0.0);  // lamN
    c[14] = real(0.0);  // applied
}

// Contacts of segment j with later segments in the CPU's candidate order
// (27 cells dz, dy, dx; bucket items ascending; a bucket reached twice is
// visited twice). Counts them, or with `fill` writes them at j's pool offset.
__device__ inline void self_query(const SelfLds& s, const real4* x4, const int* segEnd, int nS, int table,
                                  int gap, int cap, real radius, real friction, int j, bool fill) {
    int count = 0;
    if (nS > gap) {
        const real cell = bits_real(s.rmisc[0]);
        const v3 cj = f4v(s.cen[j]);
        const int cx = cell_coord(cj.x, cell), cy = cell_coord(cj.y, cell), cz = cell_coord(cj.z, cell);
        const int j0 = seg_a(segEnd, j), j1 = seg_b(segEnd, j);
        const v3 a0 = f4v(x4[j0]), a1 = f4v(x4[j1]);
        const int slotBase = fill ? s.segStart[j] : 0;
        for (int dz = -1; dz <= 1; ++dz)
            for (int dy = -1; dy <= 1; ++dy)
                for (int dx = -1; dx <= 1; ++dx) {
                    const int h = int(cell_hash(cx + dx, cy + dy, cz + dz, table));
                    const int e1 = s.cellStart[h + 1];
                    for (int e = s.cellStart[h]; e < e1; ++e) {
                        const int k = s.sorted[e];
                        if (k - j < gap) continue;
                        self_pair(s, x4, j0, j1, a0, a1, seg_a(segEnd, k), seg_b(segEnd, k), radius, friction, fill,
                                  slotBase, cap, count);
                    }
                }
    }
    if (!fill) s.segCount[j] = count;
}

// ------------------------------------------------------------------ group-parallel broadphase
//
// The contact list of self_build_hash / self_query / self_allocate, built by
// the G lanes of a rod's lane group. Counting needs no hash tables: each lane
// scans its segment's later segments in index order and weights every contact
// by the number of times the bucketed query would reach it. The tables are
// built only where some rod of the workgroup has contacts, for the fill pass
// that writes them in the CPU's order. Every step is exact integer work, an
// order-free max, or a cull that can drop only non-contacts, so the list equals
// the serial build's. Every lane of the workgroup calls these (they contain
// barriers); `on` is false for lanes whose rod has no self-collision.

// Exclusive prefix sum of in[0, n) into out[0, n) by the G lanes of a rod's
// lane group, each summing a chunk of ceil(n / G) entries: a scan of the chunk
// sums within each wavefront (lane - off is in the same wavefront when its
// position in the wavefront is at least off, and in the same group -- a
// group's lanes are consecutive -- when l >= off), then across the wavefronts
// the group spans through `scratch`, one entry per wavefront. Returns the
// total. Two barriers, whatever G, n or `on`.
__device__ inline int group_scan(const int* in, int* out, int n, int* scratch, int l, int G, bool on) {
    const int chunk = on ? (n + G - 1) / G : 0;
    const int lo = on ? min(n, l * chunk) : 0;
    const int hi = on ? min(n, lo + chunk) : 0;
    int sum = 0;
    for (int i = lo; i < hi; ++i) sum += in[i];
    int x = sum;
    const int lane = int(threadIdx.x);
    for (int off = 1; off < 64; off <<= 1) {
        const int v = __shfl_up(x, off, 64);
        if (l >= off && (lane & 63) >= off) x += v;
    }
    const int g0 = lane - l;
    const int w = (lane >> 6) - (g0 >> 6);
    if (on && ((lane & 63) == 63 || l == G - 1)) scratch[w] = x;
    __syncthreads();
    int before = 0, total = 0;
    if (on) {
        const int wLast = ((g0 + G - 1) >> 6) - (g0 >> 6);
        for (int i = 0; i <= wLast; ++i) {
            const int t = scratch[i];
            before += i < w ? t : 0;
            total += t;
        }
    }
    int run = before + x - sum;
    for (int i = lo; i < hi; ++i) {
        const int v = in[i];
        out[i] = run;
        run += v;
    }
    __syncthreads();
    return total;
}

__device__ __forceinline__ real abs_max3(v3 a) { return r_max(r_abs(a.x), r_max(r_abs(a.y), r_abs(a.z))); }
__device__ __forceinline__ int count_eq4(int4 c, int h) {
    return (c.x == h ? 1 : 0) + (c.y == h ? 1 : 0) + (c.z == h ? 1 : 0) + (c.w == h ? 1 : 0);
}

// Bucket of cell (cx + dx, cy + dy, cz + dz) for dx, dy, dz in -1..1, with the
// row term hzy = (cz + dz) p3 ^ (cy + dy) p2 given: cell_hash's arithmetic.
__device__ __forceinline__ int row_bucket(int cx, int dx, unsigned hzy, unsigned tbl, bool pow2) {
    const unsigned hh = (static_cast<unsigned>(cx + dx) * 73856093u) ^ hzy;
    return static_cast<int>(pow2 ? (hh & (tbl - 1u)) : hh % tbl);
}

// The fill pass of self_query for the lane group: segment j's contacts, in
// the CPU's candidate order (27 cells dz, dy, dx; bucket items ascending; a
// bucket reached twice is visited twice), written from slotBase. The three
// buckets of a (dz, dy) row are fetched together and their candidates tested
// four at a time; a candidate whose centre is farther from j's than the two
// half-lengths plus 2r (with a relative margin of 1e-3 and an absolute one of
// 1e-6 times the coordinate scale, far above the rounding of centres and
// lengths) cannot be a contact and skips the closest-point test. Every point of
// a segment lies within half its length of its centre, so the list is
// unchanged.
__device__ inline void self_fill_rows(const SelfLds& s, const real4* x4, const int* segEnd, int nS, int table,
                                      int gap, int cap, real radius, real friction, real cell, int j, int j0, int j1,
                                      real4 cj4, int slotBase) {
    const v3 cj = f4v(cj4);
    const real reachJ = real(0.5) * cj4.w + real(2.0) * radius;
    const real scaleJ = abs_max3(cj);
    const int cx = cell_coord(cj.x, cell), cy = cell_coord(cj.y, cell), cz = cell_coord(cj.z, cell);
    const v3 a0 = f4v(x4[j0]), a1 = f4v(x4[j1]);
    const unsigned tbl = static_cast<unsigned>(table);
    // Usually every lane group of the wave has a power-of-two table: then the
    // modulo path is not issued at all.
    const bool pow2 = __all((tbl & (tbl - 1u)) == 0u) != 0;
    int count = 0;
    for (int dz = -1; dz <= 1; ++dz)
        for (int dy = -1; dy <= 1; ++dy) {
            const unsigned hzy =
                (static_cast<unsigned>(cz + dz) * 83492791u) ^ (static_cast<unsigned>(cy + dy) * 19349663u);
            int lo[3], hi[3];
#pragma unroll
            for (int dx = 0; dx < 3; ++dx) {
                const int h = row_bucket(cx, dx - 1, hzy, tbl, pow2);
                lo[dx] = s.cellStart[h];
                hi[dx] = s.cellStart[h + 1];
            }
            const int n0 = hi[0] - lo[0], n01 = n0 + hi[1] - lo[1], total = n01 + hi[2] - lo[2];
            for (int base = 0; base < total; base += 4) {
                int kq[4];
                real4 cq[4];
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    const int q = min(base + u, total - 1);
                    const int e = q < n0 ? lo[0] + q : (q < n01 ? lo[1] + (q - n0) : lo[2] + (q - n01));
                    // Bucket items are segment indices below nS; the clamp keeps
                    // every load in range whatever the tables hold.
                    kq[u] = min(max(s.sorted[e], 0), nS - 1);
                }
#pragma unroll
                for (int u = 0; u < 4; ++u) cq[u] = s.cen[kq[u]];
                unsigned pass = 0u;
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    const v3 ck = f4v(cq[u]);
                    const real lim = real(1.001) * (reachJ + real(0.5) * cq[u].w) +
                                     real(1e-6) * (real(1.0) + r_max(scaleJ, abs_max3(ck)));
                    const bool keep = (base + u < total) & (kq[u] - j >= gap) & !(norm2(sub(ck, cj)) > lim * lim);
                    pass |= keep ? (1u << u) : 0u;
                }
                while (pass != 0u) {
                    const int u = __builtin_ctz(pass);
                    pass &= pass - 1u;
                    const int k = u == 0 ? kq[0] : u == 1 ? kq[1] : u == 2 ? kq[2] : kq[3];
                    self_pair(s, x4, j0, j1, a0, a1, seg_a(segEnd, k), seg_b(segEnd, k), radius, friction, true,
                              slotBase, cap, count);
                }
            }
        }
}

// The fill pass for a segment whose contacts are known: kk[0, nk) are its
// later segments in contact, ascending, each with a nonzero count. In the
// bucketed query's order (27 cells dz, dy, dx; within a bucket, ascending
// segment index), the list is, cell by cell, the known contacts whose bucket
// is that cell's, ascending -- a contact met at several cells of one bucket is
// written once per cell. Neither the bucket tables nor the other segments are
// read.
template <int M>
__device__ inline void self_fill_listed(const SelfLds& s, const real4* x4, const int* segEnd, int table, int cap,
                                        real radius, real friction, real cell, int j0, int j1, real4 cj4,
                                        int slotBase, const int (&kk)[M], int nk) {
    const v3 cj = f4v(cj4);
    const int cx = cell_coord(cj.x, cell), cy = cell_coord(cj.y, cell), cz = cell_coord(cj.z, cell);
    const v3 a0 = f4v(x4[j0]), a1 = f4v(x4[j1]);
    const unsigned tbl = static_cast<unsigned>(table);
    const bool pow2 = (tbl & (tbl - 1u)) == 0u;
    int hk[M];
#pragma unroll
    for (int e = 0; e < M; ++e) {
        hk[e] = -1;
        if (e < nk) {
            const v3 ck = f4v(s.cen[kk[e]]);
            hk[e] = int(cell_hash(cell_coord(ck.x, cell), cell_coord(ck.y, cell), cell_coord(ck.z, cell), table));
        }
    }
    int count = 0;
    for (int dz = -1; dz <= 1; ++dz)
        for (int dy = -1; dy <= 1; ++dy) {
            const unsigned hzy =
                (static_cast<unsigned>(cz + dz) * 83492791u) ^ (static_cast<unsigned>(cy + dy) * 19349663u);
            for (int dx = -1; dx <= 1; ++dx) {
                const int h = row_bucket(cx, dx, hzy, tbl, pow2);
#pragma unroll
                for (int e = 0; e < M; ++e)
                    if (hk[e] == h)
                        self_pair(s, x4, j0, j1, a0, a1, seg_a(segEnd, kk[e]), seg_b(segEnd, kk[e]), radius, friction,
                                  true, slotBase, cap, count);
            }
        }
}

// One later segment k whose centre passed the cull: segment j's contacts with
// it, each counted once for every one of j's 27 cells whose bucket is k's (the
// times the bucketed query reaches k). Before the closest-point test, the pair
// is dropped when segment j is farther from k's centre than k's half-length
// plus 2r (with the cull's margins): every point of segment k lies within its
// half-length of its centre. The first M contacts reached (nonzero count, in
// the order the caller visits k) are kept in kk[]; nk counts them all.
template <int M>
__device__ __forceinline__ void self_count_one(const SelfLds& s, const real4* x4, const int* segEnd, int table,
                                               real radius, real cell, real eps, v3 cj, int j0, int j1, int k,
                                               real4 ck4, int (&kk)[M > 0 ? M : 1], int& nk, int& count) {
    const v3 ck = f4v(ck4);
    const v3 a0 = f4v(x4[j0]), a1 = f4v(x4[j1]);
    const v3 dj = sub(a1, a0);
    const real a = norm2(dj);
    // Distance from segment j to k's centre (the nearest point of j up to
    // rounding, far inside the margin).
    const real tq = a > real(1e-30) ? r_max(real(0.0), min_ab(real(1.0), dot(sub(ck, a0), dj) / a)) : real(0.0);
    const v3 off = sub(add(a0, scale(dj, tq)), ck);
    const real limP = real(1.001) * (real(0.5) * ck4.w + real(2.0) * radius) + eps;
    if (norm2(off) > limP * limP) return;
    int c = 0;
    self_pair(s, x4, j0, j1, a0, a1, seg_a(segEnd, k), seg_b(segEnd, k), radius, real(0.0), false, 0, 0, c);
    if (c == 0) return;
    const unsigned tbl = static_cast<unsigned>(table);
    const bool pow2 = (tbl & (tbl - 1u)) == 0u;
    const int cx = cell_coord(cj.x, cell), cy = cell_coord(cj.y, cell), cz = cell_coord(cj.z, cell);
    const int hk = int(cell_hash(cell_coord(ck.x, cell), cell_coord(ck.y, cell), cell_coord(ck.z, cell), table));
    int mult = 0;
    for (int dz = -1; dz <= 1; ++dz)
        for (int dy = -1; dy <= 1; ++dy) {
            const unsigned hzy =
                (static_cast<unsigned>(cz + dz) * 83492791u) ^ (static_cast<unsigned>(cy + dy) * 19349663u);
#pragma unroll
            for (int dx = -1; dx <= 1; ++dx) mult += row_bucket(cx, dx, hzy, tbl, pow2) == hk ? 1 : 0;
        }
    if (M > 0 && mult > 0) {
#pragma unroll
        for (int m = 0; m < M; ++m)
            if (nk == m) kk[m] = k;
        ++nk;
    }
    count += mult;
}

// Candidate lists: the later segments k >= j + gap whose centres lie within
// the cull radius plus a skin of segment j's, found by a scan and kept while
// no particle can have moved far enough for another segment to enter the cull
// radius. One 64-bit word per segment: bits 0-2 the number of candidates (0
// to kCandidates; kCandidateOverflow when there were more), then k - j in 10
// bits each, ascending.
constexpr int kCandidates = 6;
constexpr unsigned long long kCandidateOverflow = 7ull;

// Number of contacts segment j has in the list: its later segments k >= j +
// gap scanned in index order, each passing the centre cull counted by
// self_count_one. The cull drops a pair whose centres are farther apart than
// the two half-lengths plus 2r (with a relative margin of 1e-3 and an absolute
// one of 1e-6 times the coordinate scale, far above the rounding of centres and
// lengths); it keeps every contact. When consecutive segments share a particle
// (`chain`), consecutive centres are at most the longer of the two segments
// apart, so a segment k at distance D also rules out the following segments k'
// while D stays above the cull radius at their bound: floor((D - lim) / step)
// of them, lim being the cull radius at the bound lenCap and step lenCap with
// a 1e-3 margin. No segment of such a stretch may be longer than lenCap: the
// long segments are listed in misc[4..7] (when `longs`: the rod has some and
// all are listed), the first one at or after k ends the stretch, and from a
// long segment the scan steps by one. Far parts of the rod cost one test per
// stretch of segments.
//
// With `cand`, the scan also collects segment j's candidates, those within the
// cull radius plus `skin` (the skips then use that radius), into *cand.
template <int M>
__device__ inline int self_count_scan(const SelfLds& s, const real4* x4, const int* segEnd, bool longs, int nS,
                                      int table, int gap, real radius, real cell, real lenCap, real scaleMax,
                                      bool chain, int j, int j0, int j1, real4 cj4, int (&kk)[M > 0 ? M : 1],
                                      int& nk, real skin, unsigned long long* cand) {
    const v3 cj = f4v(cj4);
    const real reachJ = real(0.5) * cj4.w + real(2.0) * radius;
    const real eps = real(1e-6) * (real(1.0) + scaleMax);
    const float limCap = float(real(1.001) * (reachJ + real(0.5) * lenCap) + eps + skin);
    const float invCap = chain && lenCap > real(0.0) ? 0.999f / (1.001f * float(lenCap)) : 0.0f;
    unsigned long long word = 0;
    int count = 0;
    int k = j + gap;
    while (k < nS) {
        const real4 ck4 = s.cen[k];
        const v3 ck = f4v(ck4);
        const real D2 = norm2(sub(ck, cj));
        const real limK = real(1.001) * (reachJ + real(0.5) * ck4.w) + eps;
        const real limS = limK + skin;
        if (D2 > limS * limS) {
            int lf = nS;
            if (longs) {
                const int4 L = *reinterpret_cast<const int4*>(s.misc + 4);
                lf = L.x >= k ? min(lf, L.x) : lf;
                lf = L.y >= k ? min(lf, L.y) : lf;
                lf = L.z >= k ? min(lf, L.z) : lf;
                lf = L.w >= k ? min(lf, L.w) : lf;
            }
            const float sk = (__builtin_amdgcn_sqrtf(float(D2)) - limCap) * invCap;
            const int next = k + 1 + (sk > 0.0f ? int(fminf(sk, float(nS))) : 0);
            k = lf > k ? min(next, lf) : k + 1;
            continue;
        }
        if (cand) {
            const unsigned long long n = word & 7ull;
            word = n < kCandidates && k - j < 1024
                       ? ((word + 1ull) | (static_cast<unsigned long long>(k - j) << (3 + 10 * n)))
                       : ((word & ~7ull) | kCandidateOverflow);
        }
        if (!(D2 > limK * limK))
            self_count_one<M>(s, x4, segEnd, table, radius, cell, eps, cj, j0, j1, k, ck4, kk, nk, count);
        ++k;
    }
    if (cand) *cand = word;
    return count;
}

// The count of self_count_scan from segment j's candidate list (not in
// overflow): the same centre cull and self_count_one on each candidate, in
// ascending order. Every segment outside the list is beyond the cull radius
// (the caller keeps the lists only while that holds), so the count and kk[]
// are the scan's.
template <int M>
__device__ inline int self_count_listed(const SelfLds& s, const real4* x4, const int* segEnd, int nS, int table,
                                        real radius, real cell, real scaleMax, int j, int j0, int j1, real4 cj4,
                                        int (&kk)[M > 0 ? M : 1], int& nk, unsigned long long word) {
    const v3 cj = f4v(cj4);
    const real reachJ = real(0.5) * cj4.w + real(2.0) * radius;
    const real eps = real(1e-6) * (real(1.0) + scaleMax);
    const int n = int(word & 7ull);
    int count = 0;
    for (int m = 0; m < n; ++m) {
        // In range whatever the word holds.
        const int k = min(j + int((word >> (3 + 10 * m)) & 1023ull), nS - 1);
        const real4 ck4 = s.cen[k];
        const real D2 = norm2(sub(f4v(ck4), cj));
        const real limK = real(1.001) * (reachJ + real(0.5) * ck4.w) + eps;
        if (!(D2 > limK * limK))
            self_count_one<M>(s, x4, segEnd, table, radius, cell, eps, cj, j0, j1, k, ck4, kk, nk, count);
    }
    return count;
}

// Broadphase and contact generation by the G lanes of a rod's lane group; lane
// l owns segments j = l + G t, t < TP (the host guarantees nS <= TP G), and
// keeps their centres in registers; s.cen holds them in segment order.
// Inclusive max (kMin: min) of v over the lanes of this lane's group in its
// wavefront, in wavefront order; the group's last lane in the wavefront ends
// with the value of the whole stretch.
//
// The scan runs on DPP lane moves, in registers: four row shifts (1, 2, 4, 8
// lanes) within each row of 16 lanes, then the last lane of row 0 and 2 into
// rows 1 and 3, then lane 31 into rows 2 and 3. A lane whose source is outside
// its row, or in a row the move does not write, receives its own value. A move
// is taken only when its source lane belongs to the same group.
template <int kCtrl, int kRowMask>
__device__ __forceinline__ rbits dpp_move(rbits v) {
    if constexpr (sizeof(rbits) == 4) {
        return __builtin_amdgcn_update_dpp(v, v, kCtrl, kRowMask, 0xF, false);
    } else {
        const int lo = int(static_cast<unsigned long long>(v) & 0xffffffffull);
        const int hi = int(static_cast<unsigned long long>(v) >> 32);
        const unsigned lo2 = unsigned(__builtin_amdgcn_update_dpp(lo, lo, kCtrl, kRowMask, 0xF, false));
        const unsigned hi2 = unsigned(__builtin_amdgcn_update_dpp(hi, hi, kCtrl, kRowMask, 0xF, false));
        return rbits((static_cast<unsigned long long>(hi2) << 32) | lo2);
    }
}

template <bool kMin>
__device__ __forceinline__ rbits group_wave_extreme(rbits v, int l) {
    const int lane = int(threadIdx.x) & 63;
    const int r = lane & 15;
    const auto pick = [](rbits a, rbits b) { return kMin ? (b < a ? b : a) : (b > a ? b : a); };
    rbits u = dpp_move<0x111, 0xF>(v);  // row_shr:1
    if (l >= 1) v = pick(v, u);
    u = dpp_move<0x112, 0xF>(v);  // row_shr:2
    if (l >= 2) v = pick(v, u);
    u = dpp_move<0x114, 0xF>(v);  // row_shr:4
    if (l >= 4) v = pick(v, u);
    u = dpp_move<0x118, 0xF>(v);  // row_shr:8
    if (l >= 8) v = pick(v, u);
    u = dpp_move<0x142, 0xA>(v);  // row_bcast:15 into rows 1 and 3: the row before's last lane
    if (l > r) v = pick(v, u);
    u = dpp_move<0x143, 0xC>(v);  // row_bcast:31 into rows 2 and 3: lane 31
    if (l > lane - 32) v = pick(v, u);
    return v;
}

// Broadphase and contact generation by the G lanes of a rod's lane group; lane
// l owns segments j = l + G t, t < TP (the host guarantees nS <= TP G), and
// keeps their centres in registers; s.cen holds them in segment order.
// Per-regeneration scalars (longest and shortest segment, coordinate scale) set
// for the next reduction, and the long-segment list emptied; lane 0, when no
// lane reads them.
__device__ __forceinline__ void self_regen_reset(const SelfLds& s) {
    s.rmisc[1] = 0;
    s.rmisc[2] = 0;
    s.rmisc[3] = real_bits(real(1e30));
}
// At launch start: no candidate lists yet, and no segment is long.
__device__ __forceinline__ void self_candidates_reset(const SelfLds& s) {
    s.rmisc[6] = 0;
    s.rmisc[7] = 0;
    s.rmisc[8] = real_bits(real(1e30));
}
__device__ __forceinline__ void self_list_reset(const SelfLds& s, int nS) {
    s.misc[2] = 0;
    for (int i = 4; i < 8; ++i) s.misc[i] = nS;
    s.rmisc[4] = 0;
}

// The bucket tables of the rods with contacts (`build`): cellOf, cellStart
// (bucket offsets) and sorted (segments by bucket, ascending within one).
template <int TP>
__device__ inline void self_build_tables(const SelfLds& s, int nS, int table, real cell, bool build,
                                         const real4 (&cj)[TP], int l, int G) {
    if (build) {
        for (int t = l; t <= table; t += G) s.cellStart[t] = 0;
#pragma unroll
        for (int t = 0; t < TP; ++t) {
            const int j = l + G * t;
            if (j < nS) {
                const v3 c = f4v(cj[t]);
                s.cellOf[j] =
                    int(cell_hash(cell_coord(c.x, cell), cell_coord(c.y, cell), cell_coord(c.z, cell), table));
            }
        }
    }
    __syncthreads();
#pragma unroll
    for (int t = 0; t < TP; ++t) {
        const int j = l + G * t;
        if (build && j < nS) atomicAdd(&s.cellStart[s.cellOf[j]], 1);
    }
    __syncthreads();
    const int items = group_scan(s.cellStart, s.cellStart, build ? table : 0, s.scan, l, G, build);
    if (build && l == 0) s.cellStart[table] = items;
    // Stable placement: j goes after the earlier segments of its bucket.
    // cellOf is 16-byte aligned, so four buckets per LDS read.
    const int4* cellOf4 = reinterpret_cast<const int4*>(s.cellOf);
#pragma unroll
    for (int t = 0; t < TP; ++t) {
        const int j = l + G * t;
        if (build && j < nS) {
            const int h = s.cellOf[j];
            int rank = 0, i = 0;
            for (; i + 16 <= j; i += 16) {
                const int4 c0 = cellOf4[(i >> 2) + 0], c1 = cellOf4[(i >> 2) + 1];
                const int4 c2 = cellOf4[(i >> 2) + 2], c3 = cellOf4[(i >> 2) + 3];
                rank += count_eq4(c0, h) + count_eq4(c1, h) + count_eq4(c2, h) + count_eq4(c3, h);
            }
            for (; i + 4 <= j; i += 4) rank += count_eq4(cellOf4[i >> 2], h);
            for (; i < j; ++i) rank += s.cellOf[i] == h ? 1 : 0;
            s.sorted[s.cellStart[h] + rank] = j;
        }
    }
    __syncthreads();
}

// The skin of the candidate lists, in rod radii.
constexpr float kSkinRadii = 1.0f;

// Returns whether some rod of the workgroup has contacts (the same on every
// lane); `rebuilt` tells the rod's lanes whether its candidate lists were
// rebuilt. x4 holds the substep-start positions.
//
// Candidate lists. The count pass scans a segment's later segments only when
// the rod's lists are rebuilt; otherwise it tests the segment's candidates,
// the segments whose centres were within the cull radius plus a skin of r
// (kSkinRadii radii) at the last rebuild. Each lane sums, over the substeps
// since then, a bound on how far its particles moved (self_lists_stale), so
// every particle moved at most D, the largest of these sums over the rod's
// lanes; every centre and every half-length then moved or changed by at most
// D, and a pair's centre distance fell and its cull radius grew by at most 2 D
// and 2.002 D (the cull radius carries the 1.001 margin). A segment outside
// the list was farther than its cull radius plus the skin, so while 4.01 D
// stays below the skin (less an absolute 1e-6 of the coordinate scale for
// rounding; 4.01 over 4.002 covers the rounding of the sums) it is still
// beyond the cull radius and the count pass would have dropped it: the lists
// give the scan's counts and contacts exactly. A
// lane whose sum passes that bound clears rmisc[7]; the lists are rebuilt at
// a launch's first regeneration, at the first regeneration after a lane
// cleared the flag, and after the bucket tables (whose arrays hold them) were
// built; a segment with more than kCandidates candidates is scanned at every
// regeneration.
template <int TP>
__device__ inline bool self_regen_group(const SelfLds& s, const real4* x4, const int* segEnd, int nS, int table,
                                        int gap, int cap, real radius, real friction, bool chain, bool on, int l,
                                        int G, bool& rebuilt) {
    real4 cj[TP];
    // Whether the rod's candidate lists are rebuilt now (the same on every
    // lane of the rod). Every lane reads the flag before lane 0 rewrites it
    // after the vote below; the lanes that cleared it did so before the
    // barrier that precedes this call.
    const bool rebuild = on && s.rmisc[7] == 0;
    // The longest and shortest segment and the largest coordinate: per lane,
    // per wavefront stretch of the group, then one atomic per stretch (on
    // values reset when the previous regeneration read them).
    rbits lenHi = 0, lenLo = real_bits(real(1e30)), scaleHi = 0;
    // Segments longer than twice the shortest of the previous regeneration
    // are "long"; the first few are listed, and skips stop at them. Any
    // threshold gives the same counts: it only decides where the scan's skips
    // stop.
    const real lenLong = on ? bits_real(s.rmisc[8]) : real(0.0);
    rbits shortHi = 0;
#pragma unroll
    for (int t = 0; t < TP; ++t) {
        const int j = l + G * t;
        cj[t] = make_real4(real(0.), real(0.), real(0.), real(0.));
        if (on && j < nS) {
            const v3 a = f4v(x4[seg_a(segEnd, j)]), b = f4v(x4[seg_b(segEnd, j)]);
            const real len = norm(sub(b, a));
            const v3 c = scale(add(a, b), real(0.5));
            cj[t] = vf4(c, len);
            s.cen[j] = cj[t];
            const rbits lb = real_bits(len), sb = real_bits(abs_max3(c));
            lenHi = lb > lenHi ? lb : lenHi;
            lenLo = lb < lenLo ? lb : lenLo;
            scaleHi = sb > scaleHi ? sb : scaleHi;
            if (len > lenLong) {
                const int slot = atomicAdd(&s.misc[2], 1);
                if (slot < kLongSlots) s.misc[4 + slot] = j;
            } else {
                shortHi = lb > shortHi ? lb : shortHi;
            }
        }
    }
    lenHi = group_wave_extreme<false>(lenHi, l);
    scaleHi = group_wave_extreme<false>(scaleHi, l);
    lenLo = group_wave_extreme<true>(lenLo, l);
    shortHi = group_wave_extreme<false>(shortHi, l);
    if (on && ((threadIdx.x & 63) == 63 || l == G - 1)) {
        atomicMax(&s.rmisc[1], lenHi);
        atomicMax(&s.rmisc[2], scaleHi);
        atomicMin(&s.rmisc[3], lenLo);
        atomicMax(&s.rmisc[4], shortHi);
    }
    __syncthreads();
    const real lenShort = on ? bits_real(s.rmisc[3]) : real(0.0);
    const real lenMax = on ? bits_real(s.rmisc[1]) : real(0.0);
    const real scaleMax = on ? bits_real(s.rmisc[2]) : real(0.0);
    const real skin = real(kSkinRadii) * radius;
    const real cell = lenMax + real(2.0) * radius;
    // With more long segments than slots, the longest segment bounds every
    // stretch and none ends one. Otherwise a stretch holds no long segment, and
    // the longest of the others bounds it.
    const int numLong = on ? s.misc[2] : 0;
    const bool listed = numLong <= kLongSlots && lenLong > real(0.0);
    const real lenCap = listed ? (on ? bits_real(s.rmisc[4]) : real(0.0)) : lenMax;
    // Contacts kept per segment for the fill: a few registers where a lane
    // owns one or two segments, none beyond.
    constexpr int kM = TP == 1 ? 6 : (TP == 2 ? 3 : 0);
    constexpr int kMr = kM > 0 ? kM : 1;
    int kk[TP][kMr];
    int nk[TP];
    bool counted = false;
#pragma unroll
    for (int t = 0; t < TP; ++t) {
        const int j = l + G * t;
        nk[t] = 0;
        if (on && j < nS) {
            int c = 0;
            if (nS > gap) {
                // Each lane reads and writes only its own segments' words.
                const bool longs = listed && numLong > 0;
                if (rebuild) {
                    unsigned long long w = 0;
                    c = self_count_scan<kM>(s, x4, segEnd, longs, nS, table, gap, radius, cell, lenCap, scaleMax,
                                            chain, j, seg_a(segEnd, j), seg_b(segEnd, j), cj[t], kk[t], nk[t], skin,
                                            &w);
                    s.cellOf[j] = int(unsigned(w));
                    s.sorted[j] = int(unsigned(w >> 32));
                } else {
                    const unsigned long long w =
                        unsigned(s.cellOf[j]) | (static_cast<unsigned long long>(unsigned(s.sorted[j])) << 32);
                    c = (w & 7ull) == kCandidateOverflow
                            ? self_count_scan<kM>(s, x4, segEnd, longs, nS, table, gap, radius, cell, lenCap,
                                                  scaleMax, chain, j, seg_a(segEnd, j), seg_b(segEnd, j), cj[t], kk[t],
                                                  nk[t], real(0.0), nullptr)
                            : self_count_listed<kM>(s, x4, segEnd, nS, table, radius, cell, scaleMax, j, seg_a(segEnd, j),
                                                    seg_b(segEnd, j), cj[t], kk[t], nk[t], w);
                }
            }
            s.segCount[j] = c;
            counted = counted || c > 0;
        }
    }
    // The barrier after the counts is a vote: when no rod of the workgroup has
    // a contact (the common case) the pool offsets are not summed and the fill
    // does not run.
    const bool any = __syncthreads_or(counted) != 0;
    // Every lane of the rod has read the reductions, the long segments and
    // the flag: record the new ones and reset the rest for the next
    // regeneration.
    rebuilt = rebuild;
    if (on && l == 0) {
        s.rmisc[6] = real_bits(scaleMax);
        s.rmisc[7] = 1;
        s.rmisc[8] = real_bits(real(2.0) * lenShort);
        self_regen_reset(s);
    }
    if (!any) {
        if (on && l == 0) {
            s.misc[0] = 0;
            s.rmisc[0] = real_bits(cell);
            self_list_reset(s, nS);
        }
        return false;
    }
    const int total = group_scan(s.segCount, s.segStart, on ? nS : 0, s.scan, l, G, on);
    if (on && l == 0) {
        s.misc[0] = total < cap ? total : cap;
        if (total > cap) s.misc[1] += total - cap;
        s.rmisc[0] = real_bits(cell);
        self_list_reset(s, nS);
    }
    // The bucket tables are needed only for a segment with more contacts than
    // it keeps.
    bool over = kM == 0;
    if constexpr (kM > 0) {
#pragma unroll
        for (int t = 0; t < TP; ++t) {
            const int j = l + G * t;
            over = over || (on && j < nS && nk[t] > kM);
        }
        over = __syncthreads_or(over) != 0;
    }
    if (over) {
        self_build_tables<TP>(s, nS, table, cell, on && total > 0, cj, l, G);
        // The tables took the arrays that held the candidate lists.
        if (on && total > 0 && l == 0) s.rmisc[7] = 0;
    }
#pragma unroll
    for (int t = 0; t < TP; ++t) {
        const int j = l + G * t;
        if (!(on && j < nS && s.segCount[j] > 0)) continue;
        if constexpr (kM > 0) {
            if (nk[t] <= kM) {
                self_fill_listed<kMr>(s, x4, segEnd, table, cap, radius, friction, cell, seg_a(segEnd, j),
                                      seg_b(segEnd, j), cj[t], s.segStart[j], kk[t], nk[t]);
                continue;
            }
        }
        self_fill_rows(s, x4, segEnd, nS, table, gap, cap, radius, friction, cell, j, seg_a(segEnd, j),
                       seg_b(segEnd, j), cj[t], s.segStart[j]);
    }
    return true;
}

// Whether this lane's displacement bound, `moved` (summed since the rod's
// lists were rebuilt), may let a segment outside the lists reach its cull
// radius (self_regen_group): 4.01 moved plus 1e-6 of the coordinate scale
// reaches the skin. The coordinate scale is the largest centre
// coordinate at the last full regeneration (rmisc[6]) plus `moved`, a bound on
// how far any of this lane's particles moved since. When every lane of the
// rod says no, the rod's largest sum D passes the bound there too.
__device__ __forceinline__ bool self_lists_stale(const SelfLds& s, real moved, real radius) {
    const real scale = bits_real(s.rmisc[6]) + moved;
    return !(real(4.01) * moved + real(1e-6) * (real(1.0) + scale) <= real(kSkinRadii) * radius);
}

// Before a regeneration, with no barrier: whether the workgroup must run the
// full one (self_regen_group) for this lane's segments. It must when the
// rod's candidate lists are not valid (not built, or `stale`: this lane's
// bound says they may miss a pair), when a segment's list overflowed (it is
// scanned at every regeneration), or when a segment is in contact with one of
// its candidates. Otherwise these segments have no contact: with valid lists
// every pair in contact is a candidate pair, and the two culls below drop
// only pairs that are not in contact (their margins, 1e-3 relative and 1e-6
// times the larger centre coordinate of the pair, are far above the rounding
// of the pair's centres, lengths and distances). The centres and lengths are
// formed here from the substep-start positions xs, which no lane writes
// during the substep, by the operations self_regen_group uses.
template <int TP>
__device__ inline bool self_quick_check(const SelfLds& s, const real4* xs, const int* segEnd, int nS, int gap,
                                        real radius, bool on, bool stale, int l, int G) {
    if (!on) return false;
    if (stale || s.rmisc[7] == 0) return true;
    if (nS <= gap) return false;
#pragma unroll
    for (int t = 0; t < TP; ++t) {
        const int j = l + G * t;
        if (j >= nS) continue;
        const unsigned long long w =
            unsigned(s.cellOf[j]) | (static_cast<unsigned long long>(unsigned(s.sorted[j])) << 32);
        if ((w & 7ull) == kCandidateOverflow) return true;
        const int n = int(w & 7ull);
        if (n == 0) continue;
        const int j0 = seg_a(segEnd, j), j1 = seg_b(segEnd, j);
        const v3 a0 = f4v(xs[j0]), a1 = f4v(xs[j1]);
        const real lenJ = norm(sub(a1, a0));
        const v3 cj = scale(add(a0, a1), real(0.5));
        const real reachJ = real(0.5) * lenJ + real(2.0) * radius;
        const v3 dj = sub(a1, a0);
        const real aj = norm2(dj);
        for (int m = 0; m < n; ++m) {
            // In range whatever the word holds.
            const int k = min(j + int((w >> (3 + 10 * m)) & 1023ull), nS - 1);
            const int k0 = seg_a(segEnd, k), k1 = seg_b(segEnd, k);
            const v3 b0 = f4v(xs[k0]), b1 = f4v(xs[k1]);
            const real lenK = norm(sub(b1, b0));
            const v3 ck = scale(add(b0, b1), real(0.5));
            const real eps = real(1e-6) * (real(1.0) + r_max(abs_max3(cj), abs_max3(ck)));
            const real limK = real(1.001) * (reachJ + real(0.5) * lenK) + eps;
            if (norm2(sub(ck, cj)) > limK * limK) continue;
            // Distance from segment j to k's centre (self_count_one's cull).
            const real tq =
                aj > real(1e-30) ? r_max(real(0.0), min_ab(real(1.0), dot(sub(ck, a0), dj) / aj)) : real(0.0);
            const v3 off = sub(add(a0, scale(dj, tq)), ck);
            const real limP = real(1.001) * (real(0.5) * lenK + real(2.0) * radius) + eps;
            if (norm2(off) > limP * limP) continue;
            int c = 0;
            self_pair(s, xs, j0, j1, a0, a1, k0, k1, radius, real(0.0), false, 0, 0, c);
            if (c > 0) return true;
        }
    }
    return false;
}

// Counts -> pool offsets; contacts past the capacity are tallied as overflow.
__device__ inline void self_allocate(const SelfLds& s, int nS, int cap) {
    int total = 0;
    for (int j = 0; j < nS; ++j) {
        s.segStart[j] = total;
        total += s.segCount[j];
    }
    s.misc[0] = total < cap ? total : cap;
    if (total > cap) s.misc[1] += total - cap;
}

__device__ inline void self_reset(const SelfLds& s) {
    const int live = s.misc[0];
    for (int c = 0; c < live; ++c) {
        s.pool[c * kSelfFloats + 13] = real(0.0);
        s.pool[c * kSelfFloats + 14] = real(0.0);
    }
}

// One four-particle contact: unilateral normal projection, then Coulomb
// friction with the per-substep budget; xp4 holds substep-start positions.
// Positions are read, corrected and written through LDS in the order the CPU
// applies them (a particle may appear twice).
__device__ inline void self_project_one_aliased(const SelfLds& s, real4* x4, const real4* xp4, int ci) {
    real* c = s.pool + ci * kSelfFloats;
    int idx[4];
    real wt[4], im[4];
    for (int m = 0; m < 4; ++m) {
        idx[m] = slot_int(c[m]);
        wt[m] = c[4 + m];
        im[m] = x4[idx[m]].w;
    }
    const v3 n = mk(c[8], c[9], c[10]);
    real wEff = real(0.0);
    for (int m = 0; m < 4; ++m) wEff += wt[m] * wt[m] * im[m];
    if (wEff <= real(0.0)) return;
    real C = -c[11];
    for (int m = 0; m < 4; ++m) C += wt[m] * dot(f4v(x4[idx[m]]), n);
    real dl = -C / wEff;
    const real lamN0 = c[13];
    const real cl = max0(lamN0 + dl);
    dl = cl - lamN0;
    c[13] = cl;
    if (dl != real(0.0))
        for (int m = 0; m < 4; ++m) {
            const v3 xm = add(f4v(x4[idx[m]]), scale(n, im[m] * wt[m] * dl));
            x4[idx[m]] = vf4(xm, im[m]);
        }
    const real friction = c[12];
    if (friction <= real(0.0) || cl <= real(0.0)) return;
    v3 dp = mk(real(0.0), real(0.0), real(0.0));
    for (int m = 0; m < 4; ++m) dp = add(dp, scale(sub(f4v(x4[idx[m]]), f4v(xp4[idx[m]])), wt[m]));
    const v3 tang = sub(dp, scale(n, dot(dp, n)));
    const real slide = norm(tang);
    if (slide < real(1e-15)) return;
    const real budget = friction * cl * wEff - c[14];
    if (budget <= real(0.0)) return;
    const real capped = min_ab(slide, budget);
    c[14] = c[14] + capped;
    const v3 corr = scale(tang, -capped / slide);
    for (int m = 0; m < 4; ++m) {
        const v3 xm = add(f4v(x4[idx[m]]), scale(corr, im[m] * wt[m] / wEff));
        x4[idx[m]] = vf4(xm, im[m]);
    }
}

// The same contact when its four particles are distinct (always so for two
// segments of a chain at least two apart): each particle's position is read
// once, receives the same corrections in the same order in registers, and is
// written once.
__device__ __forceinline__ void self_project_one(const SelfLds& s, real4* x4, const real4* xp4, int ci) {
    real* c = s.pool + ci * kSelfFloats;
    const real4* c4 = reinterpret_cast<const real4*>(c);
    const real4 r0 = c4[0], r1 = c4[1], r2 = c4[2], r3 = c4[3];
    const int idx[4] = {slot_int(r0.x), slot_int(r0.y), slot_int(r0.z), slot_int(r0.w)};
    if (idx[0] == idx[1] || idx[0] == idx[2] || idx[0] == idx[3] || idx[1] == idx[2] || idx[1] == idx[3] ||
        idx[2] == idx[3]) {
        self_project_one_aliased(s, x4, xp4, ci);
        return;
    }
    const real wt[4] = {r1.x, r1.y, r1.z, r1.w};
    const v3 n = mk(r2.x, r2.y, r2.z);
    real4 xv[4];
    for (int m = 0; m < 4; ++m) xv[m] = x4[idx[m]];
    real wEff = real(0.0);
    for (int m = 0; m < 4; ++m) wEff += wt[m] * wt[m] * xv[m].w;
    if (wEff <= real(0.0)) return;
    real C = -r2.w;
    for (int m = 0; m < 4; ++m) C += wt[m] * dot(f4v(xv[m]), n);
    real dl = -C / wEff;
    const real lamN0 = r3.y;
    const real cl = max0(lamN0 + dl);
    dl = cl - lamN0;
    c[13] = cl;
    v3 x[4];
    for (int m = 0; m < 4; ++m) x[m] = f4v(xv[m]);
    if (dl != real(0.0))
        for (int m = 0; m < 4; ++m) x[m] = add(x[m], scale(n, xv[m].w * wt[m] * dl));
    const real friction = r3.x;
    bool done = friction <= real(0.0) || cl <= real(0.0);
    v3 tang = mk(real(0.0), real(0.0), real(0.0));
    real slide = real(0.0), budget = real(0.0);
    if (!done) {
        v3 dp = mk(real(0.0), real(0.0), real(0.0));
        for (int m = 0; m < 4; ++m) dp = add(dp, scale(sub(x[m], f4v(xp4[idx[m]])), wt[m]));
        tang = sub(dp, scale(n, dot(dp, n)));
        slide = norm(tang);
        done = slide < real(1e-15);
    }
    if (!done) {
        budget = friction * cl * wEff - r3.z;
        done = budget <= real(0.0);
    }
    if (!done) {
        const real capped = min_ab(slide, budget);
        c[14] = r3.z + capped;
        const v3 corr = scale(tang, -capped / slide);
        for (int m = 0; m < 4; ++m) x[m] = add(x[m], scale(corr, xv[m].w * wt[m] / wEff));
    }
    for (int m = 0; m < 4; ++m) x4[idx[m]] = vf4(x[m], xv[m].w);
}

// Four-particle contacts, sequentially in pool order (= the CPU's list order).
__device__ inline void self_project(const SelfLds& s, real4* x4, const real4* xp4) {
    const int live = s.misc[0];
    for (int ci = 0; ci < live; ++ci) self_project_one(s, x4, xp4, ci);
}

// ------------------------------------------------------------------ coloured self-contact order
//
// An opt-in order (rp_world_desc::self_order = 1) with no counterpart in the
// reference, which projects self contacts one at a time in list order. After
// each regeneration the list is coloured greedily in list order: a contact
// takes the lowest of kSelfColours colours that no earlier contact sharing one
// of its particles holds, or the sequential tail (colour kSelfColours) when all
// are taken. The contacts are then projected colour by colour, in list order
// within a colour, the tail last. Contacts of one colour share no particle, so
// the lanes of a rod's group may project a colour's contacts at once: the result
// is that of projecting them one at a time, bit for bit.

constexpr int kSelfColours = 32;

__device__ __forceinline__ int self_colour_of(const SelfLds& s, int c) {
    return slot_int(s.pool[c * kSelfFloats + 15]);
}

// The colouring, on one lane. A contact's colour goes in its record's last
// word. misc[3] receives the number of colours in use below the tail, plus 64
// when the tail is not empty. The per-particle colour masks take the
// broadphase's per-segment arrays (4 nS ints from cellOf, at least nP; free once
// the list is built); only the contacts' particles are cleared and read. Those
// arrays also hold the candidate lists, which are then rebuilt at the next
// regeneration.
__device__ inline void self_colour(const SelfLds& s) {
    unsigned* used = reinterpret_cast<unsigned*>(s.cellOf);
    const int live = s.misc[0];
    if (live > 0) s.rmisc[7] = 0;
    for (int c = 0; c < live; ++c) {
        const real* rec = s.pool + c * kSelfFloats;
        for (int m = 0; m < 4; ++m) used[slot_int(rec[m])] = 0u;
    }
    int ncol = 0, tail = 0;
    for (int c = 0; c < live; ++c) {
        real* rec = s.pool + c * kSelfFloats;
        const int i0 = slot_int(rec[0]), i1 = slot_int(rec[1]), i2 = slot_int(rec[2]), i3 = slot_int(rec[3]);
        const unsigned mask = used[i0] | used[i1] | used[i2] | used[i3];
        const int col = mask == 0xffffffffu ? kSelfColours : __builtin_ctz(~mask);
        if (col < kSelfColours) {
            const unsigned bit = 1u << col;
            used[i0] |= bit;
            used[i1] |= bit;
            used[i2] |= bit;
            used[i3] |= bit;
            ncol = col + 1 > ncol ? col + 1 : ncol;
        } else {
            tail = 64;
        }
        rec[15] = int_slot(col);
    }
    s.misc[3] = ncol | tail;
}

// The coloured projection by the G lanes of a rod's group: one round per
// colour, lane l taking the colour's contacts at positions l, l + G, ...; then
// the tail on lane 0, in list order. Every lane of the workgroup calls this (a
// barrier ends each round); `on`: this lane's rod projects in coloured order.
__device__ inline void self_project_coloured(const SelfLds& s, real4* x4, const real4* xp4, bool on, int l, int G) {
    const int info = on ? s.misc[3] : 0;
    const int ncol = info & 63;
    const int live = on ? s.misc[0] : 0;
    for (int col = 0; __syncthreads_or(col < ncol) != 0; ++col) {
        if (col < ncol)
            for (int c = l; c < live; c += G)
                if (self_colour_of(s, c) == col) self_project_one(s, x4, xp4, c);
    }
    if ((info & 64) && l == 0)
        for (int c = 0; c < live; ++c)
            if (self_colour_of(s, c) == kSelfColours) self_project_one(s, x4, xp4, c);
}

// The same order on one thread.
__device__ inline void self_project_coloured_serial(const SelfLds& s, real4* x4, const real4* xp4) {
    const int info = s.misc[3];
    const int live = s.misc[0];
    for (int col = 0; col < (info & 63); ++col)
        for (int c = 0; c < live; ++c)
            if (self_colour_of(s, c) == col) self_project_one(s, x4, xp4, c);
    if (info & 64)
        for (int c = 0; c < live; ++c)
            if (self_colour_of(s, c) == kSelfColours) self_project_one(s, x4, xp4, c);
}

}  // namespace RP_PREC
}  // namespace rp

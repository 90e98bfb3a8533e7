// Rod arithmetic for the gfx950 kernels, in the precision of the including
// translation unit (RP_FP64 0: FP32 in namespace rp::f32; 1: FP64 in rp::f64).
//
// Every expression below evaluates in the same order as the reference CPU
// solver's (src/core/math3.h:44-353, src/core/geometry.h:17-224,
// src/core/solver.cpp:35-216,218-267). The library is compiled with
// -ffp-contract=off and IEEE division and square root, so a kernel built from
// these functions produces, operation for operation, the same FP32 roundings as
// the reference CPU path built in single precision: the GPU results are
// bitwise reproducible on the host. Literal zeros and identity entries are kept
// (0 + s, 0 * d) wherever the reference forms them, since they fix the sign of
// zero results.
#pragma once

#include <hip/hip_runtime.h>

#include "rp_layout.h"

#define RP_HD __host__ __device__ __forceinline__

#ifndef RP_FP64
#define RP_FP64 0
#endif
#if RP_FP64
#define RP_PREC f64
#else
#define RP_PREC f32
#endif

namespace rp {
namespace RP_PREC {

#if RP_FP64
using real4 = double4;
using rbits = long long;  // a real's bits, as a signed integer of its width
RP_HD real4 make_real4(real a, real b, real c, real d) { return make_double4(a, b, c, d); }
RP_HD real r_sqrt(real a) { return ::sqrt(a); }
RP_HD real r_floor(real a) { return ::floor(a); }
RP_HD real r_abs(real a) { return ::fabs(a); }
RP_HD real r_max(real a, real b) { return ::fmax(a, b); }
__device__ __forceinline__ rbits real_bits(real a) { return __double_as_longlong(a); }
__device__ __forceinline__ real bits_real(rbits b) { return __longlong_as_double(b); }
#else
using real4 = float4;
using rbits = int;
RP_HD real4 make_real4(real a, real b, real c, real d) { return make_float4(a, b, c, d); }
RP_HD real r_sqrt(real a) { return sqrtf(a); }
RP_HD real r_floor(real a) { return floorf(a); }
RP_HD real r_abs(real a) { return fabsf(a); }
RP_HD real r_max(real a, real b) { return fmaxf(a, b); }
__device__ __forceinline__ rbits real_bits(real a) { return __float_as_int(a); }
__device__ __forceinline__ real bits_real(rbits b) { return __int_as_float(b); }
#endif
RP_HD real r_inf() { return real(__builtin_huge_val()); }
// An int kept in a real slot of LDS (its bits; exact in either precision).
__device__ __forceinline__ real int_slot(int i) { return bits_real(rbits(i)); }
__device__ __forceinline__ int slot_int(real r) { return int(real_bits(r)); }

struct v3 {
    real x, y, z;
};
struct q4 {
    real w, x, y, z;
};
struct m3 {
    real m[3][3];
};

RP_HD v3 mk(real a, real b, real c) { return v3{a, b, c}; }
RP_HD v3 add(v3 a, v3 b) { return v3{a.x + b.x, a.y + b.y, a.z + b.z}; }
RP_HD v3 sub(v3 a, v3 b) { return v3{a.x - b.x, a.y - b.y, a.z - b.z}; }
RP_HD v3 neg(v3 a) { return v3{-a.x, -a.y, -a.z}; }
RP_HD v3 scale(v3 a, real s) { return v3{a.x * s, a.y * s, a.z * s}; }
// Vector over scalar: reciprocal then product.
RP_HD v3 divs(v3 a, real s) { return scale(a, real(1.0) / s); }
RP_HD real dot(v3 a, v3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
RP_HD v3 cross(v3 a, v3 b) {
    return v3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
RP_HD real norm2(v3 a) { return dot(a, a); }
RP_HD real norm(v3 a) { return r_sqrt(dot(a, a)); }
RP_HD v3 cwise(v3 a, v3 b) { return v3{a.x * b.x, a.y * b.y, a.z * b.z}; }
RP_HD real comp(v3 a, int i) { return i == 0 ? a.x : (i == 1 ? a.y : a.z); }

RP_HD m3 mzero() {
    m3 r;
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) r.m[i][j] = real(0.0);
    return r;
}
RP_HD m3 mident(real s) {
    m3 r = mzero();
    r.m[0][0] = r.m[1][1] = r.m[2][2] = s;
    return r;
}
RP_HD m3 mskew(v3 v) {
    m3 r = mzero();
    r.m[0][1] = -v.z;
    r.m[0][2] = v.y;
    r.m[1][0] = v.z;
    r.m[1][2] = -v.x;
    r.m[2][0] = -v.y;
    r.m[2][1] = v.x;
    return r;
}
RP_HD m3 madd(const m3& A, const m3& B) {
    m3 r;
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) r.m[i][j] = A.m[i][j] + B.m[i][j];
    return r;
}
RP_HD m3 mscale(const m3& A, real s) {
    m3 r;
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) r.m[i][j] = A.m[i][j] * s;
    return r;
}
// A diag(d) A^T.
RP_HD m3 sandwich(const m3& A, v3 d) {
    m3 r;
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j)
            r.m[i][j] = A.m[i][0] * d.x * A.m[j][0] + A.m[i][1] * d.y * A.m[j][1] + A.m[i][2] * d.z * A.m[j][2];
    return r;
}

// SPD 3x3 solve by Cholesky; false when a pivot is not positive.
RP_HD bool chol3(const m3& A, v3 b, v3& y) {
    real l00 = A.m[0][0];
    if (l00 <= real(0.0)) return false;
    l00 = r_sqrt(l00);
    const real l10 = A.m[1][0] / l00;
    const real l20 = A.m[2][0] / l00;
    real l11 = A.m[1][1] - l10 * l10;
    if (l11 <= real(0.0)) return false;
    l11 = r_sqrt(l11);
    const real l21 = (A.m[2][1] - l20 * l10) / l11;
    real l22 = A.m[2][2] - l20 * l20 - l21 * l21;
    if (l22 <= real(0.0)) return false;
    l22 = r_sqrt(l22);
    const real z0 = b.x / l00;
    const real z1 = (b.y - l10 * z0) / l11;
    const real z2 = (b.z - l20 * z0 - l21 * z1) / l22;
    y.z = z2 / l22;
    y.y = (z1 - l21 * y.z) / l11;
    y.x = (z0 - l10 * y.y - l20 * y.z) / l00;
    return true;
}

// ------------------------------------------------------------------ quaternions (w, v)

RP_HD v3 im(q4 q) { return v3{q.x, q.y, q.z}; }
RP_HD q4 mkq(real w, v3 v) { return q4{w, v.x, v.y, v.z}; }
RP_HD q4 qmul(q4 a, q4 b) {
    const v3 av = im(a), bv = im(b);
    // w = aw bw - av.bv ;  v = aw bv + bw av + av x bv
    return mkq(a.w * b.w - dot(av, bv), add(add(scale(bv, a.w), scale(av, b.w)), cross(av, bv)));
}
RP_HD q4 conj(q4 q) { return q4{q.w, -q.x, -q.y, -q.z}; }
RP_HD q4 qadd(q4 a, q4 b) { return q4{a.w + b.w, a.x + b.x, a.y + b.y, a.z + b.z}; }
RP_HD q4 qscale(q4 q, real s) { return q4{q.w * s, q.x * s, q.y * s, q.z * s}; }
RP_HD real qdot(q4 a, q4 b) { return a.w * b.w + a.x * b.x + a.y * b.y + a.z * b.z; }
RP_HD q4 qnormalize(q4 q) { return qscale(q, real(1.0) / r_sqrt(qdot(q, q))); }
RP_HD v3 rotate(q4 q, v3 v) {
    const v3 u = im(q);
    const v3 t = scale(cross(u, v), real(2.0));
    return add(add(v, scale(t, q.w)), cross(u, t));
}
RP_HD v3 rotate_inv(q4 q, v3 v) { return rotate(conj(q), v); }
RP_HD q4 body_delta(q4 q, v3 theta) { return qnormalize(qmul(q, mkq(real(1.0), scale(theta, real(0.5))))); }
// Representative on the identity's hemisphere. dot(q, (1,0,0,0)) is q.w plus
// three signed zeros, whose sign test equals q.w's for every finite q.
RP_HD q4 hemi_identity(q4 q) {
    const real d = q.w * real(1.0) + q.x * real(0.0) + q.y * real(0.0) + q.z * real(0.0);
    return d < real(0.0) ? qscale(q, -real(1.0)) : q;
}

// ------------------------------------------------------------------ primitives

enum : int { kPlane = 0, kSphere = 1, kCapsule = 2, kBox = 3 };

// Same field order as the reference's TPrimitive<real>.
struct prim {
    int type;
    v3 a, b, normal, half;
    q4 rot;
    real radius, friction;
};

RP_HD real abs_max3v(v3 a) { return r_max(r_abs(a.x), r_max(r_abs(a.y), r_abs(a.z))); }
RP_HD real clamp01(real v) { return v < real(0.0) ? real(0.0) : (v > real(1.0) ? real(1.0) : v); }
RP_HD real absf_(real v) { return v < real(0.0) ? -v : v; }
RP_HD real clampf_(real v, real lo, real hi) { return v < lo ? lo : (v > hi ? hi : v); }
// std::max(v, 0) and std::min(a, b) exactly, NaN and signed-zero cases included.
RP_HD real max0(real v) { return v < real(0.0) ? real(0.0) : v; }
RP_HD real min_ab(real a, real b) { return b < a ? b : a; }

RP_HD real seg_param(v3 a, v3 b, v3 p) {
    const v3 ab = sub(b, a);
    const real len2 = norm2(ab);
    if (len2 < real(1e-30)) return real(0.0);
    return clamp01(dot(sub(p, a), ab) / len2);
}
RP_HD real sphere_dist(v3 c, real r, v3 p, v3& n) {
    const v3 d = sub(p, c);
    const real len = norm(d);
    n = len > real(1e-12) ? divs(d, len) : mk(real(0.0), real(0.0), real(1.0));
    return len - r;
}
RP_HD real box_dist(const prim& pr, v3 p, v3& n) {
    const v3 loc = rotate_inv(pr.rot, sub(p, pr.a));
    const v3 h = pr.half;
    const v3 cl = mk(clampf_(loc.x, -h.x, h.x), clampf_(loc.y, -h.y, h.y), clampf_(loc.z, -h.z, h.z));
    const v3 delta = sub(loc, cl);
    const real outside = norm(delta);
    if (outside > real(1e-12)) {
        n = rotate(pr.rot, divs(delta, outside));
        return outside;
    }
    const real dx = h.x - absf_(loc.x);
    const real dy = h.y - absf_(loc.y);
    const real dz = h.z - absf_(loc.z);
    v3 f;
    real depth;
    if (dx <= dy && dx <= dz) {
        f = mk(loc.x >= real(0.0) ? real(1.0) : -real(1.0), real(0.0), real(0.0));
        depth = dx;
    } else if (dy <= dz) {
        f = mk(real(0.0), loc.y >= real(0.0) ? real(1.0) : -real(1.0), real(0.0));
        depth = dy;
    } else {
        f = mk(real(0.0), real(0.0), loc.z >= real(0.0) ? real(1.0) : -real(1.0));
        depth = dz;
    }
    n = rotate(pr.rot, f);
    return -depth;
}
// Signed distance, positive outside, with the outward normal.
RP_HD real signed_dist(const prim& pr, v3 p, v3& n) {
    switch (pr.type) {
        case kPlane:
            n = pr.normal;
            return dot(sub(p, pr.a), pr.normal);
        case kSphere:
            return sphere_dist(pr.a, pr.radius, p, n);
        case kCapsule: {
            const real t = seg_param(pr.a, pr.b, p);
            const v3 ax = add(pr.a, scale(sub(pr.b, pr.a), t));
            return sphere_dist(ax, pr.radius, p, n);
        }
        default:
            return box_dist(pr, p, n);
    }
}

// Clamped closest-point parameters between segments [p1,q1] and [p2,q2].
RP_HD void seg_seg(v3 p1, v3 q1, v3 p2, v3 q2, real& s, real& t) {
    const v3 d1 = sub(q1, p1), d2 = sub(q2, p2), r = sub(p1, p2);
    const real a = norm2(d1), e = norm2(d2), f = dot(d2, r);
    const real eps = real(1e-30);
    if (a <= eps && e <= eps) {
        s = t = real(0.0);
        return;
    }
    if (a <= eps) {
        s = real(0.0);
        t = clamp01(f / e);
        return;
    }
    const real c = dot(d1, r);
    if (e <= eps) {
        t = real(0.0);
        s = clamp01(-c / a);
        return;
    }
    const real b = dot(d1, d2);
    const real denom = a * e - b * b;
    s = denom > eps ? clamp01((b * f - c * e) / denom) : real(0.0);
    t = (b * s + f) / e;
    if (t < real(0.0)) {
        t = real(0.0);
        s = clamp01(-c / a);
    } else if (t > real(1.0)) {
        t = real(1.0);
        s = clamp01((b - c) / a);
    }
}

RP_HD unsigned cell_hash(int ix, int iy, int iz, int table) {
    const unsigned h = (static_cast<unsigned>(ix) * 73856093u) ^ (static_cast<unsigned>(iy) * 19349663u) ^
                       (static_cast<unsigned>(iz) * 83492791u);
    // A power-of-two table reduces to a mask; the result is the same.
    const unsigned t = static_cast<unsigned>(table);
    return (t & (t - 1u)) == 0u ? (h & (t - 1u)) : h % t;
}
RP_HD int cell_coord(real v, real cell) { return static_cast<int>(r_floor(v / cell)); }

// ------------------------------------------------------------------ constraint projections
//
// Value-level forms of the projections, shared by every kernel. Each returns
// false when the constraint is skipped (all participants fixed, or a failed
// Cholesky), in which case nothing is written.

// The quotients and scaled compliances a projection forms from values that
// stay fixed while a batch steps (inverse masses, rest lengths, compliances,
// the substep): a kernel that runs many substeps forms them once, by the same
// operations, so its projections give the same results.
struct StretchConst {
    real invL, wL2, w0L, w1L;  // 1/l, (w0 + w1)/(l l), w0/l, w1/l
};
RP_HD StretchConst stretch_const(real w0, real w1, real l) {
    return StretchConst{real(1.0) / l, (w0 + w1) / (l * l), w0 / l, w1 / l};
}
struct BendConst {
    real twoL, invL;  // 2/lbar, 1/lbar
};
RP_HD BendConst bend_const(real lbar) { return BendConst{real(2.0) / lbar, real(1.0) / lbar}; }
// alpha = compliance / h^2, with the batch's per-rod scale.
RP_HD v3 scaled_compliance(v3 compliance, real invH2, real cscale) { return scale(compliance, invH2 * cscale); }
// Whether a constraint is skipped whatever the state (all participants fixed).
RP_HD bool stretch_fixed(real w0, real w1, v3 iI) {
    return w0 == real(0.0) && w1 == real(0.0) && norm2(iI) == real(0.0);
}
RP_HD bool bend_fixed(v3 iIa, v3 iIb) { return norm2(iIa) == real(0.0) && norm2(iIb) == real(0.0); }

// The stretch system matrix mident(wL2) + sandwich(mskew(u), iI), plus alpha
// on the diagonal, with the entries chol3 reads (the lower triangle). Each
// sandwich entry sums three products, one or two of which have a zero factor
// of mskew(u): for finite operands they are signed zeros. The sum then equals
// the remaining products' sum whenever that is nonzero. An off-diagonal entry
// is then added to mident's +0, which turns a zero of either sign into +0 as
// it does 0 + t, so the entry is 0 + t exactly. A diagonal entry can differ
// only in the sign of a zero, and only when wL2 + entry + alpha is zero, where
// chol3 fails whatever that sign (its pivots are tested with <= 0 and any
// product subtracted from them has no negative zero to cancel). The results
// are therefore those of the full products for every finite input.
RP_HD m3 stretch_system(v3 u, v3 d, real wL2, v3 alpha) {
    m3 A;
    A.m[0][0] = (wL2 + ((u.z * d.y) * u.z + (u.y * d.z) * u.y)) + alpha.x;
    A.m[1][1] = (wL2 + ((u.z * d.x) * u.z + (u.x * d.z) * u.x)) + alpha.y;
    A.m[2][2] = (wL2 + ((u.y * d.x) * u.y + (u.x * d.y) * u.x)) + alpha.z;
    A.m[1][0] = real(0.0) + ((-u.x * d.z) * u.y);
    A.m[2][0] = real(0.0) + ((u.x * d.y) * -u.z);
    A.m[2][1] = real(0.0) + ((-u.y * d.x) * u.z);
    A.m[0][1] = A.m[1][0];
    A.m[0][2] = A.m[2][0];
    A.m[1][2] = A.m[2][1];
    return A;
}

// Stretch/shear: C = R(q)^T (x1 - x0)/l - e3, for a constraint that is not
// fixed; `rotates` is norm2(iI) > 0.
RP_HD bool project_stretch_k(v3& x0, v3& x1, q4& q, v3& lam, const StretchConst& k, v3 iI, bool rotates, v3 alpha) {
    const v3 u = scale(rotate_inv(q, sub(x1, x0)), k.invL);
    const v3 C = sub(u, mk(real(0.0), real(0.0), real(1.0)));
    const m3 A = stretch_system(u, iI, k.wL2, alpha);
    v3 dl;
    if (!chol3(A, neg(add(C, cwise(alpha, lam))), dl)) return false;
    lam = add(lam, dl);
    const v3 wv = rotate(q, dl);
    x0 = sub(x0, scale(wv, k.w0L));
    x1 = add(x1, scale(wv, k.w1L));
    if (rotates) q = body_delta(q, cwise(iI, cross(dl, u)));
    return true;
}

RP_HD bool project_stretch(v3& x0, v3& x1, q4& q, v3& lam, real w0, real w1, v3 iI, real l, v3 compliance,
                           real invH2, real cscale) {
    if (stretch_fixed(w0, w1, iI)) return false;
    return project_stretch_k(x0, x1, q, lam, stretch_const(w0, w1, l), iI, norm2(iI) > real(0.0),
                             scaled_compliance(compliance, invH2, cscale));
}

// Bend/twist: p = conj(qa) qb on the identity's hemisphere,
// C = 2 Im(p)/lbar - Omega_rest, for a constraint that is not fixed; rotA and
// rotB are norm2(iIa) > 0 and norm2(iIb) > 0.
RP_HD bool project_bend_k(q4& qa, q4& qb, v3& lam, v3 iIa, v3 iIb, const BendConst& k, v3 alpha, v3 restD,
                          bool rotA, bool rotB) {
    const q4 p = hemi_identity(qmul(conj(qa), qb));
    const v3 pv = im(p);
    const v3 C = sub(scale(pv, k.twoL), restD);
    const m3 sk = mskew(pv);
    const m3 Ja = mscale(madd(mident(-p.w), sk), k.invL);
    const m3 Jb = mscale(madd(mident(p.w), sk), k.invL);
    m3 A = madd(sandwich(Ja, iIa), sandwich(Jb, iIb));
    A.m[0][0] += alpha.x;
    A.m[1][1] += alpha.y;
    A.m[2][2] += alpha.z;
    v3 dl;
    if (!chol3(A, neg(add(C, cwise(alpha, lam))), dl)) return false;
    lam = add(lam, dl);
    if (rotA) {
        const v3 g = scale(sub(scale(dl, -p.w), cross(pv, dl)), k.invL);
        qa = body_delta(qa, cwise(iIa, g));
    }
    if (rotB) {
        const v3 g = scale(sub(scale(dl, p.w), cross(pv, dl)), k.invL);
        qb = body_delta(qb, cwise(iIb, g));
    }
    return true;
}

RP_HD bool project_bend(q4& qa, q4& qb, v3& lam, v3 iIa, v3 iIb, real lbar, v3 compliance, v3 restD, real invH2,
                        real cscale) {
    if (bend_fixed(iIa, iIb)) return false;
    return project_bend_k(qa, qb, lam, iIa, iIb, bend_const(lbar), scaled_compliance(compliance, invH2, cscale),
                          restD, norm2(iIa) > real(0.0), norm2(iIb) > real(0.0));
}

// One rod-vs-world contact slot (a single particle, weight 1).
struct wcontact {
    v3 n;
    real offset, friction, lamN, applied;
    int active;
};

// Regenerate a slot from the particle position at the start of the substep.
RP_HD void world_contact_generate(wcontact& c, const prim& pr, v3 x, real w, real radius, real margin,
                                  real fscale) {
    c.active = 0;
    if (w == real(0.0)) return;
    v3 n;
    const real gap = signed_dist(pr, x, n) - radius;
    if (gap >= margin) return;
    c.active = 1;
    c.n = n;
    // makeContact with weight 1: offset = 1 * dot(x, n) - gap.
    c.offset = (real(0.0) + real(1.0) * dot(x, n)) - gap;
    c.friction = pr.friction * fscale;
}

// Unilateral normal projection, then Coulomb friction with the per-substep
// budget, for the single-particle contact (weights 1, w_eff = w).
RP_HD void world_contact_project(wcontact& c, v3& x, v3 xPrev, real w) {
    if (!c.active) return;
    const real wEff = real(0.0) + real(1.0) * real(1.0) * w;
    if (wEff <= real(0.0)) return;
    real C = -c.offset;
    C += real(1.0) * dot(x, c.n);
    real dl = -C / wEff;
    const real cl = max0(c.lamN + dl);
    dl = cl - c.lamN;
    c.lamN = cl;
    if (dl != real(0.0)) x = add(x, scale(c.n, w * real(1.0) * dl));
    if (c.friction <= real(0.0) || c.lamN <= real(0.0)) return;
    v3 dp = mk(real(0.0), real(0.0), real(0.0));
    dp = add(dp, scale(sub(x, xPrev), real(1.0)));
    const v3 tang = sub(dp, scale(c.n, dot(dp, c.n)));
    const real slide = norm(tang);
    if (slide < real(1e-15)) return;
    const real budget = c.friction * c.lamN * wEff - c.applied;
    if (budget <= real(0.0)) return;
    const real capped = min_ab(slide, budget);
    c.applied += capped;
    const v3 corr = scale(tang, -capped / slide);
    x = add(x, scale(corr, w * real(1.0) / wEff));
}

}  // namespace RP_PREC
}  // namespace rp

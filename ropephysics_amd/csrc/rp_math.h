// FP32 rod arithmetic for the gfx950 kernels.
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

#define RP_HD __host__ __device__ __forceinline__

namespace rp {

struct v3 {
    float x, y, z;
};
struct q4 {
    float w, x, y, z;
};
struct m3 {
    float m[3][3];
};

RP_HD v3 mk(float a, float b, float c) { return v3{a, b, c}; }
RP_HD v3 add(v3 a, v3 b) { return v3{a.x + b.x, a.y + b.y, a.z + b.z}; }
RP_HD v3 sub(v3 a, v3 b) { return v3{a.x - b.x, a.y - b.y, a.z - b.z}; }
RP_HD v3 neg(v3 a) { return v3{-a.x, -a.y, -a.z}; }
RP_HD v3 scale(v3 a, float s) { return v3{a.x * s, a.y * s, a.z * s}; }
// Vector over scalar: reciprocal then product.
RP_HD v3 divs(v3 a, float s) { return scale(a, 1.0f / s); }
RP_HD float dot(v3 a, v3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
RP_HD v3 cross(v3 a, v3 b) {
    return v3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
RP_HD float norm2(v3 a) { return dot(a, a); }
RP_HD float norm(v3 a) { return sqrtf(dot(a, a)); }
RP_HD v3 cwise(v3 a, v3 b) { return v3{a.x * b.x, a.y * b.y, a.z * b.z}; }
RP_HD float comp(v3 a, int i) { return i == 0 ? a.x : (i == 1 ? a.y : a.z); }

RP_HD m3 mzero() {
    m3 r;
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) r.m[i][j] = 0.0f;
    return r;
}
RP_HD m3 mident(float s) {
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
RP_HD m3 mscale(const m3& A, float s) {
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
    float l00 = A.m[0][0];
    if (l00 <= 0.0f) return false;
    l00 = sqrtf(l00);
    const float l10 = A.m[1][0] / l00;
    const float l20 = A.m[2][0] / l00;
    float l11 = A.m[1][1] - l10 * l10;
    if (l11 <= 0.0f) return false;
    l11 = sqrtf(l11);
    const float l21 = (A.m[2][1] - l20 * l10) / l11;
    float l22 = A.m[2][2] - l20 * l20 - l21 * l21;
    if (l22 <= 0.0f) return false;
    l22 = sqrtf(l22);
    const float z0 = b.x / l00;
    const float z1 = (b.y - l10 * z0) / l11;
    const float z2 = (b.z - l20 * z0 - l21 * z1) / l22;
    y.z = z2 / l22;
    y.y = (z1 - l21 * y.z) / l11;
    y.x = (z0 - l10 * y.y - l20 * y.z) / l00;
    return true;
}

// ------------------------------------------------------------------ quaternions (w, v)

RP_HD v3 im(q4 q) { return v3{q.x, q.y, q.z}; }
RP_HD q4 mkq(float w, v3 v) { return q4{w, v.x, v.y, v.z}; }
RP_HD q4 qmul(q4 a, q4 b) {
    const v3 av = im(a), bv = im(b);
    // w = aw bw - av.bv ;  v = aw bv + bw av + av x bv
    return mkq(a.w * b.w - dot(av, bv), add(add(scale(bv, a.w), scale(av, b.w)), cross(av, bv)));
}
RP_HD q4 conj(q4 q) { return q4{q.w, -q.x, -q.y, -q.z}; }
RP_HD q4 qadd(q4 a, q4 b) { return q4{a.w + b.w, a.x + b.x, a.y + b.y, a.z + b.z}; }
RP_HD q4 qscale(q4 q, float s) { return q4{q.w * s, q.x * s, q.y * s, q.z * s}; }
RP_HD float qdot(q4 a, q4 b) { return a.w * b.w + a.x * b.x + a.y * b.y + a.z * b.z; }
RP_HD q4 qnormalize(q4 q) { return qscale(q, 1.0f / sqrtf(qdot(q, q))); }
RP_HD v3 rotate(q4 q, v3 v) {
    const v3 u = im(q);
    const v3 t = scale(cross(u, v), 2.0f);
    return add(add(v, scale(t, q.w)), cross(u, t));
}
RP_HD v3 rotate_inv(q4 q, v3 v) { return rotate(conj(q), v); }
RP_HD q4 body_delta(q4 q, v3 theta) { return qnormalize(qmul(q, mkq(1.0f, scale(theta, 0.5f)))); }
// Representative on the identity's hemisphere. dot(q, (1,0,0,0)) is q.w plus
// three signed zeros, whose sign test equals q.w's for every finite q.
RP_HD q4 hemi_identity(q4 q) {
    const float d = q.w * 1.0f + q.x * 0.0f + q.y * 0.0f + q.z * 0.0f;
    return d < 0.0f ? qscale(q, -1.0f) : q;
}

// ------------------------------------------------------------------ primitives

enum : int { kPlane = 0, kSphere = 1, kCapsule = 2, kBox = 3 };

// Same field order as the reference's TPrimitive<float>.
struct prim {
    int type;
    v3 a, b, normal, half;
    q4 rot;
    float radius, friction;
};

RP_HD float clamp01(float v) { return v < 0.0f ? 0.0f : (v > 1.0f ? 1.0f : v); }
RP_HD float absf_(float v) { return v < 0.0f ? -v : v; }
RP_HD float clampf_(float v, float lo, float hi) { return v < lo ? lo : (v > hi ? hi : v); }
// std::max(v, 0) and std::min(a, b) exactly, NaN and signed-zero cases included.
RP_HD float max0(float v) { return v < 0.0f ? 0.0f : v; }
RP_HD float min_ab(float a, float b) { return b < a ? b : a; }

RP_HD float seg_param(v3 a, v3 b, v3 p) {
    const v3 ab = sub(b, a);
    const float len2 = norm2(ab);
    if (len2 < 1e-30f) return 0.0f;
    return clamp01(dot(sub(p, a), ab) / len2);
}
RP_HD float sphere_dist(v3 c, float r, v3 p, v3& n) {
    const v3 d = sub(p, c);
    const float len = norm(d);
    n = len > 1e-12f ? divs(d, len) : mk(0.0f, 0.0f, 1.0f);
    return len - r;
}
RP_HD float box_dist(const prim& pr, v3 p, v3& n) {
    const v3 loc = rotate_inv(pr.rot, sub(p, pr.a));
    const v3 h = pr.half;
    const v3 cl = mk(clampf_(loc.x, -h.x, h.x), clampf_(loc.y, -h.y, h.y), clampf_(loc.z, -h.z, h.z));
    const v3 delta = sub(loc, cl);
    const float outside = norm(delta);
    if (outside > 1e-12f) {
        n = rotate(pr.rot, divs(delta, outside));
        return outside;
    }
    const float dx = h.x - absf_(loc.x);
    const float dy = h.y - absf_(loc.y);
    const float dz = h.z - absf_(loc.z);
    v3 f;
    float depth;
    if (dx <= dy && dx <= dz) {
        f = mk(loc.x >= 0.0f ? 1.0f : -1.0f, 0.0f, 0.0f);
        depth = dx;
    } else if (dy <= dz) {
        f = mk(0.0f, loc.y >= 0.0f ? 1.0f : -1.0f, 0.0f);
        depth = dy;
    } else {
        f = mk(0.0f, 0.0f, loc.z >= 0.0f ? 1.0f : -1.0f);
        depth = dz;
    }
    n = rotate(pr.rot, f);
    return -depth;
}
// Signed distance, positive outside, with the outward normal.
RP_HD float signed_dist(const prim& pr, v3 p, v3& n) {
    switch (pr.type) {
        case kPlane:
            n = pr.normal;
            return dot(sub(p, pr.a), pr.normal);
        case kSphere:
            return sphere_dist(pr.a, pr.radius, p, n);
        case kCapsule: {
            const float t = seg_param(pr.a, pr.b, p);
            const v3 ax = add(pr.a, scale(sub(pr.b, pr.a), t));
            return sphere_dist(ax, pr.radius, p, n);
        }
        default:
            return box_dist(pr, p, n);
    }
}

// Clamped closest-point parameters between segments [p1,q1] and [p2,q2].
RP_HD void seg_seg(v3 p1, v3 q1, v3 p2, v3 q2, float& s, float& t) {
    const v3 d1 = sub(q1, p1), d2 = sub(q2, p2), r = sub(p1, p2);
    const float a = norm2(d1), e = norm2(d2), f = dot(d2, r);
    const float eps = 1e-30f;
    if (a <= eps && e <= eps) {
        s = t = 0.0f;
        return;
    }
    if (a <= eps) {
        s = 0.0f;
        t = clamp01(f / e);
        return;
    }
    const float c = dot(d1, r);
    if (e <= eps) {
        t = 0.0f;
        s = clamp01(-c / a);
        return;
    }
    const float b = dot(d1, d2);
    const float denom = a * e - b * b;
    s = denom > eps ? clamp01((b * f - c * e) / denom) : 0.0f;
    t = (b * s + f) / e;
    if (t < 0.0f) {
        t = 0.0f;
        s = clamp01(-c / a);
    } else if (t > 1.0f) {
        t = 1.0f;
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
RP_HD int cell_coord(float v, float cell) { return static_cast<int>(floorf(v / cell)); }

// ------------------------------------------------------------------ constraint projections
//
// Value-level forms of the projections, shared by every kernel. Each returns
// false when the constraint is skipped (all participants fixed, or a failed
// Cholesky), in which case nothing is written.

// Stretch/shear: C = R(q)^T (x1 - x0)/l - e3.
RP_HD bool project_stretch(v3& x0, v3& x1, q4& q, v3& lam, float w0, float w1, v3 iI, float l, v3 compliance,
                           float invH2, float cscale) {
    if (w0 == 0.0f && w1 == 0.0f && norm2(iI) == 0.0f) return false;
    const v3 u = divs(rotate_inv(q, sub(x1, x0)), l);
    const v3 C = sub(u, mk(0.0f, 0.0f, 1.0f));
    const v3 alpha = scale(compliance, invH2 * cscale);
    m3 A = madd(mident((w0 + w1) / (l * l)), sandwich(mskew(u), iI));
    A.m[0][0] += alpha.x;
    A.m[1][1] += alpha.y;
    A.m[2][2] += alpha.z;
    v3 dl;
    if (!chol3(A, neg(add(C, cwise(alpha, lam))), dl)) return false;
    lam = add(lam, dl);
    const v3 wv = rotate(q, dl);
    x0 = sub(x0, scale(wv, w0 / l));
    x1 = add(x1, scale(wv, w1 / l));
    if (norm2(iI) > 0.0f) q = body_delta(q, cwise(iI, cross(dl, u)));
    return true;
}

// Bend/twist: p = conj(qa) qb on the identity's hemisphere,
// C = 2 Im(p)/lbar - Omega_rest.
RP_HD bool project_bend(q4& qa, q4& qb, v3& lam, v3 iIa, v3 iIb, float lbar, v3 compliance, v3 restD, float invH2,
                        float cscale) {
    if (norm2(iIa) == 0.0f && norm2(iIb) == 0.0f) return false;
    const q4 p = hemi_identity(qmul(conj(qa), qb));
    const v3 pv = im(p);
    const v3 C = sub(scale(pv, 2.0f / lbar), restD);
    const m3 sk = mskew(pv);
    const m3 Ja = mscale(madd(mident(-p.w), sk), 1.0f / lbar);
    const m3 Jb = mscale(madd(mident(p.w), sk), 1.0f / lbar);
    const v3 alpha = scale(compliance, invH2 * cscale);
    m3 A = madd(sandwich(Ja, iIa), sandwich(Jb, iIb));
    A.m[0][0] += alpha.x;
    A.m[1][1] += alpha.y;
    A.m[2][2] += alpha.z;
    v3 dl;
    if (!chol3(A, neg(add(C, cwise(alpha, lam))), dl)) return false;
    lam = add(lam, dl);
    if (norm2(iIa) > 0.0f) {
        const v3 g = scale(sub(scale(dl, -p.w), cross(pv, dl)), 1.0f / lbar);
        qa = body_delta(qa, cwise(iIa, g));
    }
    if (norm2(iIb) > 0.0f) {
        const v3 g = scale(sub(scale(dl, p.w), cross(pv, dl)), 1.0f / lbar);
        qb = body_delta(qb, cwise(iIb, g));
    }
    return true;
}

// One rod-vs-world contact slot (a single particle, weight 1).
struct wcontact {
    v3 n;
    float offset, friction, lamN, applied;
    int active;
};

// Regenerate a slot from the particle position at the start of the substep.
RP_HD void world_contact_generate(wcontact& c, const prim& pr, v3 x, float w, float radius, float margin,
                                  float fscale) {
    c.active = 0;
    if (w == 0.0f) return;
    v3 n;
    const float gap = signed_dist(pr, x, n) - radius;
    if (gap >= margin) return;
    c.active = 1;
    c.n = n;
    // makeContact with weight 1: offset = 1 * dot(x, n) - gap.
    c.offset = (0.0f + 1.0f * dot(x, n)) - gap;
    c.friction = pr.friction * fscale;
}

// Unilateral normal projection, then Coulomb friction with the per-substep
// budget, for the single-particle contact (weights 1, w_eff = w).
RP_HD void world_contact_project(wcontact& c, v3& x, v3 xPrev, float w) {
    if (!c.active) return;
    const float wEff = 0.0f + 1.0f * 1.0f * w;
    if (wEff <= 0.0f) return;
    float C = -c.offset;
    C += 1.0f * dot(x, c.n);
    float dl = -C / wEff;
    const float cl = max0(c.lamN + dl);
    dl = cl - c.lamN;
    c.lamN = cl;
    if (dl != 0.0f) x = add(x, scale(c.n, w * 1.0f * dl));
    if (c.friction <= 0.0f || c.lamN <= 0.0f) return;
    v3 dp = mk(0.0f, 0.0f, 0.0f);
    dp = add(dp, scale(sub(x, xPrev), 1.0f));
    const v3 tang = sub(dp, scale(c.n, dot(dp, c.n)));
    const float slide = norm(tang);
    if (slide < 1e-15f) return;
    const float budget = c.friction * c.lamN * wEff - c.applied;
    if (budget <= 0.0f) return;
    const float capped = min_ab(slide, budget);
    c.applied += capped;
    const v3 corr = scale(tang, -capped / slide);
    x = add(x, scale(corr, w * 1.0f / wEff));
}

}  // namespace rp

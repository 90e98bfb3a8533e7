// CPU oracle for the batched XPBD Cosserat rod stepper -- TEST INFRASTRUCTURE.
//
// This is the checker, not the product. Only tests/, the smoke check and the
// CPU-baseline leg of bench.py may load it. The product path (HIP kernels
// behind include/ropephysics.h) never links or calls anything in oracle/.
//
// It restates the reference's CPU solver, templated on the scalar so that one
// source gives the FP64 oracle and an FP32 instantiation. The arithmetic order of
// every expression follows the reference so that, compiled with
// -ffp-contract=off, the FP64 build reproduces the reference CPU bit for bit and
// the FP32 build reproduces the reference built with CRS_REAL_FLOAT. Reference
// files followed (paths relative to the RopePhysics checkout):
//   math       src/core/math3.h:44-353
//   geometry   src/core/geometry.h:17-224
//   rod        src/core/rod.h:26-170, src/core/rod.cpp:11-319
//   colouring  src/core/coloring.cpp:11-83
//   collision  src/core/collision.h:35-121, src/core/collision.cpp:10-152
//   solver     src/core/solver.h:15-37, src/core/solver.cpp:35-290
// Two device-side semantics of the reference GPU path are modelled too, because
// the batch API exposes them: a per-rod compliance scale multiplying every alpha
// (src/gpu/gpu_solver.cu:61,93) and a per-rod friction scale multiplying every
// contact friction (src/gpu/gpu_solver.cu:140,302). Both default to 1, which
// leaves the arithmetic identical to the CPU reference.
#pragma once

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <unordered_map>
#include <vector>

namespace orc {

// ------------------------------------------------------------------ scalars

inline float oSqrt(float v) { return std::sqrt(v); }
inline double oSqrt(double v) { return std::sqrt(v); }
inline float oFloor(float v) { return std::floor(v); }
inline double oFloor(double v) { return std::floor(v); }
inline float oExp(float v) { return std::exp(v); }
inline double oExp(double v) { return std::exp(v); }
inline float oCos(float v) { return std::cos(v); }
inline double oCos(double v) { return std::cos(v); }
inline float oSin(float v) { return std::sin(v); }
inline double oSin(double v) { return std::sin(v); }

template <typename T>
constexpr T kPiT = T(3.14159265358979323846);

// ------------------------------------------------------------------ Vec3

template <typename T>
struct V3 {
    T x = 0, y = 0, z = 0;
    V3() = default;
    V3(T a, T b, T c) : x(a), y(b), z(c) {}
    T operator[](int i) const { return i == 0 ? x : (i == 1 ? y : z); }
};

template <typename T> inline V3<T> operator+(V3<T> a, V3<T> b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
template <typename T> inline V3<T> operator-(V3<T> a, V3<T> b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
template <typename T> inline V3<T> operator-(V3<T> a) { return {-a.x, -a.y, -a.z}; }
template <typename T> inline V3<T> operator*(V3<T> a, T s) { return {a.x * s, a.y * s, a.z * s}; }
template <typename T> inline V3<T> operator*(T s, V3<T> a) { return a * s; }
// Division is a reciprocal followed by a product, as in the reference.
template <typename T> inline V3<T> operator/(V3<T> a, T s) { return a * (T(1) / s); }
template <typename T> inline V3<T>& operator+=(V3<T>& a, V3<T> b) { a = a + b; return a; }
template <typename T> inline V3<T>& operator-=(V3<T>& a, V3<T> b) { a = a - b; return a; }

template <typename T> inline T dot(V3<T> a, V3<T> b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
template <typename T> inline V3<T> cross(V3<T> a, V3<T> b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
template <typename T> inline T norm2(V3<T> a) { return dot(a, a); }
template <typename T> inline T norm(V3<T> a) { return oSqrt(dot(a, a)); }
template <typename T> inline V3<T> normalize(V3<T> a) { return a / norm(a); }
template <typename T> inline V3<T> cwise(V3<T> a, V3<T> b) { return {a.x * b.x, a.y * b.y, a.z * b.z}; }

// ------------------------------------------------------------------ Mat3 (row-major)

template <typename T>
struct M3 {
    T m[3][3];
    static M3 zero() {
        M3 r;
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) r.m[i][j] = T(0);
        return r;
    }
    static M3 ident(T s) {
        M3 r = zero();
        r.m[0][0] = r.m[1][1] = r.m[2][2] = s;
        return r;
    }
    static M3 cols(V3<T> c0, V3<T> c1, V3<T> c2) {
        M3 r;
        for (int i = 0; i < 3; ++i) {
            r.m[i][0] = c0[i];
            r.m[i][1] = c1[i];
            r.m[i][2] = c2[i];
        }
        return r;
    }
    static M3 skew(V3<T> v) {
        M3 r = zero();
        r.m[0][1] = -v.z;
        r.m[0][2] = v.y;
        r.m[1][0] = v.z;
        r.m[1][2] = -v.x;
        r.m[2][0] = -v.y;
        r.m[2][1] = v.x;
        return r;
    }
};

template <typename T> inline M3<T> operator+(const M3<T>& A, const M3<T>& B) {
    M3<T> r;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) r.m[i][j] = A.m[i][j] + B.m[i][j];
    return r;
}
template <typename T> inline M3<T> operator*(const M3<T>& A, T s) {
    M3<T> r;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) r.m[i][j] = A.m[i][j] * s;
    return r;
}
// A diag(d) A^T: the shape of every J M^-1 J^T block.
template <typename T> inline M3<T> sandwich(const M3<T>& A, V3<T> d) {
    M3<T> r;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
            r.m[i][j] = A.m[i][0] * d.x * A.m[j][0] + A.m[i][1] * d.y * A.m[j][1] +
                        A.m[i][2] * d.z * A.m[j][2];
    return r;
}

// 3x3 Cholesky solve; false when a pivot is not positive (reference
// src/core/math3.h:206-229).
template <typename T>
inline bool cholSolve3(const M3<T>& A, V3<T> b, V3<T>& y) {
    T l00 = A.m[0][0];
    if (l00 <= T(0)) return false;
    l00 = oSqrt(l00);
    const T l10 = A.m[1][0] / l00;
    const T l20 = A.m[2][0] / l00;
    T l11 = A.m[1][1] - l10 * l10;
    if (l11 <= T(0)) return false;
    l11 = oSqrt(l11);
    const T l21 = (A.m[2][1] - l20 * l10) / l11;
    T l22 = A.m[2][2] - l20 * l20 - l21 * l21;
    if (l22 <= T(0)) return false;
    l22 = oSqrt(l22);
    const T z0 = b.x / l00;
    const T z1 = (b.y - l10 * z0) / l11;
    const T z2 = (b.z - l20 * z0 - l21 * z1) / l22;
    y.z = z2 / l22;
    y.y = (z1 - l21 * y.z) / l11;
    y.x = (z0 - l10 * y.y - l20 * y.z) / l00;
    return true;
}

// ------------------------------------------------------------------ Quat (w, v), body -> world

template <typename T>
struct Q4 {
    T w = 1, x = 0, y = 0, z = 0;
    Q4() = default;
    Q4(T w_, T x_, T y_, T z_) : w(w_), x(x_), y(y_), z(z_) {}
    Q4(T w_, V3<T> v) : w(w_), x(v.x), y(v.y), z(v.z) {}
    V3<T> im() const { return {x, y, z}; }
};

template <typename T> inline Q4<T> operator*(Q4<T> a, Q4<T> b) {
    const V3<T> av = a.im(), bv = b.im();
    return Q4<T>(a.w * b.w - dot(av, bv), a.w * bv + b.w * av + cross(av, bv));
}
template <typename T> inline Q4<T> conj(Q4<T> q) { return {q.w, -q.x, -q.y, -q.z}; }
template <typename T> inline Q4<T> operator+(Q4<T> a, Q4<T> b) {
    return {a.w + b.w, a.x + b.x, a.y + b.y, a.z + b.z};
}
template <typename T> inline Q4<T> operator*(Q4<T> q, T s) { return {q.w * s, q.x * s, q.y * s, q.z * s}; }
template <typename T> inline T dot(Q4<T> a, Q4<T> b) { return a.w * b.w + a.x * b.x + a.y * b.y + a.z * b.z; }
template <typename T> inline Q4<T> normalize(Q4<T> q) { return q * (T(1) / oSqrt(dot(q, q))); }

template <typename T> inline V3<T> rotate(Q4<T> q, V3<T> v) {
    const V3<T> u = q.im();
    const V3<T> t = cross(u, v) * T(2);
    return v + t * q.w + cross(u, t);
}
template <typename T> inline V3<T> rotateInv(Q4<T> q, V3<T> v) { return rotate(conj(q), v); }

// q <- normalize(q * (1, theta/2)): a body-frame rotation increment.
template <typename T> inline Q4<T> bodyDelta(Q4<T> q, V3<T> theta) {
    return normalize(q * Q4<T>(T(1), theta * T(0.5)));
}

template <typename T> inline Q4<T> quatFromTo(V3<T> from, V3<T> to) {
    const T d = dot(from, to);
    if (d > T(1) - T(1e-12)) return Q4<T>();
    if (d < T(-1) + T(1e-12)) {
        V3<T> axis = cross(from, V3<T>(1, 0, 0));
        if (norm2(axis) < T(1e-12)) axis = cross(from, V3<T>(0, 1, 0));
        axis = normalize(axis);
        return Q4<T>(T(0), axis);
    }
    const V3<T> c = cross(from, to);
    const T s = oSqrt((T(1) + d) * T(2));
    return normalize(Q4<T>(s * T(0.5), c / s));
}

template <typename T> inline Q4<T> quatFromMat(const M3<T>& R) {
    const T t = R.m[0][0] + R.m[1][1] + R.m[2][2];
    Q4<T> q;
    if (t > T(0)) {
        T s = oSqrt(t + T(1)) * T(2);
        q = Q4<T>(T(0.25) * s, (R.m[2][1] - R.m[1][2]) / s, (R.m[0][2] - R.m[2][0]) / s,
                  (R.m[1][0] - R.m[0][1]) / s);
    } else if (R.m[0][0] > R.m[1][1] && R.m[0][0] > R.m[2][2]) {
        T s = oSqrt(T(1) + R.m[0][0] - R.m[1][1] - R.m[2][2]) * T(2);
        q = Q4<T>((R.m[2][1] - R.m[1][2]) / s, T(0.25) * s, (R.m[0][1] + R.m[1][0]) / s,
                  (R.m[0][2] + R.m[2][0]) / s);
    } else if (R.m[1][1] > R.m[2][2]) {
        T s = oSqrt(T(1) + R.m[1][1] - R.m[0][0] - R.m[2][2]) * T(2);
        q = Q4<T>((R.m[0][2] - R.m[2][0]) / s, (R.m[0][1] + R.m[1][0]) / s, T(0.25) * s,
                  (R.m[1][2] + R.m[2][1]) / s);
    } else {
        T s = oSqrt(T(1) + R.m[2][2] - R.m[0][0] - R.m[1][1]) * T(2);
        q = Q4<T>((R.m[1][0] - R.m[0][1]) / s, (R.m[0][2] + R.m[2][0]) / s,
                  (R.m[1][2] + R.m[2][1]) / s, T(0.25) * s);
    }
    return normalize(q);
}

template <typename T> inline Q4<T> quatAxisAngle(V3<T> axis, T angle) {
    const T h = angle * T(0.5);
    return Q4<T>(oCos(h), normalize(axis) * oSin(h));
}

template <typename T> inline Q4<T> hemi(Q4<T> q, Q4<T> ref) { return dot(q, ref) < T(0) ? q * T(-1) : q; }

// ------------------------------------------------------------------ primitives and distances

enum PrimType { kPlane = 0, kSphere = 1, kCapsule = 2, kBox = 3 };

template <typename T>
struct Prim {
    int type = kPlane;
    V3<T> a, b, normal, half;
    Q4<T> rot;
    T radius = 0;
    T friction = 0;
};

template <typename T> inline T clamp01(T v) { return v < T(0) ? T(0) : (v > T(1) ? T(1) : v); }
template <typename T> inline T absT(T v) { return v < T(0) ? -v : v; }
template <typename T> inline T clampT(T v, T lo, T hi) { return v < lo ? lo : (v > hi ? hi : v); }

template <typename T> inline T segParam(V3<T> a, V3<T> b, V3<T> p) {
    const V3<T> ab = b - a;
    const T len2 = norm2(ab);
    if (len2 < T(1e-30)) return T(0);
    return clamp01(dot(p - a, ab) / len2);
}

template <typename T> inline T sphereDist(V3<T> c, T r, V3<T> p, V3<T>& n) {
    const V3<T> d = p - c;
    const T len = norm(d);
    n = len > T(1e-12) ? d / len : V3<T>(T(0), T(0), T(1));
    return len - r;
}

template <typename T> inline T boxDist(const Prim<T>& pr, V3<T> p, V3<T>& n) {
    const V3<T> loc = rotateInv(pr.rot, p - pr.a);
    const V3<T> h = pr.half;
    const V3<T> cl(clampT(loc.x, -h.x, h.x), clampT(loc.y, -h.y, h.y), clampT(loc.z, -h.z, h.z));
    const V3<T> delta = loc - cl;
    const T outside = norm(delta);
    if (outside > T(1e-12)) {
        n = rotate(pr.rot, delta / outside);
        return outside;
    }
    const T dx = h.x - absT(loc.x);
    const T dy = h.y - absT(loc.y);
    const T dz = h.z - absT(loc.z);
    V3<T> f;
    T depth;
    if (dx <= dy && dx <= dz) {
        f = V3<T>(loc.x >= T(0) ? T(1) : T(-1), T(0), T(0));
        depth = dx;
    } else if (dy <= dz) {
        f = V3<T>(T(0), loc.y >= T(0) ? T(1) : T(-1), T(0));
        depth = dy;
    } else {
        f = V3<T>(T(0), T(0), loc.z >= T(0) ? T(1) : T(-1));
        depth = dz;
    }
    n = rotate(pr.rot, f);
    return -depth;
}

// Signed distance (positive outside) and outward normal (reference
// src/core/geometry.h:147-165).
template <typename T> inline T signedDist(const Prim<T>& pr, V3<T> p, V3<T>& n) {
    switch (pr.type) {
        case kPlane:
            n = pr.normal;
            return dot(p - pr.a, pr.normal);
        case kSphere:
            return sphereDist(pr.a, pr.radius, p, n);
        case kCapsule: {
            const T t = segParam(pr.a, pr.b, p);
            const V3<T> ax = pr.a + (pr.b - pr.a) * t;
            return sphereDist(ax, pr.radius, p, n);
        }
        default:
            return boxDist(pr, p, n);
    }
}

// Clamped closest-point parameters between segments (reference
// src/core/geometry.h:169-207).
template <typename T>
inline void segSegParams(V3<T> p1, V3<T> q1, V3<T> p2, V3<T> q2, T& s, T& t) {
    const V3<T> d1 = q1 - p1, d2 = q2 - p2, r = p1 - p2;
    const T a = norm2(d1), e = norm2(d2), f = dot(d2, r);
    const T eps = T(1e-30);
    if (a <= eps && e <= eps) {
        s = t = T(0);
        return;
    }
    if (a <= eps) {
        s = T(0);
        t = clamp01(f / e);
        return;
    }
    const T c = dot(d1, r);
    if (e <= eps) {
        t = T(0);
        s = clamp01(-c / a);
        return;
    }
    const T b = dot(d1, d2);
    const T denom = a * e - b * b;
    s = denom > eps ? clamp01((b * f - c * e) / denom) : T(0);
    t = (b * s + f) / e;
    if (t < T(0)) {
        t = T(0);
        s = clamp01(-c / a);
    } else if (t > T(1)) {
        t = T(1);
        s = clamp01((b - c) / a);
    }
}

inline unsigned cellHash(int ix, int iy, int iz, int table) {
    const unsigned h = (static_cast<unsigned>(ix) * 73856093u) ^ (static_cast<unsigned>(iy) * 19349663u) ^
                       (static_cast<unsigned>(iz) * 83492791u);
    return h % static_cast<unsigned>(table);
}
template <typename T> inline int cellOf(T v, T cell) { return static_cast<int>(oFloor(v / cell)); }

// ------------------------------------------------------------------ rod

template <typename T>
struct Material {
    T youngs = T(1e7), poisson = T(0.35), density = T(1000), radius = T(5e-3);
    T shear() const { return youngs / (T(2) * (T(1) + poisson)); }
    T area() const { return kPiT<T> * radius * radius; }
    T second() const {
        const T r2 = radius * radius;
        return kPiT<T> * r2 * r2 / T(4);
    }
    T polar() const { return T(2) * second(); }
    T kappa() const { return T(6) * (T(1) + poisson) / (T(7) + T(6) * poisson); }
    T EI() const { return youngs * second(); }
    T GJ() const { return shear() * polar(); }
};

template <typename T>
struct Rod {
    // particles
    std::vector<V3<T>> x, xPrev, v, extForce, kinVel;
    std::vector<T> invMass, mass;
    // frames (segments, then any fixed ghost frames)
    std::vector<Q4<T>> q, qPrev;
    std::vector<V3<T>> omega, invInertia, inertia, extTorque;
    // stretch/shear constraints
    std::vector<int> sP0, sP1, sSeg;
    std::vector<T> sRest;
    std::vector<V3<T>> sComp, sLam;
    // bend/twist constraints
    std::vector<int> bA, bB;
    std::vector<T> bRest;
    std::vector<V3<T>> bRestD, bComp, bLam;
    Material<T> mat;

    int nP() const { return int(x.size()); }
    int nF() const { return int(q.size()); }
    int nStretch() const { return int(sSeg.size()); }
    int nBend() const { return int(bA.size()); }
};

template <typename T> inline V3<T> stretchComp(const Material<T>& m, T l) {
    const T A = m.area();
    const T EA = m.youngs * A;
    const T GA = m.kappa() * m.shear() * A;
    return {T(1) / (GA * l), T(1) / (GA * l), T(1) / (EA * l)};
}
template <typename T> inline V3<T> bendComp(const Material<T>& m, T l) {
    const T EI = m.EI();
    const T GJ = m.GJ();
    return {T(1) / (EI * l), T(1) / (EI * l), T(1) / (GJ * l)};
}
template <typename T> inline V3<T> darbouxOf(Q4<T> qa, Q4<T> qb, T restLen) {
    Q4<T> rel = conj(qa) * qb;
    rel = hemi(rel, Q4<T>());
    return rel.im() * (T(2) / restLen);
}

template <typename T> inline void setRestFromCurrent(Rod<T>& rod) {
    for (int k = 0; k < rod.nStretch(); ++k) rod.sRest[k] = norm(rod.x[rod.sP1[k]] - rod.x[rod.sP0[k]]);
    const int nSeg = rod.nStretch();
    for (int k = 0; k < rod.nBend(); ++k) {
        const int a = rod.bA[k], b = rod.bB[k];
        if (a < nSeg && b < nSeg) rod.bRest[k] = T(0.5) * (rod.sRest[a] + rod.sRest[b]);
        rod.bRestD[k] = darbouxOf(rod.q[a], rod.q[b], rod.bRest[k]);
    }
}

// Masses, inertias, topology, rest state and compliance of a chain through
// `pts` with segment frames `frames` (reference src/core/rod.cpp:219-283).
template <typename T>
inline Rod<T> buildChain(const std::vector<V3<T>>& pts, const std::vector<Q4<T>>& frames,
                         const Material<T>& mat) {
    const int nP = int(pts.size());
    const int nS = nP - 1;
    Rod<T> rod;
    rod.mat = mat;
    rod.x = pts;
    rod.xPrev = pts;
    rod.v.assign(nP, V3<T>());
    rod.extForce.assign(nP, V3<T>());
    rod.kinVel.assign(nP, V3<T>());
    rod.mass.assign(nP, T(0));
    rod.invMass.assign(nP, T(0));
    rod.q = frames;
    rod.qPrev = frames;
    rod.omega.assign(nS, V3<T>());
    rod.extTorque.assign(nS, V3<T>());
    rod.inertia.assign(nS, V3<T>());
    rod.invInertia.assign(nS, V3<T>());
    const T rho = mat.density, A = mat.area();
    for (int j = 0; j < nS; ++j) {
        const T l = norm(pts[j + 1] - pts[j]);
        const T segMass = rho * A * l;
        rod.mass[j] += T(0.5) * segMass;
        rod.mass[j + 1] += T(0.5) * segMass;
        const T tr = segMass * (T(3) * mat.radius * mat.radius + l * l) / T(12);
        const T ax = T(0.5) * segMass * mat.radius * mat.radius;
        rod.inertia[j] = V3<T>(tr, tr, ax);
        const V3<T> d = rod.inertia[j];
        rod.invInertia[j] = V3<T>(d.x > T(0) ? T(1) / d.x : T(0), d.y > T(0) ? T(1) / d.y : T(0),
                                  d.z > T(0) ? T(1) / d.z : T(0));
    }
    for (int i = 0; i < nP; ++i) rod.invMass[i] = T(1) / rod.mass[i];
    rod.sP0.resize(nS);
    rod.sP1.resize(nS);
    rod.sSeg.resize(nS);
    rod.sRest.assign(nS, T(0));
    rod.sComp.assign(nS, V3<T>());
    rod.sLam.assign(nS, V3<T>());
    for (int j = 0; j < nS; ++j) {
        rod.sP0[j] = j;
        rod.sP1[j] = j + 1;
        rod.sSeg[j] = j;
    }
    const int nB = nS - 1;
    rod.bA.resize(nB);
    rod.bB.resize(nB);
    rod.bRest.assign(nB, T(0));
    rod.bRestD.assign(nB, V3<T>());
    rod.bComp.assign(nB, V3<T>());
    rod.bLam.assign(nB, V3<T>());
    for (int j = 0; j < nB; ++j) {
        rod.bA[j] = j;
        rod.bB[j] = j + 1;
    }
    setRestFromCurrent(rod);
    for (int k = 0; k < rod.nStretch(); ++k) rod.sComp[k] = stretchComp(mat, rod.sRest[k]);
    for (int k = 0; k < rod.nBend(); ++k) rod.bComp[k] = bendComp(mat, rod.bRest[k]);
    return rod;
}

// Straight rod along `dir` (reference src/core/rod.cpp:285-299).
template <typename T>
inline Rod<T> straightRod(int n, T length, const Material<T>& mat, V3<T> origin, V3<T> dir) {
    const V3<T> d = normalize(dir);
    const T l = length / T(n);
    std::vector<V3<T>> pts(n + 1);
    for (int i = 0; i <= n; ++i) pts[i] = origin + d * (l * T(i));
    const Q4<T> frame = quatFromTo(V3<T>(0, 0, 1), d);
    std::vector<Q4<T>> frames(n, frame);
    return buildChain(pts, frames, mat);
}

// Rod along a polyline with parallel-transported frames (reference
// src/core/rod.cpp:301-337).
template <typename T>
inline Rod<T> centerlineRod(const std::vector<V3<T>>& points, const Material<T>& mat, V3<T> seed,
                            T twistPerLength) {
    const int nS = int(points.size()) - 1;
    std::vector<V3<T>> tan(nS);
    for (int j = 0; j < nS; ++j) tan[j] = normalize(points[j + 1] - points[j]);
    V3<T> d1 = seed - tan[0] * dot(seed, tan[0]);
    if (norm2(d1) < T(1e-20)) {
        d1 = V3<T>(1, 0, 0) - tan[0] * dot(V3<T>(1, 0, 0), tan[0]);
        if (norm2(d1) < T(1e-20)) d1 = V3<T>(0, 1, 0) - tan[0] * dot(V3<T>(0, 1, 0), tan[0]);
    }
    d1 = normalize(d1);
    std::vector<Q4<T>> frames(nS);
    frames[0] = quatFromMat(M3<T>::cols(d1, cross(tan[0], d1), tan[0]));
    for (int j = 1; j < nS; ++j) frames[j] = normalize(quatFromTo(tan[j - 1], tan[j]) * frames[j - 1]);
    if (twistPerLength != T(0)) {
        T s = 0;
        for (int j = 0; j < nS; ++j) {
            const T l = norm(points[j + 1] - points[j]);
            const T sMid = s + T(0.5) * l;
            frames[j] = normalize(frames[j] * quatAxisAngle(V3<T>(0, 0, 1), twistPerLength * sMid));
            s += l;
        }
    }
    return buildChain(points, frames, mat);
}

template <typename T> inline void pinParticle(Rod<T>& rod, int i) {
    rod.invMass[i] = T(0);
    rod.v[i] = V3<T>();
}

template <typename T> inline int addFixedFrame(Rod<T>& rod, Q4<T> orientation) {
    const Q4<T> q = normalize(orientation);
    rod.q.push_back(q);
    rod.qPrev.push_back(q);
    rod.omega.push_back(V3<T>());
    rod.extTorque.push_back(V3<T>());
    rod.inertia.push_back(V3<T>());
    rod.invInertia.push_back(V3<T>());
    return rod.nF() - 1;
}

template <typename T> inline void addBend(Rod<T>& rod, int a, int b, T len, V3<T> restD) {
    rod.bA.push_back(a);
    rod.bB.push_back(b);
    rod.bRest.push_back(len);
    rod.bRestD.push_back(restD);
    rod.bComp.push_back(bendComp(rod.mat, len));
    rod.bLam.push_back(V3<T>());
}

// Clamp one end through a fixed ghost frame and a half-length bend element
// (reference src/core/rod.cpp:169-184).
template <typename T> inline int clampEnd(Rod<T>& rod, bool tip) {
    const int lastSeg = rod.nStretch() - 1;
    const int seg = tip ? lastSeg : 0;
    const int particle = tip ? rod.nP() - 1 : 0;
    const T half = T(0.5) * rod.sRest[seg];
    pinParticle(rod, particle);
    const int ghost = addFixedFrame(rod, rod.q[seg]);
    if (tip)
        addBend(rod, seg, ghost, half, V3<T>());
    else
        addBend(rod, ghost, seg, half, V3<T>());
    return ghost;
}

// ------------------------------------------------------------------ colouring

struct Colouring {
    std::vector<int> order;
    std::vector<int> start;
    int colours() const { return int(start.size()) - 1; }
};

// Greedy lowest-free-colour assignment in constraint index order (reference
// src/core/coloring.cpp:11-57).
inline Colouring greedyColour(int count, const std::vector<std::vector<int>>& res) {
    Colouring c;
    if (count == 0) {
        c.start.assign(1, 0);
        return c;
    }
    std::unordered_map<int, std::vector<char>> used;
    std::vector<int> colOf(count, -1);
    int nCol = 0;
    for (int k = 0; k < count; ++k) {
        int col = 0;
        for (;; ++col) {
            bool clash = false;
            for (int rid : res[k]) {
                auto it = used.find(rid);
                if (it != used.end() && col < int(it->second.size()) && it->second[col]) {
                    clash = true;
                    break;
                }
            }
            if (!clash) break;
        }
        colOf[k] = col;
        nCol = std::max(nCol, col + 1);
        for (int rid : res[k]) {
            std::vector<char>& u = used[rid];
            if (int(u.size()) <= col) u.resize(col + 1, 0);
            u[col] = 1;
        }
    }
    c.start.assign(nCol + 1, 0);
    for (int k = 0; k < count; ++k) ++c.start[colOf[k] + 1];
    for (int i = 0; i < nCol; ++i) c.start[i + 1] += c.start[i];
    c.order.resize(count);
    std::vector<int> cur(c.start.begin(), c.start.end() - 1);
    for (int k = 0; k < count; ++k) c.order[cur[colOf[k]]++] = k;
    return c;
}

template <typename T> inline Colouring colourStretch(const Rod<T>& rod) {
    std::vector<std::vector<int>> res(rod.nStretch());
    for (int k = 0; k < rod.nStretch(); ++k) res[k] = {rod.sP0[k], rod.sP1[k], rod.nP() + rod.sSeg[k]};
    return greedyColour(rod.nStretch(), res);
}
template <typename T> inline Colouring colourBend(const Rod<T>& rod) {
    std::vector<std::vector<int>> res(rod.nBend());
    for (int k = 0; k < rod.nBend(); ++k) res[k] = {rod.nP() + rod.bA[k], rod.nP() + rod.bB[k]};
    return greedyColour(rod.nBend(), res);
}

// ------------------------------------------------------------------ contacts and broadphase

template <typename T>
struct Contact {
    int idx[4] = {-1, -1, -1, -1};
    T w[4] = {0, 0, 0, 0};
    V3<T> n;
    T offset = 0, friction = 0, lamN = 0, applied = 0;
    int count = 0;
};

template <typename T>
struct World {
    std::vector<Prim<T>> prims;
    bool self = false;
    T selfFriction = 0;
    int selfSkip = 2;
    T marginRadii = T(0.25);
    int table = 4096;
    int selfOrder = 0;  // 0: self contacts in list order; 1: coloured order (selfColourOrder)
};

template <typename T>
struct Hash {
    T cell = 1;
    int table = 0;
    std::vector<uint32_t> cellOfItem;
    std::vector<int> sorted;
    std::vector<int> start;

    // Counting sort of item centres into `table` buckets (reference
    // src/core/collision.cpp:10-30).
    void build(const std::vector<V3<T>>& c, T cs, int tbl) {
        cell = cs;
        table = tbl;
        const size_t n = c.size();
        cellOfItem.resize(n);
        sorted.resize(n);
        start.assign(size_t(table) + 1, 0);
        for (size_t i = 0; i < n; ++i) {
            cellOfItem[i] = cellHash(cellOf(c[i].x, cell), cellOf(c[i].y, cell), cellOf(c[i].z, cell), table);
            ++start[cellOfItem[i] + 1];
        }
        for (int i = 0; i < table; ++i) start[i + 1] += start[i];
        std::vector<int> cur(start.begin(), start.end() - 1);
        for (size_t i = 0; i < n; ++i) sorted[cur[cellOfItem[i]]++] = int(i);
    }
    // Items in the 27 cells around p, dz outermost, then dy, then dx.
    void query(V3<T> p, std::vector<int>& out) const {
        const int cx = cellOf(p.x, cell), cy = cellOf(p.y, cell), cz = cellOf(p.z, cell);
        for (int dz = -1; dz <= 1; ++dz)
            for (int dy = -1; dy <= 1; ++dy)
                for (int dx = -1; dx <= 1; ++dx) {
                    const uint32_t h = cellHash(cx + dx, cy + dy, cz + dz, table);
                    for (int k = start[h]; k < start[h + 1]; ++k) out.push_back(sorted[k]);
                }
    }
};

template <typename T> inline int selfIndexGap(const Rod<T>& rod, const World<T>& w) {
    T lMin = rod.sRest.empty() ? T(1) : rod.sRest[0];
    for (T l : rod.sRest) lMin = std::min(lMin, l);
    const int geometric = static_cast<int>(oFloor(T(2) * rod.mat.radius / lMin)) + 2;
    return std::max(w.selfSkip, geometric);
}

template <typename T>
inline Contact<T> makeContact(const Rod<T>& rod, const int* idx, const T* wt, int count, V3<T> n, T gap,
                              T friction) {
    Contact<T> c;
    c.count = count;
    c.n = n;
    c.friction = friction;
    T dotSum = 0;
    for (int k = 0; k < count; ++k) {
        c.idx[k] = idx[k];
        c.w[k] = wt[k];
        dotSum += wt[k] * dot(rod.x[idx[k]], n);
    }
    c.offset = dotSum - gap;
    return c;
}

// World contacts per free particle and primitive, then self contacts in hash
// query order (reference src/core/collision.cpp:76-152). `frictionScale` is the
// per-rod scale of the batch API.
template <typename T>
inline void generateContacts(const Rod<T>& rod, const World<T>& w, T frictionScale,
                             std::vector<Contact<T>>& out, Hash<T>& hash, int* worldCount) {
    out.clear();
    const T r = rod.mat.radius;
    for (int i = 0; i < rod.nP(); ++i) {
        if (rod.invMass[i] == T(0)) continue;
        for (const Prim<T>& pr : w.prims) {
            V3<T> n;
            const T d = signedDist(pr, rod.x[i], n);
            const T gap = d - r;
            if (gap >= w.marginRadii * r) continue;
            const int idx[1] = {i};
            const T wt[1] = {T(1)};
            out.push_back(makeContact(rod, idx, wt, 1, n, gap, pr.friction * frictionScale));
        }
    }
    if (worldCount) *worldCount = int(out.size());
    if (!w.self) return;
    const int nSeg = rod.nStretch();
    const int gap = selfIndexGap(rod, w);
    if (nSeg <= gap) return;
    std::vector<V3<T>> centres(nSeg);
    T maxSeg = 0;
    for (int j = 0; j < nSeg; ++j) {
        const V3<T> a = rod.x[rod.sP0[j]], b = rod.x[rod.sP1[j]];
        centres[j] = (a + b) * T(0.5);
        maxSeg = std::max(maxSeg, norm(b - a));
    }
    const T cell = maxSeg + 2 * r;
    hash.build(centres, cell, w.table);
    std::vector<int> cand;
    for (int j = 0; j < nSeg; ++j) {
        cand.clear();
        hash.query(centres[j], cand);
        for (int k : cand) {
            if (k - j < gap) continue;
            T u, v;
            const V3<T> a0 = rod.x[rod.sP0[j]], a1 = rod.x[rod.sP1[j]];
            const V3<T> b0 = rod.x[rod.sP0[k]], b1 = rod.x[rod.sP1[k]];
            segSegParams(a0, a1, b0, b1, u, v);
            const V3<T> pa = a0 + (a1 - a0) * u;
            const V3<T> pb = b0 + (b1 - b0) * v;
            const V3<T> delta = pa - pb;
            const T dist = norm(delta);
            const T sep = dist - 2 * r;
            if (sep >= T(0)) continue;
            const V3<T> n = dist > T(1e-12) ? delta / dist : V3<T>(T(0), T(0), T(1));
            const int idx[4] = {rod.sP0[j], rod.sP1[j], rod.sP0[k], rod.sP1[k]};
            const T wt[4] = {T(1) - u, u, -(T(1) - v), -v};
            out.push_back(makeContact(rod, idx, wt, 4, n, sep, w.selfFriction * frictionScale));
        }
    }
}

// ------------------------------------------------------------------ solver

template <typename T>
struct Params {
    T dt = T(1e-3);
    int substeps = 10;
    int iterations = 1;
    V3<T> gravity = V3<T>(0, 0, T(-9.81));
    T linDamp = 0, angDamp = 0;
    int contactInterval = 1;
    bool coloured = false;
    T complianceScale = T(1);
    T frictionScale = T(1);
};

// Stretch/shear projection (reference src/core/solver.cpp:35-71).
template <typename T>
inline void projectStretch(Rod<T>& rod, T h, T compScale, const Colouring* col) {
    const T invH2 = T(1) / (h * h);
    const int n = rod.nStretch();
    const V3<T> e3(0, 0, 1);
    for (int ii = 0; ii < n; ++ii) {
        const int k = col ? col->order[ii] : ii;
        const int i0 = rod.sP0[k], i1 = rod.sP1[k], j = rod.sSeg[k];
        const T w0 = rod.invMass[i0], w1 = rod.invMass[i1];
        const V3<T> iI = rod.invInertia[j];
        if (w0 == T(0) && w1 == T(0) && norm2(iI) == T(0)) continue;
        const T l = rod.sRest[k];
        const Q4<T> q = rod.q[j];
        const V3<T> u = rotateInv(q, rod.x[i1] - rod.x[i0]) / l;
        const V3<T> C = u - e3;
        const V3<T> alpha = rod.sComp[k] * (invH2 * compScale);
        M3<T> A = M3<T>::ident((w0 + w1) / (l * l)) + sandwich(M3<T>::skew(u), iI);
        A.m[0][0] += alpha.x;
        A.m[1][1] += alpha.y;
        A.m[2][2] += alpha.z;
        V3<T> dl;
        if (!cholSolve3(A, -(C + cwise(alpha, rod.sLam[k])), dl)) continue;
        rod.sLam[k] += dl;
        const V3<T> wv = rotate(q, dl);
        rod.x[i0] -= wv * (w0 / l);
        rod.x[i1] += wv * (w1 / l);
        if (norm2(iI) > T(0)) rod.q[j] = bodyDelta(q, cwise(iI, cross(dl, u)));
    }
}

// Bend/twist projection (reference src/core/solver.cpp:81-123).
template <typename T>
inline void projectBend(Rod<T>& rod, T h, T compScale, const Colouring* col) {
    const T invH2 = T(1) / (h * h);
    const int n = rod.nBend();
    for (int ii = 0; ii < n; ++ii) {
        const int k = col ? col->order[ii] : ii;
        const int a = rod.bA[k], b = rod.bB[k];
        const V3<T> iIa = rod.invInertia[a], iIb = rod.invInertia[b];
        if (norm2(iIa) == T(0) && norm2(iIb) == T(0)) continue;
        const T lbar = rod.bRest[k];
        const Q4<T> p = hemi(conj(rod.q[a]) * rod.q[b], Q4<T>());
        const V3<T> C = p.im() * (T(2) / lbar) - rod.bRestD[k];
        const M3<T> sk = M3<T>::skew(p.im());
        const M3<T> Ja = (M3<T>::ident(-p.w) + sk) * (T(1) / lbar);
        const M3<T> Jb = (M3<T>::ident(p.w) + sk) * (T(1) / lbar);
        const V3<T> alpha = rod.bComp[k] * (invH2 * compScale);
        M3<T> A = sandwich(Ja, iIa) + sandwich(Jb, iIb);
        A.m[0][0] += alpha.x;
        A.m[1][1] += alpha.y;
        A.m[2][2] += alpha.z;
        V3<T> dl;
        if (!cholSolve3(A, -(C + cwise(alpha, rod.bLam[k])), dl)) continue;
        rod.bLam[k] += dl;
        if (norm2(iIa) > T(0)) {
            const V3<T> g = (dl * -p.w - cross(p.im(), dl)) * (T(1) / lbar);
            rod.q[a] = bodyDelta(rod.q[a], cwise(iIa, g));
        }
        if (norm2(iIb) > T(0)) {
            const V3<T> g = (dl * p.w - cross(p.im(), dl)) * (T(1) / lbar);
            rod.q[b] = bodyDelta(rod.q[b], cwise(iIb, g));
        }
    }
}

// The coloured projection order of a contact list (an opt-in order of this
// build; the reference projects in list order): world contacts first, as
// listed, then the self contacts (from index `first`) coloured greedily in
// list order -- each takes the lowest of 32 colours that no earlier self
// contact sharing one of its particles holds, or joins a sequential tail when
// all 32 are taken -- and ordered by colour, then by list position, the tail
// last. Contacts of one colour share no particle, so projecting a colour's
// contacts in any order, or at once, gives the same result.
template <typename T>
inline void selfColourOrder(const Rod<T>& rod, const std::vector<Contact<T>>& set, int first,
                            std::vector<int>& order, std::vector<int>& colour) {
    const int n = int(set.size());
    order.clear();
    for (int i = 0; i < first; ++i) order.push_back(i);
    std::vector<uint32_t> used(size_t(rod.nP()), 0u);
    colour.assign(size_t(n), -1);
    for (int i = first; i < n; ++i) {
        const Contact<T>& c = set[i];
        uint32_t mask = 0u;
        for (int k = 0; k < c.count; ++k) mask |= used[c.idx[k]];
        int col = 32;
        for (int b = 0; b < 32; ++b)
            if (!(mask & (1u << b))) {
                col = b;
                break;
            }
        colour[i] = col;
        if (col < 32)
            for (int k = 0; k < c.count; ++k) used[c.idx[k]] |= 1u << col;
    }
    for (int col = 0; col <= 32; ++col)
        for (int i = first; i < n; ++i)
            if (colour[i] == col) order.push_back(i);
}

// Unilateral contact plus position-level Coulomb friction capped per substep
// (reference src/core/solver.cpp:163-216), in list order or in `order`.
template <typename T>
inline void projectContacts(Rod<T>& rod, std::vector<Contact<T>>& set, const std::vector<int>* order = nullptr) {
    const int n = int(set.size());
    for (int ii = 0; ii < n; ++ii) {
        Contact<T>& c = set[order ? (*order)[ii] : ii];
        T wEff = 0;
        for (int k = 0; k < c.count; ++k) wEff += c.w[k] * c.w[k] * rod.invMass[c.idx[k]];
        if (wEff <= T(0)) continue;
        T C = -c.offset;
        for (int k = 0; k < c.count; ++k) C += c.w[k] * dot(rod.x[c.idx[k]], c.n);
        T dl = -C / wEff;
        const T clamped = std::max(c.lamN + dl, T(0));
        dl = clamped - c.lamN;
        c.lamN = clamped;
        if (dl != T(0))
            for (int k = 0; k < c.count; ++k) rod.x[c.idx[k]] += c.n * (rod.invMass[c.idx[k]] * c.w[k] * dl);
        if (c.friction <= T(0) || c.lamN <= T(0)) continue;
        V3<T> dp;
        for (int k = 0; k < c.count; ++k) dp += (rod.x[c.idx[k]] - rod.xPrev[c.idx[k]]) * c.w[k];
        const V3<T> tang = dp - c.n * dot(dp, c.n);
        const T slide = norm(tang);
        if (slide < T(1e-15)) continue;
        const T budget = c.friction * c.lamN * wEff - c.applied;
        if (budget <= T(0)) continue;
        const T capped = std::min(slide, budget);
        c.applied += capped;
        const V3<T> corr = tang * (-capped / slide);
        for (int k = 0; k < c.count; ++k) rod.x[c.idx[k]] += corr * (rod.invMass[c.idx[k]] * c.w[k] / wEff);
    }
}

// One substep of size h (reference src/core/solver.cpp:218-267).
template <typename T>
inline void substep(Rod<T>& rod, const Params<T>& p, T h, std::vector<Contact<T>>* contacts,
                    const Colouring* sc, const Colouring* bc, const std::vector<int>* contactOrder = nullptr) {
    const int nP = rod.nP(), nF = rod.nF();
    for (int i = 0; i < nP; ++i) {
        rod.xPrev[i] = rod.x[i];
        if (rod.invMass[i] == T(0)) {
            rod.x[i] += rod.kinVel[i] * h;
            continue;
        }
        rod.v[i] += (p.gravity + rod.extForce[i] * rod.invMass[i]) * h;
        rod.x[i] += rod.v[i] * h;
    }
    for (int j = 0; j < nF; ++j) {
        rod.qPrev[j] = rod.q[j];
        if (norm2(rod.invInertia[j]) == T(0)) continue;
        const V3<T> w = rod.omega[j];
        const V3<T> tau = rotateInv(rod.q[j], rod.extTorque[j]);
        rod.omega[j] = w + cwise(rod.invInertia[j], tau - cross(w, cwise(rod.inertia[j], w))) * h;
        rod.q[j] = normalize(rod.q[j] + (rod.q[j] * Q4<T>(T(0), rod.omega[j])) * (T(0.5) * h));
    }
    for (int it = 0; it < p.iterations; ++it) {
        projectStretch(rod, h, p.complianceScale, sc);
        projectBend(rod, h, p.complianceScale, bc);
        if (contacts) projectContacts(rod, *contacts, contactOrder);
    }
    const T linDecay = oExp(-p.linDamp * h);
    const T angDecay = oExp(-p.angDamp * h);
    for (int i = 0; i < nP; ++i) {
        if (rod.invMass[i] == T(0)) {
            rod.v[i] = rod.kinVel[i];
            continue;
        }
        rod.v[i] = (rod.x[i] - rod.xPrev[i]) / h * linDecay;
    }
    for (int j = 0; j < nF; ++j) {
        if (norm2(rod.invInertia[j]) == T(0)) {
            rod.omega[j] = V3<T>();
            continue;
        }
        const Q4<T> dq = hemi(conj(rod.qPrev[j]) * rod.q[j], Q4<T>());
        rod.omega[j] = dq.im() * (T(2) / h) * angDecay;
    }
}

// Per-rod scratch carried across steps (reference src/core/solver.h:41-44).
template <typename T>
struct Context {
    std::vector<Contact<T>> contacts;
    Hash<T> hash;
    Colouring sc, bc;
    int worldContacts = 0;
    std::vector<int> contactOrder;   // the coloured order, when the world asks for it
    std::vector<int> contactColour;  // per contact: its colour (32: the tail), -1 for world contacts
};

// One step of p.dt (reference src/core/solver.cpp:269-290). With `world` null
// this is the contact-free step.
template <typename T>
inline void step(Rod<T>& rod, const Params<T>& p, const World<T>* world, Context<T>& ctx) {
    const T h = p.dt / T(p.substeps);
    const Colouring* sc = p.coloured ? &ctx.sc : nullptr;
    const Colouring* bc = p.coloured ? &ctx.bc : nullptr;
    if (p.coloured && ctx.sc.start.empty()) {
        ctx.sc = colourStretch(rod);
        ctx.bc = colourBend(rod);
    }
    const int interval = std::max(1, p.contactInterval);
    for (int i = 0; i < p.substeps; ++i) {
        if (world && i % interval == 0) {
            generateContacts(rod, *world, p.frictionScale, ctx.contacts, ctx.hash, &ctx.worldContacts);
            if (world->selfOrder)
                selfColourOrder(rod, ctx.contacts, ctx.worldContacts, ctx.contactOrder, ctx.contactColour);
        }
        std::fill(rod.sLam.begin(), rod.sLam.end(), V3<T>());
        std::fill(rod.bLam.begin(), rod.bLam.end(), V3<T>());
        if (world)
            for (Contact<T>& c : ctx.contacts) {
                c.lamN = 0;
                c.applied = 0;
            }
        substep(rod, p, h, world ? &ctx.contacts : nullptr, sc, bc,
                world && world->selfOrder ? &ctx.contactOrder : nullptr);
    }
}

}  // namespace orc

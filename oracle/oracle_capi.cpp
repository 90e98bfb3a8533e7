// C interface over the oracle restatement -- TEST INFRASTRUCTURE (see
// oracle_capi.h). Instantiates rp_oracle.hpp in FP64 and FP32.
#include "oracle_capi.h"

#include <thread>
#include <vector>

#include "opcount.hpp"
#include "rp_oracle.hpp"

namespace {

template <typename T>
struct Holder {
    orc::Rod<T> rod;
    orc::Context<T> ctx;
};

struct Handle {
    int precision = 64;
    Holder<double> d;
    Holder<float> f;
};

template <typename T>
orc::V3<T> v3(const double* p) {
    return orc::V3<T>(T(p[0]), T(p[1]), T(p[2]));
}
template <typename T>
orc::Q4<T> q4(const double* p) {
    return orc::Q4<T>(T(p[0]), T(p[1]), T(p[2]), T(p[3]));
}
template <typename T>
orc::Material<T> matOf(const double m[4]) {
    orc::Material<T> mat;
    mat.youngs = T(m[0]);
    mat.poisson = T(m[1]);
    mat.density = T(m[2]);
    mat.radius = T(m[3]);
    return mat;
}

template <typename T>
orc::Prim<T> primOf(const orc_prim& p) {
    orc::Prim<T> r;
    r.type = p.type;
    r.a = v3<T>(p.a);
    r.b = v3<T>(p.b);
    r.normal = v3<T>(p.normal);
    r.half = v3<T>(p.half);
    r.rot = q4<T>(p.rot);
    r.radius = T(p.radius);
    r.friction = T(p.friction);
    return r;
}

template <typename T>
orc::World<T> worldOf(const orc_world& w) {
    orc::World<T> r;
    for (int i = 0; i < w.numPrims; ++i) r.prims.push_back(primOf<T>(w.prims[i]));
    r.self = w.selfCollision != 0;
    r.selfFriction = T(w.selfFriction);
    r.selfSkip = w.selfSkip;
    r.marginRadii = T(w.marginRadii);
    r.table = w.hashTableSize;
    r.selfOrder = w.selfOrder;
    return r;
}

template <typename T>
orc::Params<T> paramsOf(const orc_params& p) {
    orc::Params<T> r;
    r.dt = T(p.dt);
    r.substeps = p.substeps;
    r.iterations = p.iterations;
    r.gravity = v3<T>(p.gravity);
    r.linDamp = T(p.linDamp);
    r.angDamp = T(p.angDamp);
    r.contactInterval = p.contactInterval;
    r.coloured = p.coloured != 0;
    r.complianceScale = T(p.complianceScale);
    r.frictionScale = T(p.frictionScale);
    return r;
}

template <typename T, typename Fn>
void putVec(const std::vector<T>& src, double* dst, int comps, Fn get) {
    if (!dst) return;
    for (size_t i = 0; i < src.size(); ++i)
        for (int c = 0; c < comps; ++c) dst[i * comps + c] = double(get(src[i], c));
}

template <typename T>
void getRod(const orc::Rod<T>& r, orc_rod_view* o) {
    o->nP = r.nP();
    o->nF = r.nF();
    o->nStretch = r.nStretch();
    o->nBend = r.nBend();
    auto vc = [](const orc::V3<T>& v, int c) { return v[c]; };
    auto qc = [](const orc::Q4<T>& q, int c) { return c == 0 ? q.w : (c == 1 ? q.x : (c == 2 ? q.y : q.z)); };
    auto sc = [](const T& s, int) { return s; };
    putVec(r.x, o->x, 3, vc);
    putVec(r.xPrev, o->xPrev, 3, vc);
    putVec(r.v, o->v, 3, vc);
    putVec(r.invMass, o->invMass, 1, sc);
    putVec(r.mass, o->mass, 1, sc);
    putVec(r.extForce, o->extForce, 3, vc);
    putVec(r.kinVel, o->kinVel, 3, vc);
    putVec(r.q, o->q, 4, qc);
    putVec(r.qPrev, o->qPrev, 4, qc);
    putVec(r.omega, o->omega, 3, vc);
    putVec(r.invInertia, o->invInertia, 3, vc);
    putVec(r.inertia, o->inertia, 3, vc);
    putVec(r.extTorque, o->extTorque, 3, vc);
    for (int k = 0; k < r.nStretch(); ++k) {
        if (o->sP0) o->sP0[k] = r.sP0[k];
        if (o->sP1) o->sP1[k] = r.sP1[k];
        if (o->sSeg) o->sSeg[k] = r.sSeg[k];
    }
    putVec(r.sRest, o->sRest, 1, sc);
    putVec(r.sComp, o->sComp, 3, vc);
    putVec(r.sLam, o->sLam, 3, vc);
    for (int k = 0; k < r.nBend(); ++k) {
        if (o->bA) o->bA[k] = r.bA[k];
        if (o->bB) o->bB[k] = r.bB[k];
    }
    putVec(r.bRest, o->bRest, 1, sc);
    putVec(r.bRestD, o->bRestD, 3, vc);
    putVec(r.bComp, o->bComp, 3, vc);
    putVec(r.bLam, o->bLam, 3, vc);
    o->material[0] = double(r.mat.youngs);
    o->material[1] = double(r.mat.poisson);
    o->material[2] = double(r.mat.density);
    o->material[3] = double(r.mat.radius);
}

template <typename T>
void takeV(std::vector<orc::V3<T>>& dst, const double* src) {
    if (!src) return;
    for (size_t i = 0; i < dst.size(); ++i) dst[i] = v3<T>(src + 3 * i);
}
template <typename T>
void takeQ(std::vector<orc::Q4<T>>& dst, const double* src) {
    if (!src) return;
    for (size_t i = 0; i < dst.size(); ++i) dst[i] = q4<T>(src + 4 * i);
}
template <typename T>
void takeS(std::vector<T>& dst, const double* src) {
    if (!src) return;
    for (size_t i = 0; i < dst.size(); ++i) dst[i] = T(src[i]);
}
void takeI(std::vector<int>& dst, const int* src) {
    if (!src) return;
    for (size_t i = 0; i < dst.size(); ++i) dst[i] = src[i];
}

template <typename T>
int setRod(orc::Rod<T>& r, const orc_rod_view* in) {
    // Resizing is allowed: a rod may be rebuilt wholesale from arrays.
    r.x.resize(in->nP);
    r.xPrev.resize(in->nP);
    r.v.resize(in->nP);
    r.invMass.resize(in->nP);
    r.mass.resize(in->nP);
    r.extForce.resize(in->nP);
    r.kinVel.resize(in->nP);
    r.q.resize(in->nF);
    r.qPrev.resize(in->nF);
    r.omega.resize(in->nF);
    r.invInertia.resize(in->nF);
    r.inertia.resize(in->nF);
    r.extTorque.resize(in->nF);
    r.sP0.resize(in->nStretch);
    r.sP1.resize(in->nStretch);
    r.sSeg.resize(in->nStretch);
    r.sRest.resize(in->nStretch);
    r.sComp.resize(in->nStretch);
    r.sLam.resize(in->nStretch);
    r.bA.resize(in->nBend);
    r.bB.resize(in->nBend);
    r.bRest.resize(in->nBend);
    r.bRestD.resize(in->nBend);
    r.bComp.resize(in->nBend);
    r.bLam.resize(in->nBend);
    takeV(r.x, in->x);
    takeV(r.xPrev, in->xPrev);
    takeV(r.v, in->v);
    takeS(r.invMass, in->invMass);
    takeS(r.mass, in->mass);
    takeV(r.extForce, in->extForce);
    takeV(r.kinVel, in->kinVel);
    takeQ(r.q, in->q);
    takeQ(r.qPrev, in->qPrev);
    takeV(r.omega, in->omega);
    takeV(r.invInertia, in->invInertia);
    takeV(r.inertia, in->inertia);
    takeV(r.extTorque, in->extTorque);
    takeI(r.sP0, in->sP0);
    takeI(r.sP1, in->sP1);
    takeI(r.sSeg, in->sSeg);
    takeS(r.sRest, in->sRest);
    takeV(r.sComp, in->sComp);
    takeV(r.sLam, in->sLam);
    takeI(r.bA, in->bA);
    takeI(r.bB, in->bB);
    takeS(r.bRest, in->bRest);
    takeV(r.bRestD, in->bRestD);
    takeV(r.bComp, in->bComp);
    takeV(r.bLam, in->bLam);
    r.mat = matOf<T>(in->material);
    return 0;
}

template <typename T>
void stepHolder(Holder<T>& h, const orc_params* p, const orc_world* w, int steps) {
    const orc::Params<T> pp = paramsOf<T>(*p);
    if (w) {
        const orc::World<T> ww = worldOf<T>(*w);
        for (int s = 0; s < steps; ++s) orc::step(h.rod, pp, &ww, h.ctx);
    } else {
        for (int s = 0; s < steps; ++s) orc::step(h.rod, pp, static_cast<const orc::World<T>*>(nullptr), h.ctx);
    }
}

Handle* H(void* p) { return static_cast<Handle*>(p); }

template <typename To, typename From>
std::vector<To> conv(const std::vector<From>& v) {
    std::vector<To> r(v.size());
    for (size_t i = 0; i < v.size(); ++i) r[i] = To(double(v[i]));
    return r;
}
template <typename To, typename From>
std::vector<orc::V3<To>> convV(const std::vector<orc::V3<From>>& v) {
    std::vector<orc::V3<To>> r(v.size());
    for (size_t i = 0; i < v.size(); ++i) r[i] = orc::V3<To>(To(double(v[i].x)), To(double(v[i].y)), To(double(v[i].z)));
    return r;
}
template <typename To, typename From>
std::vector<orc::Q4<To>> convQ(const std::vector<orc::Q4<From>>& v) {
    std::vector<orc::Q4<To>> r(v.size());
    for (size_t i = 0; i < v.size(); ++i)
        r[i] = orc::Q4<To>(To(double(v[i].w)), To(double(v[i].x)), To(double(v[i].y)), To(double(v[i].z)));
    return r;
}

orc::Rod<orc::Cnt> countingRod(const orc::Rod<double>& s) {
    using orc::Cnt;
    orc::Rod<Cnt> r;
    r.x = convV<Cnt>(s.x);
    r.xPrev = convV<Cnt>(s.xPrev);
    r.v = convV<Cnt>(s.v);
    r.extForce = convV<Cnt>(s.extForce);
    r.kinVel = convV<Cnt>(s.kinVel);
    r.invMass = conv<Cnt>(s.invMass);
    r.mass = conv<Cnt>(s.mass);
    r.q = convQ<Cnt>(s.q);
    r.qPrev = convQ<Cnt>(s.qPrev);
    r.omega = convV<Cnt>(s.omega);
    r.invInertia = convV<Cnt>(s.invInertia);
    r.inertia = convV<Cnt>(s.inertia);
    r.extTorque = convV<Cnt>(s.extTorque);
    r.sP0 = s.sP0;
    r.sP1 = s.sP1;
    r.sSeg = s.sSeg;
    r.sRest = conv<Cnt>(s.sRest);
    r.sComp = convV<Cnt>(s.sComp);
    r.sLam = convV<Cnt>(s.sLam);
    r.bA = s.bA;
    r.bB = s.bB;
    r.bRest = conv<Cnt>(s.bRest);
    r.bRestD = convV<Cnt>(s.bRestD);
    r.bComp = convV<Cnt>(s.bComp);
    r.bLam = convV<Cnt>(s.bLam);
    r.mat.youngs = s.mat.youngs;
    r.mat.poisson = s.mat.poisson;
    r.mat.density = s.mat.density;
    r.mat.radius = s.mat.radius;
    return r;
}

}  // namespace

template <typename T>
int selfContactsOf(const Holder<T>& h, int* idx4, int* colour, int* rank, int capacity) {
    const auto& set = h.ctx.contacts;
    const int first = h.ctx.worldContacts, n = int(set.size()) - first;
    const bool coloured = int(h.ctx.contactColour.size()) == int(set.size()) && !set.empty();
    std::vector<int> pos(set.size(), -1);
    if (coloured)
        for (size_t ii = size_t(first); ii < h.ctx.contactOrder.size(); ++ii)
            pos[size_t(h.ctx.contactOrder[ii])] = int(ii) - first;
    for (int i = 0; i < n && i < capacity; ++i) {
        const auto& c = set[size_t(first + i)];
        for (int k = 0; k < 4; ++k) idx4[4 * i + k] = k < c.count ? c.idx[k] : -1;
        colour[i] = coloured ? h.ctx.contactColour[size_t(first + i)] : -1;
        rank[i] = coloured ? pos[size_t(first + i)] : -1;
    }
    return n;
}

extern "C" {

void* orc_rod_straight(int precision, int n, double length, const double material[4], const double origin[3],
                       const double dir[3]) {
    Handle* h = new Handle();
    h->precision = precision;
    if (precision == 32)
        h->f.rod = orc::straightRod<float>(n, float(length), matOf<float>(material), v3<float>(origin),
                                           v3<float>(dir));
    else
        h->d.rod = orc::straightRod<double>(n, length, matOf<double>(material), v3<double>(origin),
                                            v3<double>(dir));
    return h;
}

void* orc_rod_centerline(int precision, int numPoints, const double* points, const double material[4],
                         const double seed[3], double twistPerLength) {
    Handle* h = new Handle();
    h->precision = precision;
    if (precision == 32) {
        std::vector<orc::V3<float>> pts(numPoints);
        for (int i = 0; i < numPoints; ++i) pts[i] = v3<float>(points + 3 * i);
        h->f.rod = orc::centerlineRod<float>(pts, matOf<float>(material), v3<float>(seed), float(twistPerLength));
    } else {
        std::vector<orc::V3<double>> pts(numPoints);
        for (int i = 0; i < numPoints; ++i) pts[i] = v3<double>(points + 3 * i);
        h->d.rod = orc::centerlineRod<double>(pts, matOf<double>(material), v3<double>(seed), twistPerLength);
    }
    return h;
}

int orc_rod_clamp_end(void* rod, int tip) {
    Handle* h = H(rod);
    return h->precision == 32 ? orc::clampEnd(h->f.rod, tip != 0) : orc::clampEnd(h->d.rod, tip != 0);
}

void orc_rod_pin(void* rod, int particle) {
    Handle* h = H(rod);
    if (h->precision == 32)
        orc::pinParticle(h->f.rod, particle);
    else
        orc::pinParticle(h->d.rod, particle);
}

void orc_rod_set_rest_from_current(void* rod) {
    Handle* h = H(rod);
    if (h->precision == 32)
        orc::setRestFromCurrent(h->f.rod);
    else
        orc::setRestFromCurrent(h->d.rod);
}

void orc_rod_sizes(void* rod, int out[4]) {
    Handle* h = H(rod);
    if (h->precision == 32) {
        out[0] = h->f.rod.nP();
        out[1] = h->f.rod.nF();
        out[2] = h->f.rod.nStretch();
        out[3] = h->f.rod.nBend();
    } else {
        out[0] = h->d.rod.nP();
        out[1] = h->d.rod.nF();
        out[2] = h->d.rod.nStretch();
        out[3] = h->d.rod.nBend();
    }
}

void orc_rod_get(void* rod, orc_rod_view* out) {
    Handle* h = H(rod);
    if (h->precision == 32)
        getRod(h->f.rod, out);
    else
        getRod(h->d.rod, out);
}

int orc_rod_set(void* rod, const orc_rod_view* in) {
    Handle* h = H(rod);
    // A new topology invalidates cached colourings.
    h->f.ctx.sc = orc::Colouring();
    h->f.ctx.bc = orc::Colouring();
    h->d.ctx.sc = orc::Colouring();
    h->d.ctx.bc = orc::Colouring();
    return h->precision == 32 ? setRod(h->f.rod, in) : setRod(h->d.rod, in);
}

void orc_rod_free(void* rod) { delete H(rod); }

int orc_step(void* rod, const orc_params* p, const orc_world* world, int steps) {
    Handle* h = H(rod);
    if (h->precision == 32)
        stepHolder(h->f, p, world, steps);
    else
        stepHolder(h->d, p, world, steps);
    return 0;
}

void orc_contact_counts(void* rod, int* worldContacts, int* selfContacts) {
    Handle* h = H(rod);
    int total = 0, world = 0;
    if (h->precision == 32) {
        total = int(h->f.ctx.contacts.size());
        world = h->f.ctx.worldContacts;
    } else {
        total = int(h->d.ctx.contacts.size());
        world = h->d.ctx.worldContacts;
    }
    *worldContacts = world;
    *selfContacts = total - world;
}

int orc_step_many(void** rods, int count, const orc_params* p, const orc_world* const* worlds, int steps,
                  int threads) {
    if (threads < 1) threads = 1;
    std::vector<std::thread> pool;
    // Static block partition over rods, one block per thread.
    for (int t = 0; t < threads; ++t) {
        const int begin = int((long long)count * t / threads);
        const int end = int((long long)count * (t + 1) / threads);
        pool.emplace_back([=]() {
            for (int r = begin; r < end; ++r) orc_step(rods[r], p, worlds ? worlds[r] : nullptr, steps);
        });
    }
    for (std::thread& th : pool) th.join();
    return 0;
}

int orc_step_many_driven(void** rods, int count, const orc_params* p, const orc_world* const* worlds, int steps,
                         int threads, const int* driven, const double* kin_vel) {
    if (threads < 1) threads = 1;
    std::vector<std::thread> pool;
    for (int t = 0; t < threads; ++t) {
        const int begin = int((long long)count * t / threads);
        const int end = int((long long)count * (t + 1) / threads);
        pool.emplace_back([=]() {
            for (int r = begin; r < end; ++r) {
                Handle* h = H(rods[r]);
                for (int s = 0; s < steps; ++s) {
                    if (driven[r] >= 0) {
                        const double* v = kin_vel + (size_t(r) * steps + s) * 3;
                        if (h->precision == 32)
                            h->f.rod.kinVel[driven[r]] = v3<float>(v);
                        else
                            h->d.rod.kinVel[driven[r]] = v3<double>(v);
                    }
                    orc_step(rods[r], p, worlds ? worlds[r] : nullptr, 1);
                }
            }
        });
    }
    for (std::thread& th : pool) th.join();
    return 0;
}

int orc_chol_solve3(int precision, const double A[9], const double b[3], double y[3]) {
    if (precision == 32) {
        orc::M3<float> M;
        for (int i = 0; i < 9; ++i) M.m[i / 3][i % 3] = float(A[i]);
        orc::V3<float> out;
        const bool ok = orc::cholSolve3(M, v3<float>(b), out);
        y[0] = out.x;
        y[1] = out.y;
        y[2] = out.z;
        return ok ? 1 : 0;
    }
    orc::M3<double> M;
    for (int i = 0; i < 9; ++i) M.m[i / 3][i % 3] = A[i];
    orc::V3<double> out;
    const bool ok = orc::cholSolve3(M, v3<double>(b), out);
    y[0] = out.x;
    y[1] = out.y;
    y[2] = out.z;
    return ok ? 1 : 0;
}

double orc_signed_distance(int precision, const orc_prim* prim, const double p[3], double n[3]) {
    if (precision == 32) {
        orc::V3<float> nn;
        const float d = orc::signedDist(primOf<float>(*prim), v3<float>(p), nn);
        n[0] = nn.x;
        n[1] = nn.y;
        n[2] = nn.z;
        return d;
    }
    orc::V3<double> nn;
    const double d = orc::signedDist(primOf<double>(*prim), v3<double>(p), nn);
    n[0] = nn.x;
    n[1] = nn.y;
    n[2] = nn.z;
    return d;
}

void orc_seg_seg(int precision, const double p1[3], const double q1[3], const double p2[3], const double q2[3],
                 double st[2]) {
    if (precision == 32) {
        float s, t;
        orc::segSegParams(v3<float>(p1), v3<float>(q1), v3<float>(p2), v3<float>(q2), s, t);
        st[0] = s;
        st[1] = t;
        return;
    }
    double s, t;
    orc::segSegParams(v3<double>(p1), v3<double>(q1), v3<double>(p2), v3<double>(q2), s, t);
    st[0] = s;
    st[1] = t;
}

unsigned orc_hash_cell(int ix, int iy, int iz, int table) { return orc::cellHash(ix, iy, iz, table); }

int orc_self_contacts(void* rod, int* idx4, int* colour, int* rank, int capacity) {
    Handle* h = H(rod);
    return h->precision == 32 ? selfContactsOf(h->f, idx4, colour, rank, capacity)
                              : selfContactsOf(h->d, idx4, colour, rank, capacity);
}

int orc_cell_coord(int precision, double v, double cell) {
    return precision == 32 ? orc::cellOf(float(v), float(cell)) : orc::cellOf(v, cell);
}

int orc_self_gap(void* rod, const orc_world* world) {
    Handle* h = H(rod);
    if (h->precision == 32) return orc::selfIndexGap(h->f.rod, worldOf<float>(*world));
    return orc::selfIndexGap(h->d.rod, worldOf<double>(*world));
}

int orc_count_ops(void* rod, const orc_params* p, const orc_world* world, int steps, long long out[5]) {
    Handle* h = H(rod);
    if (h->precision != 64) return -1;
    orc::Rod<orc::Cnt> r = countingRod(h->d.rod);
    orc::Context<orc::Cnt> ctx;
    const orc::Params<orc::Cnt> pp = paramsOf<orc::Cnt>(*p);
    orc::tally() = orc::OpTally();
    if (world) {
        const orc::World<orc::Cnt> w = worldOf<orc::Cnt>(*world);
        for (int s = 0; s < steps; ++s) orc::step(r, pp, &w, ctx);
    } else {
        for (int s = 0; s < steps; ++s) orc::step(r, pp, static_cast<const orc::World<orc::Cnt>*>(nullptr), ctx);
    }
    const orc::OpTally t = orc::tally();
    out[0] = t.add;
    out[1] = t.mul;
    out[2] = t.div;
    out[3] = t.sqrt;
    out[4] = t.cmp;
    return 0;
}

int orc_colour(void* rod, int family, int* order, int* start, int maxColours) {
    Handle* h = H(rod);
    orc::Colouring c;
    if (h->precision == 32)
        c = family == 0 ? orc::colourStretch(h->f.rod) : orc::colourBend(h->f.rod);
    else
        c = family == 0 ? orc::colourStretch(h->d.rod) : orc::colourBend(h->d.rod);
    for (size_t i = 0; i < c.order.size(); ++i) order[i] = c.order[i];
    const int n = c.colours();
    for (int i = 0; i <= n && i <= maxColours; ++i) start[i] = c.start[i];
    return n;
}

}  // extern "C"

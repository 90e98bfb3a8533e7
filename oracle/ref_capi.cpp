// The oracle's C interface implemented over the REFERENCE's own CPU sources --
// TEST INFRASTRUCTURE. oracle/Makefile compiles this file together with the
// reference checkout's src/core/{rod,coloring,collision,solver}.cpp (read in
// place, never copied) into oracle/_ref/libref64.so and, with CRS_REAL_FLOAT,
// oracle/_ref/libref32.so. Tests load them to pin the restatement in
// rp_oracle.hpp bit for bit. Per-rod compliance/friction scales do not exist in
// the reference CPU path, so anything but 1 is rejected.
#include <thread>
#include <vector>

#include "core/coloring.h"
#include "core/collision.h"
#include "core/geometry.h"
#include "core/rod.h"
#include "core/solver.h"
#include "oracle_capi.h"

using namespace crs;

namespace {

#ifdef CRS_REAL_FLOAT
constexpr int kPrecision = 32;
#else
constexpr int kPrecision = 64;
#endif

struct Handle {
    Rod rod;
    SolverContext ctx;
    Coloring sc, bc;
    bool coloured = false;
};

Handle* H(void* p) { return static_cast<Handle*>(p); }

Vec3 v3(const double* p) { return Vec3(Real(p[0]), Real(p[1]), Real(p[2])); }
Quat q4(const double* p) { return Quat(Real(p[0]), Real(p[1]), Real(p[2]), Real(p[3])); }

RodMaterial matOf(const double m[4]) {
    RodMaterial mat;
    mat.youngs = Real(m[0]);
    mat.poisson = Real(m[1]);
    mat.density = Real(m[2]);
    mat.radius = Real(m[3]);
    return mat;
}

Primitive primOf(const orc_prim& p) {
    Primitive r;
    r.type = p.type;
    r.a = v3(p.a);
    r.b = v3(p.b);
    r.normal = v3(p.normal);
    r.halfExtents = v3(p.half);
    r.rotation = q4(p.rot);
    r.radius = Real(p.radius);
    r.friction = Real(p.friction);
    return r;
}

CollisionWorld worldOf(const orc_world& w) {
    CollisionWorld r;
    for (int i = 0; i < w.numPrims; ++i) r.primitives.push_back(primOf(w.prims[i]));
    r.selfCollision = w.selfCollision != 0;
    r.selfFriction = Real(w.selfFriction);
    r.selfSkip = w.selfSkip;
    r.contactMarginRadii = Real(w.marginRadii);
    r.hashTableSize = w.hashTableSize;
    return r;
}

void putV(const std::vector<Vec3>& s, double* d) {
    if (!d) return;
    for (size_t i = 0; i < s.size(); ++i) {
        d[3 * i] = double(s[i].x);
        d[3 * i + 1] = double(s[i].y);
        d[3 * i + 2] = double(s[i].z);
    }
}
void putQ(const std::vector<Quat>& s, double* d) {
    if (!d) return;
    for (size_t i = 0; i < s.size(); ++i) {
        d[4 * i] = double(s[i].w);
        d[4 * i + 1] = double(s[i].x);
        d[4 * i + 2] = double(s[i].y);
        d[4 * i + 3] = double(s[i].z);
    }
}
void putS(const std::vector<Real>& s, double* d) {
    if (!d) return;
    for (size_t i = 0; i < s.size(); ++i) d[i] = double(s[i]);
}
void putI(const std::vector<int>& s, int* d) {
    if (!d) return;
    for (size_t i = 0; i < s.size(); ++i) d[i] = s[i];
}
void takeV(std::vector<Vec3>& d, const double* s) {
    if (!s) return;
    for (size_t i = 0; i < d.size(); ++i) d[i] = v3(s + 3 * i);
}
void takeQ(std::vector<Quat>& d, const double* s) {
    if (!s) return;
    for (size_t i = 0; i < d.size(); ++i) d[i] = q4(s + 4 * i);
}
void takeS(std::vector<Real>& d, const double* s) {
    if (!s) return;
    for (size_t i = 0; i < d.size(); ++i) d[i] = Real(s[i]);
}
void takeI(std::vector<int>& d, const int* s) {
    if (!s) return;
    for (size_t i = 0; i < d.size(); ++i) d[i] = s[i];
}

}  // namespace

extern "C" {

void* orc_rod_straight(int precision, int n, double length, const double material[4], const double origin[3],
                       const double dir[3]) {
    if (precision != kPrecision) return nullptr;
    Handle* h = new Handle();
    h->rod = makeStraightRod(n, Real(length), matOf(material), v3(origin), v3(dir));
    return h;
}

void* orc_rod_centerline(int precision, int numPoints, const double* points, const double material[4],
                         const double seed[3], double twistPerLength) {
    if (precision != kPrecision) return nullptr;
    std::vector<Vec3> pts(numPoints);
    for (int i = 0; i < numPoints; ++i) pts[i] = v3(points + 3 * i);
    Handle* h = new Handle();
    h->rod = makeRodFromCenterline(pts, matOf(material), v3(seed), Real(twistPerLength));
    return h;
}

int orc_rod_clamp_end(void* rod, int tip) { return clampEndExact(H(rod)->rod, tip != 0); }
void orc_rod_pin(void* rod, int particle) { H(rod)->rod.pinParticle(particle); }
void orc_rod_set_rest_from_current(void* rod) { setRestFromCurrent(H(rod)->rod); }

void orc_rod_sizes(void* rod, int out[4]) {
    const Rod& r = H(rod)->rod;
    out[0] = int(r.state.numParticles());
    out[1] = int(r.state.numSegments());
    out[2] = int(r.stretch.size());
    out[3] = int(r.bend.size());
}

void orc_rod_get(void* rod, orc_rod_view* o) {
    const Rod& r = H(rod)->rod;
    const RodState& s = r.state;
    o->nP = int(s.numParticles());
    o->nF = int(s.numSegments());
    o->nStretch = int(r.stretch.size());
    o->nBend = int(r.bend.size());
    putV(s.x, o->x);
    putV(s.xPrev, o->xPrev);
    putV(s.v, o->v);
    putS(s.invMass, o->invMass);
    putS(s.mass, o->mass);
    putV(s.extForce, o->extForce);
    putV(s.kinematicVelocity, o->kinVel);
    putQ(s.q, o->q);
    putQ(s.qPrev, o->qPrev);
    putV(s.omega, o->omega);
    putV(s.invInertia, o->invInertia);
    putV(s.inertia, o->inertia);
    putV(s.extTorque, o->extTorque);
    putI(r.stretch.p0, o->sP0);
    putI(r.stretch.p1, o->sP1);
    putI(r.stretch.seg, o->sSeg);
    putS(r.stretch.restLength, o->sRest);
    putV(r.stretch.compliance, o->sComp);
    putV(r.stretch.lambda, o->sLam);
    putI(r.bend.segA, o->bA);
    putI(r.bend.segB, o->bB);
    putS(r.bend.restLength, o->bRest);
    putV(r.bend.restDarboux, o->bRestD);
    putV(r.bend.compliance, o->bComp);
    putV(r.bend.lambda, o->bLam);
    o->material[0] = double(r.material.youngs);
    o->material[1] = double(r.material.poisson);
    o->material[2] = double(r.material.density);
    o->material[3] = double(r.material.radius);
}

int orc_rod_set(void* rod, const orc_rod_view* in) {
    Handle* h = H(rod);
    Rod& r = h->rod;
    RodState& s = r.state;
    s.x.resize(in->nP);
    s.xPrev.resize(in->nP);
    s.v.resize(in->nP);
    s.invMass.resize(in->nP);
    s.mass.resize(in->nP);
    s.extForce.resize(in->nP);
    s.kinematicVelocity.resize(in->nP);
    s.q.resize(in->nF);
    s.qPrev.resize(in->nF);
    s.omega.resize(in->nF);
    s.invInertia.resize(in->nF);
    s.inertia.resize(in->nF);
    s.extTorque.resize(in->nF);
    r.stretch.p0.resize(in->nStretch);
    r.stretch.p1.resize(in->nStretch);
    r.stretch.seg.resize(in->nStretch);
    r.stretch.restLength.resize(in->nStretch);
    r.stretch.compliance.resize(in->nStretch);
    r.stretch.lambda.resize(in->nStretch);
    r.bend.segA.resize(in->nBend);
    r.bend.segB.resize(in->nBend);
    r.bend.restLength.resize(in->nBend);
    r.bend.restDarboux.resize(in->nBend);
    r.bend.compliance.resize(in->nBend);
    r.bend.lambda.resize(in->nBend);
    takeV(s.x, in->x);
    takeV(s.xPrev, in->xPrev);
    takeV(s.v, in->v);
    takeS(s.invMass, in->invMass);
    takeS(s.mass, in->mass);
    takeV(s.extForce, in->extForce);
    takeV(s.kinematicVelocity, in->kinVel);
    takeQ(s.q, in->q);
    takeQ(s.qPrev, in->qPrev);
    takeV(s.omega, in->omega);
    takeV(s.invInertia, in->invInertia);
    takeV(s.inertia, in->inertia);
    takeV(s.extTorque, in->extTorque);
    takeI(r.stretch.p0, in->sP0);
    takeI(r.stretch.p1, in->sP1);
    takeI(r.stretch.seg, in->sSeg);
    takeS(r.stretch.restLength, in->sRest);
    takeV(r.stretch.compliance, in->sComp);
    takeV(r.stretch.lambda, in->sLam);
    takeI(r.bend.segA, in->bA);
    takeI(r.bend.segB, in->bB);
    takeS(r.bend.restLength, in->bRest);
    takeV(r.bend.restDarboux, in->bRestD);
    takeV(r.bend.compliance, in->bComp);
    takeV(r.bend.lambda, in->bLam);
    r.material = matOf(in->material);
    h->coloured = false;
    return 0;
}

void orc_rod_free(void* rod) { delete H(rod); }

int orc_step(void* rod, const orc_params* p, const orc_world* world, int steps) {
    if (p->complianceScale != 1.0 || p->frictionScale != 1.0) return -1;
    if (world && world->selfOrder != 0) return -1;  // the reference projects self contacts in list order only
    Handle* h = H(rod);
    SolverParams sp;
    sp.dt = Real(p->dt);
    sp.substeps = p->substeps;
    sp.iterations = p->iterations;
    sp.gravity = v3(p->gravity);
    sp.linearDamping = Real(p->linDamp);
    sp.angularDamping = Real(p->angDamp);
    sp.contactInterval = p->contactInterval;
    if (p->coloured) {
        if (!h->coloured) {
            h->sc = colorStretchConstraints(h->rod);
            h->bc = colorBendConstraints(h->rod);
            h->coloured = true;
        }
        sp.stretchColoring = &h->sc;
        sp.bendColoring = &h->bc;
    }
    if (world) {
        const CollisionWorld w = worldOf(*world);
        for (int s = 0; s < steps; ++s) step(h->rod, sp, w, h->ctx);
    } else {
        for (int s = 0; s < steps; ++s) step(h->rod, sp);
    }
    return 0;
}

void orc_contact_counts(void* rod, int* worldContacts, int* selfContacts) {
    int w = 0, s = 0;
    for (const Contact& c : H(rod)->ctx.contacts.contacts) (c.count == 4 ? s : w) += 1;
    *worldContacts = w;
    *selfContacts = s;
}

int orc_step_many(void** rods, int count, const orc_params* p, const orc_world* const* worlds, int steps,
                  int threads) {
    if (threads < 1) threads = 1;
    std::vector<std::thread> pool;
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
                for (int s = 0; s < steps; ++s) {
                    if (driven[r] >= 0)
                        H(rods[r])->rod.state.kinematicVelocity[driven[r]] =
                            v3(kin_vel + (size_t(r) * steps + s) * 3);
                    orc_step(rods[r], p, worlds ? worlds[r] : nullptr, 1);
                }
            }
        });
    }
    for (std::thread& th : pool) th.join();
    return 0;
}

int orc_chol_solve3(int precision, const double A[9], const double b[3], double y[3]) {
    if (precision != kPrecision) return -1;
    Mat3 M;
    for (int i = 0; i < 9; ++i) M.m[i / 3][i % 3] = Real(A[i]);
    Vec3 out;
    const bool ok = solveSPD3(M, v3(b), out);
    y[0] = double(out.x);
    y[1] = double(out.y);
    y[2] = double(out.z);
    return ok ? 1 : 0;
}

double orc_signed_distance(int precision, const orc_prim* prim, const double p[3], double n[3]) {
    (void)precision;
    Vec3 nn;
    const Real d = signedDistance(primOf(*prim), v3(p), nn);
    n[0] = double(nn.x);
    n[1] = double(nn.y);
    n[2] = double(nn.z);
    return double(d);
}

void orc_seg_seg(int precision, const double p1[3], const double q1[3], const double p2[3], const double q2[3],
                 double st[2]) {
    (void)precision;
    Real s, t;
    closestPointsBetweenSegments(v3(p1), v3(q1), v3(p2), v3(q2), s, t);
    st[0] = double(s);
    st[1] = double(t);
}

unsigned orc_hash_cell(int ix, int iy, int iz, int table) { return hashCell(ix, iy, iz, table); }

int orc_cell_coord(int precision, double v, double cell) {
    (void)precision;
    return cellCoord(Real(v), Real(cell));
}

int orc_self_gap(void* rod, const orc_world* world) { return selfCollisionIndexGap(H(rod)->rod, worldOf(*world)); }

int orc_colour(void* rod, int family, int* order, int* start, int maxColours) {
    const Coloring c = family == 0 ? colorStretchConstraints(H(rod)->rod) : colorBendConstraints(H(rod)->rod);
    for (size_t i = 0; i < c.order.size(); ++i) order[i] = c.order[i];
    const int n = c.numColors();
    for (int i = 0; i <= n && i <= maxColours; ++i) start[i] = c.colorStart[i];
    return n;
}

}  // extern "C"

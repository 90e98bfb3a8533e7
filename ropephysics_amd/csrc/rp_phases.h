// Substep phases shared by the PLAIN and TILED kernels: predict, finish and
// the rod's self-collision (broadphase, pool allocation, sequential
// projection). Arithmetic order follows the reference CPU substep
// (src/core/solver.cpp:218-267) and contact generation
// (src/core/collision.cpp:10-152); see rp_math.h.
#pragma once

#include "rp_layout.h"
#include "rp_math.h"

namespace rp {

__device__ __forceinline__ v3 f4v(float4 a) { return mk(a.x, a.y, a.z); }
__device__ __forceinline__ q4 f4q(float4 a) { return q4{a.x, a.y, a.z, a.w}; }
__device__ __forceinline__ float4 vf4(v3 a, float w) { return make_float4(a.x, a.y, a.z, w); }
__device__ __forceinline__ float4 qf4(q4 a) { return make_float4(a.w, a.x, a.y, a.z); }
__device__ __forceinline__ v3 ld3(const float* p, int i) { return mk(p[3 * i], p[3 * i + 1], p[3 * i + 2]); }
__device__ __forceinline__ void st3(float* p, int i, v3 a) {
    p[3 * i] = a.x;
    p[3 * i + 1] = a.y;
    p[3 * i + 2] = a.z;
}
__device__ __forceinline__ q4 ld4q(const float* p, int i) {
    return q4{p[4 * i], p[4 * i + 1], p[4 * i + 2], p[4 * i + 3]};
}
__device__ __forceinline__ void st4q(float* p, int i, q4 a) {
    p[4 * i] = a.w;
    p[4 * i + 1] = a.x;
    p[4 * i + 2] = a.y;
    p[4 * i + 3] = a.z;
}

// Symplectic-Euler prediction of one particle. x is updated in place and the
// substep-start position returned through xPrev.
__device__ __forceinline__ void predict_particle(v3& x, v3& v, v3& xPrev, float w, v3 force, v3 kinVel, float h,
                                                 v3 g) {
    xPrev = x;
    if (w == 0.0f) {
        x = add(x, scale(kinVel, h));
        return;
    }
    v = add(v, scale(add(g, scale(force, w)), h));
    x = add(x, scale(v, h));
}

// Gyroscopic angular-velocity update and quaternion integration of one frame.
__device__ __forceinline__ void predict_frame(q4& q, v3& omega, q4& qPrev, v3 iI, v3 I, v3 torque, float h) {
    qPrev = q;
    if (norm2(iI) == 0.0f) return;
    const v3 w = omega;
    const v3 tau = rotate_inv(q, torque);
    omega = add(w, scale(cwise(iI, sub(tau, cross(w, cwise(I, w)))), h));
    q = qnormalize(qadd(q, qscale(qmul(q, mkq(0.0f, omega)), 0.5f * h)));
}

__device__ __forceinline__ v3 finish_particle(v3 x, v3 xPrev, float w, v3 kinVel, float h, float decay) {
    if (w == 0.0f) return kinVel;
    return scale(divs(sub(x, xPrev), h), decay);
}

__device__ __forceinline__ v3 finish_frame(q4 q, q4 qPrev, v3 iI, float h, float decay) {
    if (norm2(iI) == 0.0f) return mk(0.0f, 0.0f, 0.0f);
    const q4 dq = hemi_identity(qmul(conj(qPrev), q));
    return scale(scale(im(dq), 2.0f / h), decay);
}

__device__ __forceinline__ prim load_prim(const float* s) {
    prim p;
    p.type = __float_as_int(s[0]);
    p.a = mk(s[1], s[2], s[3]);
    p.b = mk(s[4], s[5], s[6]);
    p.normal = mk(s[7], s[8], s[9]);
    p.half = mk(s[10], s[11], s[12]);
    p.rot = q4{s[13], s[14], s[15], s[16]};
    p.radius = s[17];
    p.friction = s[18];
    return p;
}
__device__ __forceinline__ void stage_prim(float* s, const Prim& g) {
    s[0] = __int_as_float(g.type);
    for (int c = 0; c < 3; ++c) {
        s[1 + c] = g.a[c];
        s[4 + c] = g.b[c];
        s[7 + c] = g.normal[c];
        s[10 + c] = g.half[c];
    }
    for (int c = 0; c < 4; ++c) s[13 + c] = g.rot[c];
    s[17] = g.radius;
    s[18] = g.friction;
    s[19] = 0.0f;
}

// World contacts of one particle: regenerate (or just reset) its P slots.
__device__ __forceinline__ void world_contacts_begin(float* slots, const float* prims, int nPrim, v3 x, float w,
                                                     float radius, float margin, float fscale, bool regen) {
    for (int k = 0; k < nPrim; ++k) {
        float* s = slots + k * kContactFloats;
        s[5] = 0.0f;  // lamN
        s[6] = 0.0f;  // applied
        if (!regen) continue;
        wcontact c;
        world_contact_generate(c, load_prim(prims + k * kPrimFloats), x, w, radius, margin, fscale);
        s[7] = __int_as_float(c.active);
        if (c.active) {
            s[0] = c.n.x;
            s[1] = c.n.y;
            s[2] = c.n.z;
            s[3] = c.offset;
            s[4] = c.friction;
        }
    }
}

__device__ __forceinline__ void world_contacts_project(float* slots, int nPrim, v3& x, v3 xPrev, float w) {
    for (int k = 0; k < nPrim; ++k) {
        float* s = slots + k * kContactFloats;
        wcontact c;
        c.active = __float_as_int(s[7]);
        if (!c.active) continue;
        c.n = mk(s[0], s[1], s[2]);
        c.offset = s[3];
        c.friction = s[4];
        c.lamN = s[5];
        c.applied = s[6];
        world_contact_project(c, x, xPrev, w);
        s[5] = c.lamN;
        s[6] = c.applied;
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
// each control tick, (P(t_{k+1}) - P(t_k)) / period, rounded to FP32.
__device__ inline v3 path_velocity(const PathRec& p, int step) {
    const int tick = step / p.stepsPerControl;
    double x1, y1, x0, y0;
    path_point(p, double(tick + 1) * p.period, x1, y1);
    path_point(p, double(tick) * p.period, x0, y0);
    const double inv = 1.0 / p.period;
    return mk(float((x1 - x0) * inv), float((y1 - y0) * inv), float((p.height - p.height) * inv));
}

// ------------------------------------------------------------------ self-collision
//
// One rod's working set in LDS. Segments are visited in the rod's original
// stretch-constraint order, as the CPU broadphase does.
struct SelfLds {
    float4* cen;     // nS centres and lengths (after a lane-group build: in bucket order)
    int* cellOf;     // nS
    int* sorted;     // nS
    int* segCount;   // nS
    int* segStart;   // nS
    int* cellStart;  // table + 1
    float* pool;     // cap * kSelfFloats
    int* misc;       // [0] live, [1] overflow, [2] cell size, [3] longest segment (float bits)
    int* scan;       // kScanLanes, scratch of the group-parallel scans
};

constexpr int kScanLanes = 256;

__host__ __device__ inline int self_lds_bytes(int nS, int table, int cap) {
    return nS * 16 + 4 * nS * 4 + (table + 1) * 4 + cap * kSelfFloats * 4 + 16 + kScanLanes * 4;
}

__device__ inline SelfLds self_carve(char* base, int nS, int table, int cap) {
    SelfLds s;
    s.cen = reinterpret_cast<float4*>(base);
    s.pool = reinterpret_cast<float*>(base + nS * 16);
    s.misc = reinterpret_cast<int*>(base + nS * 16 + cap * kSelfFloats * 4);
    s.cellOf = s.misc + 4;
    s.sorted = s.cellOf + nS;
    s.segCount = s.sorted + nS;
    s.segStart = s.segCount + nS;
    s.cellStart = s.segStart + nS;
    s.scan = s.cellStart + table + 1;
    return s;
}

// Centres, cell size and the counting sort into buckets; one lane.
__device__ inline void self_build_hash(const SelfLds& s, const float4* x4, const int* segEnd, int nS, int table,
                                       float radius) {
    float maxSeg = 0.0f;
    for (int j = 0; j < nS; ++j) {
        const v3 a = f4v(x4[segEnd[2 * j]]), b = f4v(x4[segEnd[2 * j + 1]]);
        s.cen[j] = vf4(scale(add(a, b), 0.5f), 0.0f);
        const float len = norm(sub(b, a));
        maxSeg = maxSeg < len ? len : maxSeg;
    }
    const float cell = maxSeg + 2.0f * radius;
    s.misc[2] = __float_as_int(cell);
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
__device__ __forceinline__ void self_pair(const SelfLds& s, const float4* x4, int j0, int j1, v3 a0, v3 a1, int k0,
                                          int k1, float radius, float friction, bool fill, int slotBase, int cap,
                                          int& count) {
    const v3 b0 = f4v(x4[k0]), b1 = f4v(x4[k1]);
    float u, v;
    seg_seg(a0, a1, b0, b1, u, v);
    const v3 pa = add(a0, scale(sub(a1, a0), u));
    const v3 pb = add(b0, scale(sub(b1, b0), v));
    const v3 delta = sub(pa, pb);
    const float dist = norm(delta);
    const float sep = dist - 2.0f * radius;
    if (sep >= 0.0f) return;
    const int slot = slotBase + count++;
    if (!fill || slot >= cap) return;
    const v3 n = dist > 1e-12f ? divs(delta, dist) : mk(0.0f, 0.0f, 1.0f);
    const int idx[4] = {j0, j1, k0, k1};
    const float wt[4] = {1.0f - u, u, -(1.0f - v), -v};
    float dotSum = 0.0f;
    for (int m = 0; m < 4; ++m) dotSum += wt[m] * dot(f4v(x4[idx[m]]), n);
    float* c = s.pool + slot * kSelfFloats;
    for (int m = 0; m < 4; ++m) {
        c[m] = __int_as_float(idx[m]);
        c[4 + m] = wt[m];
    }
    c[8] = n.x;
    c[9] = n.y;
    c[10] = n.z;
    c[11] = dotSum - sep;  // offset
    c[12] = friction;
    c[13] = 0.0f;  // lamN
    c[14] = 0.0f;  // applied
}

// Contacts of segment j with later segments in the CPU's candidate order
// (27 cells dz, dy, dx; bucket items ascending; a bucket reached twice is
// visited twice). Counts them, or with `fill` writes them at j's pool offset.
__device__ inline void self_query(const SelfLds& s, const float4* x4, const int* segEnd, int nS, int table,
                                  int gap, int cap, float radius, float friction, int j, bool fill) {
    int count = 0;
    if (nS > gap) {
        const float cell = __int_as_float(s.misc[2]);
        const v3 cj = f4v(s.cen[j]);
        const int cx = cell_coord(cj.x, cell), cy = cell_coord(cj.y, cell), cz = cell_coord(cj.z, cell);
        const int j0 = segEnd[2 * j], j1 = segEnd[2 * j + 1];
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
                        self_pair(s, x4, j0, j1, a0, a1, segEnd[2 * k], segEnd[2 * k + 1], radius, friction, fill,
                                  slotBase, cap, count);
                    }
                }
    }
    if (!fill) s.segCount[j] = count;
}

// ------------------------------------------------------------------ group-parallel broadphase
//
// The same hash, candidate order and pool layout as self_build_hash /
// self_allocate, built by the G lanes of a rod's lane group: the longest
// segment by an atomic max on the (non-negative) float bits, bucket counts by
// atomic adds, prefix sums by a chunked scan, and each bucket's items placed
// in ascending segment order by counting the earlier segments in the same
// bucket. Every step is exact integer work or an order-free max, so the
// tables equal the serial build's. Every lane of the workgroup calls these
// (they contain barriers); `on` is false for lanes whose rod has no
// self-collision.

// Exclusive prefix sum of in[0, n) into out[0, n) by the first
// min(G, kScanLanes) lanes of the group (each summing a chunk); returns the
// total. Fixed barrier count whatever G, n or `on`; lanes past kScanLanes only
// take part in the barriers.
__device__ inline int group_scan(const int* in, int* out, int n, int* scratch, int l, int G, bool on) {
    const int P = min(G, kScanLanes);
    const bool part = on && l < P;
    const int chunk = (n + P - 1) / max(P, 1);
    const int lo = part ? min(n, l * chunk) : 0;
    const int hi = part ? min(n, lo + chunk) : 0;
    int sum = 0;
    for (int i = lo; i < hi; ++i) sum += in[i];
    if (part) scratch[l] = sum;
    __syncthreads();
    for (int off = 1; off < kScanLanes; off <<= 1) {
        const int v = (part && l >= off) ? scratch[l - off] : 0;
        __syncthreads();
        if (part) scratch[l] += v;
        __syncthreads();
    }
    const int total = on ? scratch[P - 1] : 0;
    int run = part ? scratch[l] - sum : 0;
    for (int i = lo; i < hi; ++i) {
        const int v = in[i];
        out[i] = run;
        run += v;
    }
    __syncthreads();
    return total;
}

__device__ __forceinline__ float abs_max3(v3 a) { return fmaxf(fabsf(a.x), fmaxf(fabsf(a.y), fabsf(a.z))); }
__device__ __forceinline__ int count_eq4(int4 c, int h) {
    return (c.x == h ? 1 : 0) + (c.y == h ? 1 : 0) + (c.z == h ? 1 : 0) + (c.w == h ? 1 : 0);
}

// self_query for the lane group: segment j's centre and length (cj4, held by
// the lane) and candidates' centres stored in bucket order (s.cen[e] belongs
// to s.sorted[e]), so a candidate costs one LDS round trip; the three buckets
// of a (dz, dy) row are fetched together. Same candidates in the same order as
// self_query, with one addition: a candidate whose centre is farther from j's
// than the two half-lengths plus 2r (with a relative margin of 1e-3 and an
// absolute one of 1e-6 times the coordinate scale, far above the FP32
// rounding of centres and lengths) is dropped before the closest-point test.
// Every point of a segment lies within half its length of its centre, so such
// a pair's computed separation is positive and it would not be a contact: the
// contact list is unchanged. Returns the number of contacts.
template <bool kFill>
__device__ inline int self_query_rows(const SelfLds& s, const float4* x4, const int* segEnd, int table, int gap,
                                      int cap, float radius, float friction, float cell, int j, float4 cj4,
                                      int slotBase) {
    const v3 cj = f4v(cj4);
    const float reachJ = 0.5f * cj4.w + 2.0f * radius;
    const float scaleJ = abs_max3(cj);
    const int cx = cell_coord(cj.x, cell), cy = cell_coord(cj.y, cell), cz = cell_coord(cj.z, cell);
    const int j0 = segEnd[2 * j], j1 = segEnd[2 * j + 1];
    const v3 a0 = f4v(x4[j0]), a1 = f4v(x4[j1]);
    const unsigned tbl = static_cast<unsigned>(table);
    // Usually every lane group of the wave has a power-of-two table: then the
    // modulo path is not issued at all.
    const bool pow2 = __all((tbl & (tbl - 1u)) == 0u) != 0;
    int count = 0;
    for (int dz = -1; dz <= 1; ++dz)
        for (int dy = -1; dy <= 1; ++dy) {
            // cell_hash's XOR of the three products, the x term varying fastest.
            const unsigned hzy =
                (static_cast<unsigned>(cz + dz) * 83492791u) ^ (static_cast<unsigned>(cy + dy) * 19349663u);
            int lo[3], hi[3];
#pragma unroll
            for (int dx = 0; dx < 3; ++dx) {
                const unsigned hh = (static_cast<unsigned>(cx + dx - 1) * 73856093u) ^ hzy;
                const int h = static_cast<int>(pow2 ? (hh & (tbl - 1u)) : hh % tbl);
                lo[dx] = s.cellStart[h];
                hi[dx] = s.cellStart[h + 1];
            }
            // The row's candidates, buckets dx = -1, 0, 1 concatenated, four
            // at a time: their loads go out together (indices past the end
            // repeat the last candidate, so every load is in range and none
            // waits on a test), the culls run side by side without branches,
            // and survivors reach the narrow phase in candidate order.
            const int n0 = hi[0] - lo[0], n01 = n0 + hi[1] - lo[1], total = n01 + hi[2] - lo[2];
            for (int base = 0; base < total; base += 4) {
                int kq[4];
                float4 cq[4];
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    const int q = min(base + u, total - 1);
                    const int e = q < n0 ? lo[0] + q : (q < n01 ? lo[1] + (q - n0) : lo[2] + (q - n01));
                    kq[u] = s.sorted[e];
                    cq[u] = s.cen[e];
                }
                unsigned pass = 0u;
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    const v3 ck = f4v(cq[u]);
                    const float lim =
                        1.001f * (reachJ + 0.5f * cq[u].w) + 1e-6f * (1.0f + fmaxf(scaleJ, abs_max3(ck)));
                    const bool keep = (base + u < total) & (kq[u] - j >= gap) & !(norm2(sub(ck, cj)) > lim * lim);
                    pass |= keep ? (1u << u) : 0u;
                }
                while (pass != 0u) {
                    const int u = __builtin_ctz(pass);
                    pass &= pass - 1u;
                    const int k = u == 0 ? kq[0] : u == 1 ? kq[1] : u == 2 ? kq[2] : kq[3];
                    self_pair(s, x4, j0, j1, a0, a1, segEnd[2 * k], segEnd[2 * k + 1], radius, friction, kFill,
                              slotBase, cap, count);
                }
            }
        }
    return count;
}

// Broadphase and contact generation by the G lanes of a rod's lane group; lane
// l owns segments j = l + G t, t < TP (the host guarantees nS <= TP G), and
// keeps their centres in registers. s.cen ends up in bucket order.
template <int TP>
__device__ inline void self_regen_group(const SelfLds& s, const float4* x4, const int* segEnd, int nS, int table,
                                        int gap, int cap, float radius, float friction, bool on, int l, int G) {
    float4 cj[TP];
    if (on) {
        for (int t = l; t <= table; t += G) s.cellStart[t] = 0;
        if (l == 0) s.misc[3] = 0;
    }
    __syncthreads();
#pragma unroll
    for (int t = 0; t < TP; ++t) {
        const int j = l + G * t;
        cj[t] = make_float4(0.f, 0.f, 0.f, 0.f);
        if (on && j < nS) {
            const v3 a = f4v(x4[segEnd[2 * j]]), b = f4v(x4[segEnd[2 * j + 1]]);
            const float len = norm(sub(b, a));
            cj[t] = vf4(scale(add(a, b), 0.5f), len);
            atomicMax(&s.misc[3], __float_as_int(len));
        }
    }
    __syncthreads();
    const float cell = on ? __int_as_float(s.misc[3]) + 2.0f * radius : 0.0f;
#pragma unroll
    for (int t = 0; t < TP; ++t) {
        const int j = l + G * t;
        if (on && j < nS) {
            const v3 c = f4v(cj[t]);
            const int h = int(cell_hash(cell_coord(c.x, cell), cell_coord(c.y, cell), cell_coord(c.z, cell), table));
            s.cellOf[j] = h;
            atomicAdd(&s.cellStart[h], 1);
        }
    }
    if (on && l == 0) s.misc[2] = __float_as_int(cell);
    __syncthreads();
    const int items = group_scan(s.cellStart, s.cellStart, on ? table : 0, s.scan, l, G, on);
    if (on && l == 0) s.cellStart[table] = items;
    // Stable placement: j goes after the earlier segments of its bucket.
    // cellOf is 16-byte aligned, so four buckets per LDS read.
    const int4* cellOf4 = reinterpret_cast<const int4*>(s.cellOf);
#pragma unroll
    for (int t = 0; t < TP; ++t) {
        const int j = l + G * t;
        if (on && j < nS) {
            const int h = s.cellOf[j];
            int rank = 0, i = 0;
            for (; i + 16 <= j; i += 16) {
                const int4 c0 = cellOf4[(i >> 2) + 0], c1 = cellOf4[(i >> 2) + 1];
                const int4 c2 = cellOf4[(i >> 2) + 2], c3 = cellOf4[(i >> 2) + 3];
                rank += count_eq4(c0, h) + count_eq4(c1, h) + count_eq4(c2, h) + count_eq4(c3, h);
            }
            for (; i + 4 <= j; i += 4) rank += count_eq4(cellOf4[i >> 2], h);
            for (; i < j; ++i) rank += s.cellOf[i] == h ? 1 : 0;
            const int e = s.cellStart[h] + rank;
            s.sorted[e] = j;
            s.cen[e] = cj[t];
        }
    }
    __syncthreads();
#pragma unroll
    for (int t = 0; t < TP; ++t) {
        const int j = l + G * t;
        if (on && j < nS)
            s.segCount[j] = nS > gap ? self_query_rows<false>(s, x4, segEnd, table, gap, cap, radius, friction,
                                                              cell, j, cj[t], 0)
                                     : 0;
    }
    __syncthreads();
    const int total = group_scan(s.segCount, s.segStart, on ? nS : 0, s.scan, l, G, on);
    if (on && l == 0) {
        s.misc[0] = total < cap ? total : cap;
        if (total > cap) s.misc[1] += total - cap;
    }
    __syncthreads();
#pragma unroll
    for (int t = 0; t < TP; ++t) {
        const int j = l + G * t;
        if (on && j < nS && s.segCount[j] > 0)
            self_query_rows<true>(s, x4, segEnd, table, gap, cap, radius, friction, cell, j, cj[t], s.segStart[j]);
    }
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
        s.pool[c * kSelfFloats + 13] = 0.0f;
        s.pool[c * kSelfFloats + 14] = 0.0f;
    }
}

// Four-particle contacts, sequentially in pool order (= the CPU's list order).
// xp4 holds substep-start positions.
__device__ inline void self_project(const SelfLds& s, float4* x4, const float4* xp4) {
    const int live = s.misc[0];
    for (int ci = 0; ci < live; ++ci) {
        float* c = s.pool + ci * kSelfFloats;
        int idx[4];
        float wt[4], im[4];
        for (int m = 0; m < 4; ++m) {
            idx[m] = __float_as_int(c[m]);
            wt[m] = c[4 + m];
            im[m] = x4[idx[m]].w;
        }
        const v3 n = mk(c[8], c[9], c[10]);
        float wEff = 0.0f;
        for (int m = 0; m < 4; ++m) wEff += wt[m] * wt[m] * im[m];
        if (wEff <= 0.0f) continue;
        float C = -c[11];
        for (int m = 0; m < 4; ++m) C += wt[m] * dot(f4v(x4[idx[m]]), n);
        float dl = -C / wEff;
        const float lamN0 = c[13];
        const float cl = max0(lamN0 + dl);
        dl = cl - lamN0;
        c[13] = cl;
        if (dl != 0.0f)
            for (int m = 0; m < 4; ++m) {
                const v3 xm = add(f4v(x4[idx[m]]), scale(n, im[m] * wt[m] * dl));
                x4[idx[m]] = vf4(xm, im[m]);
            }
        const float friction = c[12];
        if (friction <= 0.0f || cl <= 0.0f) continue;
        v3 dp = mk(0.0f, 0.0f, 0.0f);
        for (int m = 0; m < 4; ++m) dp = add(dp, scale(sub(f4v(x4[idx[m]]), f4v(xp4[idx[m]])), wt[m]));
        const v3 tang = sub(dp, scale(n, dot(dp, n)));
        const float slide = norm(tang);
        if (slide < 1e-15f) continue;
        const float budget = friction * cl * wEff - c[14];
        if (budget <= 0.0f) continue;
        const float capped = min_ab(slide, budget);
        c[14] = c[14] + capped;
        const v3 corr = scale(tang, -capped / slide);
        for (int m = 0; m < 4; ++m) {
            const v3 xm = add(f4v(x4[idx[m]]), scale(corr, im[m] * wt[m] / wEff));
            x4[idx[m]] = vf4(xm, im[m]);
        }
    }
}

}  // namespace rp

// Host-side pieces shared by the batch and the serving loop: errors, the
// tile-shape table, record building (validation, colouring, layout), LDS
// limits, tile packing and step arguments.
//
// Record building follows the reference's Batch::create
// (src/gpu/gpu_batch.cpp:150-316) generalised to rods of any topology:
// greedy colouring per rod in constraint index order
// (src/core/coloring.cpp:11-57), constraints stored in colour order, and the
// self-collision index gap and pool size as the reference forms them.
#include "rp_host.h"

#include <algorithm>
#include <array>
#include <cmath>
#include <cstring>
#include <map>
#include <string>
#include <vector>

namespace rp {
namespace host {

namespace {
thread_local std::string g_error;
}  // namespace

int fail(int code, const std::string& msg) {
    g_error = msg;
    return code;
}

const char* last_error() { return g_error.c_str(); }

// ------------------------------------------------------------------ tile configurations

namespace {
// id = index + 1. Every (tp, tc) here is instantiated in rp_kernels.hip.
const TileConfig kConfigs[] = {
    {2, 1, 1, "tp2-tc1-w1"}, {3, 1, 1, "tp3-tc1-w1"}, {3, 2, 1, "tp3-tc2-w1"}, {5, 2, 1, "tp5-tc2-w1"},
    {2, 1, 2, "tp2-tc1-w2"}, {3, 1, 2, "tp3-tc1-w2"}, {3, 2, 2, "tp3-tc2-w2"}, {5, 2, 2, "tp5-tc2-w2"},
    {2, 1, 4, "tp2-tc1-w4"}, {3, 1, 4, "tp3-tc1-w4"}, {1, 1, 2, "tp1-tc1-w2"}, {1, 1, 4, "tp1-tc1-w4"},
    {1, 1, 8, "tp1-tc1-w8"},
};
constexpr int kNumConfigs = int(sizeof(kConfigs) / sizeof(kConfigs[0]));

// Greedy colouring, lowest free colour in constraint index order: the
// reference's colorConstraints (src/core/coloring.cpp:11-57). `res` lists the
// state ids each constraint writes.
int greedy_colours(int count, const std::vector<std::array<int, 3>>& res, int width, int numIds,
                   std::vector<int>& order, std::vector<int>& start) {
    std::vector<std::vector<char>> used(numIds);
    std::vector<int> colOf(count);
    int nCol = 0;
    for (int k = 0; k < count; ++k) {
        int c = 0;
        for (;; ++c) {
            bool clash = false;
            for (int m = 0; m < width; ++m) {
                const std::vector<char>& u = used[res[k][m]];
                if (c < int(u.size()) && u[c]) {
                    clash = true;
                    break;
                }
            }
            if (!clash) break;
        }
        colOf[k] = c;
        nCol = std::max(nCol, c + 1);
        for (int m = 0; m < width; ++m) {
            std::vector<char>& u = used[res[k][m]];
            if (int(u.size()) <= c) u.resize(c + 1, 0);
            u[c] = 1;
        }
    }
    start.assign(nCol + 1, 0);
    for (int k = 0; k < count; ++k) ++start[colOf[k] + 1];
    for (int c = 0; c < nCol; ++c) start[c + 1] += start[c];
    order.resize(count);
    std::vector<int> cur(start.begin(), start.end() - 1);
    for (int k = 0; k < count; ++k) order[cur[colOf[k]]++] = k;
    return nCol;
}

}  // namespace

int num_tile_configs() { return kNumConfigs; }

const TileConfig& tile_config(int id) { return kConfigs[id - 1]; }

template <typename Real>
bool tile_config_available(int id) {
    if (id < 1 || id > kNumConfigs) return false;
    const TileConfig& c = kConfigs[id - 1];
    return Prec<Real>::tiled_shape_available(c.tp, c.tc, 64 * c.waves);
}

// Lanes a rod needs under (tp, tc): every particle and frame within tp trips,
// every colour within tc trips, and -- with self-collision, whose broadphase
// gives each lane the segments l, l + G, ... -- every stretch constraint
// within tp trips.
template <typename Real>
int group_lanes(const RodInfoT<Real>& r, const std::vector<int>& colStart, int tp, int tc) {
    int g = std::max((r.nP + tp - 1) / tp, (r.nF + tp - 1) / tp);
    if (r.self) g = std::max(g, (r.nS + tp - 1) / tp);
    const int* cs = colStart.data() + r.colOff;
    for (int c = 0; c < r.nSCol; ++c) g = std::max(g, (cs[c + 1] - cs[c] + tc - 1) / tc);
    const int* cb = cs + r.nSCol + 1;
    for (int c = 0; c < r.nBCol; ++c) g = std::max(g, (cb[c + 1] - cb[c] + tc - 1) / tc);
    return std::max(g, 1);
}

template <typename Real>
int rod_tiled_lds(const RodInfoT<Real>& r) {
    return Prec<Real>::tiled_lds_bytes(r.nP, r.nF, r.nS, r.nPrim, r.self, r.table, r.selfCap);
}

int lds_limits(LdsLimits& out) {
    hipDeviceProp_t prop;
    RP_HIP(hipGetDeviceProperties(&prop, 0));
    // Workgroups may use the whole LDS of a compute unit; where the runtime
    // refuses the larger dynamic allocation, stay within 64 KiB.
    out.perCU = prop.sharedMemPerBlock;
    out.computeUnits = prop.multiProcessorCount;
    int granted32 = 0, granted64 = 0;
    if (Prec<float>::allow_dynamic_lds(int(prop.sharedMemPerBlock), &granted32) == hipSuccess &&
        Prec<double>::allow_dynamic_lds(int(prop.sharedMemPerBlock), &granted64) == hipSuccess) {
        out.perWorkgroup = size_t(std::min(granted32, granted64));
    } else {
        (void)hipGetLastError();
        out.perCU = out.perWorkgroup = 65536 - 1024;
    }
    return RP_OK;
}

// A 512-byte allocation granule is left for the static part.
int lds_budget(const LdsLimits& l, int perCU) {
    return std::min(int(l.perWorkgroup), (int(l.perCU / size_t(perCU)) & ~511) - 512);
}

// Best fit over open workgroups, rods taken widest first; a workgroup closes
// when neither lanes nor LDS are left for the narrowest remaining rod.
template <typename Real>
int pack_rods(const std::vector<RodInfoT<Real>>& rods, const std::vector<int>& ids, const std::vector<int>& colStart,
              int configId, int ldsCap, Packing& out) {
    const TileConfig& cfg = tile_config(configId);
    const int laneCap = 64 * cfg.waves;
    const int n = int(ids.size());
    std::vector<int> lanes(rods.size()), lds(rods.size()), order(n);
    for (int i = 0; i < n; ++i) {
        const int r = ids[i];
        const RodInfoT<Real>& ri = rods[r];
        if (ri.nSCol > kTiledColours || ri.nBCol > kTiledColours) return RP_ERR_CAPACITY;
        lanes[r] = group_lanes(ri, colStart, cfg.tp, cfg.tc);
        lds[r] = rod_tiled_lds(ri);
        if (lanes[r] > laneCap || lds[r] > ldsCap) return RP_ERR_CAPACITY;
        order[i] = r;
    }
    std::stable_sort(order.begin(), order.end(), [&](int a, int c) { return lanes[a] > lanes[c]; });
    struct Open {
        int lanesUsed, ldsUsed;
        std::vector<int> rods;
    };
    std::vector<Open> bins;
    // Open bins keyed by free lanes, for best fit.
    std::multimap<int, int> byFree;
    for (int r : order) {
        int chosen = -1;
        for (auto it = byFree.lower_bound(lanes[r]); it != byFree.end(); ++it) {
            Open& o = bins[it->second];
            if (o.ldsUsed + lds[r] <= ldsCap) {
                chosen = it->second;
                byFree.erase(it);
                break;
            }
        }
        if (chosen < 0) {
            bins.push_back(Open{0, 0, {}});
            chosen = int(bins.size()) - 1;
        }
        Open& o = bins[chosen];
        o.rods.push_back(r);
        o.lanesUsed += lanes[r];
        o.ldsUsed += lds[r];
        if (laneCap - o.lanesUsed > 0) byFree.emplace(laneCap - o.lanesUsed, chosen);
    }
    out.config = configId;
    out.tiles.clear();
    out.slots.clear();
    out.maxLds = 0;
    out.lanesBusy = 0;
    for (const Open& o : bins) {
        Tile t;
        t.firstSlot = int(out.slots.size());
        t.numSlots = int(o.rods.size());
        // A lane group starts on a wavefront boundary when the lanes left in
        // the workgroup allow it: a colour pass then spans fewer wavefronts.
        int lane = 0, off = 0, rest = o.lanesUsed;
        for (int r : o.rods) {
            const int aligned = (lane + 63) & ~63;
            if (aligned + rest <= laneCap) lane = aligned;
            rest -= lanes[r];
            out.slots.push_back(TileSlot{r, lane, lanes[r], off});
            lane += lanes[r];
            off += lds[r];
        }
        t.ldsBytes = off;
        t.pad = 0;
        out.maxLds = std::max(out.maxLds, off);
        out.lanesBusy += o.lanesUsed;
        out.tiles.push_back(t);
    }
    return RP_OK;
}

// A launch of n tiles on `units` compute units with units < n <= 2 units, two
// workgroups fitting on a unit, runs workgroups b and b + units on one unit
// for b < n - units, and workgroups n - units .. units - 1 alone. The launch lasts as long as its slowest unit.
// A tile's cost is estimated as the sum over its rods of lanes squared: a
// rod's self-contact search (the candidate scan over the hash cells its
// segments reach, and the lists it keeps) grows with the square of its length
// once it folds, while the constraint passes grow with its lanes. The
// costliest tiles go to the units they hold alone and the rest pair costliest
// with cheapest. Measured on 1024 routing cables (490 tiles of 256 lanes on
// 256 units): 683-688 ms in packing order, 657-659 ms placed. Only the order of
// the tiles changes, and a rod's arithmetic does not depend on its tile's
// index, so every result is the same bit for bit.
void place_tiles(Packing& p, int units) {
    const int n = int(p.tiles.size());
    if (units <= 0 || n <= units || n > 2 * units) return;
    const int pairs = n - units;
    std::vector<long long> cost(size_t(n), 0);
    std::vector<int> byCost(static_cast<size_t>(n));
    for (int i = 0; i < n; ++i) {
        const Tile& t = p.tiles[size_t(i)];
        for (int k = 0; k < t.numSlots; ++k) {
            const long long l = p.slots[size_t(t.firstSlot + k)].lanes;
            cost[size_t(i)] += l * l;
        }
        byCost[size_t(i)] = i;
    }
    std::stable_sort(byCost.begin(), byCost.end(), [&](int a, int c) { return cost[size_t(a)] > cost[size_t(c)]; });
    std::vector<Tile> placed(static_cast<size_t>(n));
    int next = 0;
    for (int b = pairs; b < units; ++b) placed[size_t(b)] = p.tiles[size_t(byCost[size_t(next++)])];
    for (int b = 0; b < pairs; ++b) placed[size_t(b)] = p.tiles[size_t(byCost[size_t(next++)])];
    for (int b = pairs - 1; b >= 0; --b) placed[size_t(units + b)] = p.tiles[size_t(byCost[size_t(next++)])];
    p.tiles.swap(placed);
}

// h = dt / substeps and the velocity decays exp(-damping h) in the working
// precision, as the reference's step forms them (src/core/solver.cpp:218-251).
template <typename Real>
StepArgsT<Real> step_args(const typename Prec<Real>::StepParams& p, int steps) {
    StepArgsT<Real> a;
    a.h = p.dt / Real(p.substeps);
    a.invH2 = Real(1) / (a.h * a.h);
    a.gx = p.gravity[0];
    a.gy = p.gravity[1];
    a.gz = p.gravity[2];
    a.linDecay = std::exp(-p.linear_damping * a.h);
    a.angDecay = std::exp(-p.angular_damping * a.h);
    a.substeps = p.substeps;
    a.iterations = p.iterations;
    a.contactInterval = p.contact_interval > 0 ? p.contact_interval : 1;
    a.steps = steps;
    a.sliceTicks = 0;
    a.progress = nullptr;
    a.pairUnits = 0;
    return a;
}

template <typename Real>
int build_records(const typename Prec<Real>::BatchDesc* desc, RecordsT<Real>& out) {
    using Vec4 = typename Prec<Real>::Vec4;
    const auto make4 = &Prec<Real>::make4;
    if (!desc) return fail(RP_ERR_INVALID, "null argument");
    const int R = desc->num_rods;
    if (R <= 0) return fail(RP_ERR_INVALID, "num_rods must be positive");
    if (!desc->particle_start || !desc->frame_start || !desc->stretch_start || !desc->bend_start)
        return fail(RP_ERR_INVALID, "missing CSR start arrays");
    out.pStart.assign(desc->particle_start, desc->particle_start + R + 1);
    out.fStart.assign(desc->frame_start, desc->frame_start + R + 1);
    out.sStart.assign(desc->stretch_start, desc->stretch_start + R + 1);
    out.bStart.assign(desc->bend_start, desc->bend_start + R + 1);
    for (const std::vector<int>* v : {&out.pStart, &out.fStart, &out.sStart, &out.bStart}) {
        if ((*v)[0] != 0) return fail(RP_ERR_INVALID, "CSR start arrays must begin at 0");
        for (int r = 0; r < R; ++r)
            if ((*v)[r + 1] < (*v)[r]) return fail(RP_ERR_INVALID, "CSR start arrays must be non-decreasing");
    }
    const int P = out.pStart[R], F = out.fStart[R], S = out.sStart[R], B = out.bStart[R];
    out.R = R;
    out.P = P;
    out.F = F;
    out.S = S;
    out.B = B;
    if (P > 0 && (!desc->inv_mass)) return fail(RP_ERR_INVALID, "missing inv_mass");
    if (F > 0 && (!desc->inv_inertia || !desc->inertia)) return fail(RP_ERR_INVALID, "missing inertia");
    if (S > 0 && (!desc->stretch_p0 || !desc->stretch_p1 || !desc->stretch_seg || !desc->stretch_rest ||
                  !desc->stretch_compliance))
        return fail(RP_ERR_INVALID, "missing stretch constraint arrays");
    if (B > 0 && (!desc->bend_a || !desc->bend_b || !desc->bend_rest || !desc->bend_compliance ||
                  !desc->bend_rest_darboux))
        return fail(RP_ERR_INVALID, "missing bend constraint arrays");
    if (!desc->radius) return fail(RP_ERR_INVALID, "missing radius");
    if (desc->num_worlds < 0 || (desc->num_worlds > 0 && !desc->worlds))
        return fail(RP_ERR_INVALID, "bad world table");
    for (int w = 0; w < desc->num_worlds; ++w) {
        const auto& wd = desc->worlds[w];
        if (wd.num_prims < 0 || wd.num_prims > kMaxPrims || (wd.num_prims > 0 && !wd.prims))
            return fail(RP_ERR_INVALID, "world " + std::to_string(w) + ": 0.." + std::to_string(kMaxPrims) +
                                            " primitives supported");
        for (int k = 0; k < wd.num_prims; ++k)
            if (wd.prims[k].type < RP_PLANE || wd.prims[k].type > RP_BOX)
                return fail(RP_ERR_INVALID, "world " + std::to_string(w) + ": unknown primitive type");
        if (wd.self_collision && wd.hash_table_size < 1)
            return fail(RP_ERR_INVALID, "world " + std::to_string(w) + ": hash_table_size must be positive");
        if (wd.self_capacity < 0 || !(wd.self_capacity_per_segment >= 0) || wd.self_capacity_per_segment > 64)
            return fail(RP_ERR_INVALID, "world " + std::to_string(w) + ": self capacity out of range");
        if (wd.self_order != 0 && wd.self_order != 1)
            return fail(RP_ERR_INVALID, "world " + std::to_string(w) + ": self_order is 0 (list) or 1 (coloured)");
    }

    // ---- per-rod colouring, constraint reordering and rod info
    std::vector<int4>& sIdx = out.sIdx;
    std::vector<int4>& bIdx = out.bIdx;
    std::vector<Vec4>& sPar = out.sPar;
    std::vector<Vec4>& bPar = out.bPar;
    std::vector<Vec4>& bRest4 = out.bRest4;
    std::vector<int>& segEnd = out.segEnd;
    sIdx.assign(S, int4{});
    bIdx.assign(B, int4{});
    sPar.assign(S, Vec4{});
    bPar.assign(B, Vec4{});
    bRest4.assign(B, Vec4{});
    segEnd.assign(2 * size_t(S), 0);
    std::vector<int>& colStart = out.colStart;
    colStart.clear();
    out.rods.resize(R);
    out.colourFit.assign(R, 1);
    std::vector<PrimT<Real>>& prims = out.prims;
    std::vector<int> primOffOfWorld(std::max(desc->num_worlds, 1), 0);
    for (int w = 0; w < desc->num_worlds; ++w) {
        primOffOfWorld[w] = int(prims.size());
        for (int k = 0; k < desc->worlds[w].num_prims; ++k) {
            const auto& s = desc->worlds[w].prims[k];
            PrimT<Real> p;
            p.type = s.type;
            for (int c = 0; c < 3; ++c) {
                p.a[c] = s.a[c];
                p.b[c] = s.b[c];
                p.normal[c] = s.normal[c];
                p.half[c] = s.half_extents[c];
            }
            for (int c = 0; c < 4; ++c) p.rot[c] = s.rotation[c];
            p.radius = s.radius;
            p.friction = s.friction;
            prims.push_back(p);
        }
    }
    std::vector<int> order, start;
    std::vector<std::array<int, 3>> res;
    int maxPlainLds = 0;
    for (int r = 0; r < R; ++r) {
        RodInfoT<Real>& ri = out.rods[r];
        std::memset(&ri, 0, sizeof(ri));
        ri.pStart = out.pStart[r];
        ri.nP = out.pStart[r + 1] - out.pStart[r];
        ri.fStart = out.fStart[r];
        ri.nF = out.fStart[r + 1] - out.fStart[r];
        ri.sStart = out.sStart[r];
        ri.nS = out.sStart[r + 1] - out.sStart[r];
        ri.bStart = out.bStart[r];
        ri.nB = out.bStart[r + 1] - out.bStart[r];
        if (ri.nP < 1) return fail(RP_ERR_INVALID, "rod " + std::to_string(r) + " has no particles");
        ri.radius = desc->radius[r];
        const int w = desc->num_worlds == 0 ? -1 : (desc->world_of_rod ? desc->world_of_rod[r] : 0);
        if (w >= desc->num_worlds) return fail(RP_ERR_INVALID, "rod " + std::to_string(r) + ": world out of range");
        ri.world = w;
        // Stretch: validate, colour, reorder.
        res.resize(ri.nS);
        for (int k = 0; k < ri.nS; ++k) {
            const int g = ri.sStart + k;
            const int a = desc->stretch_p0[g], c = desc->stretch_p1[g], f = desc->stretch_seg[g];
            if (a < 0 || a >= ri.nP || c < 0 || c >= ri.nP || f < 0 || f >= ri.nF)
                return fail(RP_ERR_INVALID, "rod " + std::to_string(r) + ": stretch constraint " +
                                                std::to_string(k) + " index out of range");
            res[k] = {a, c, ri.nP + f};
            segEnd[2 * size_t(g)] = a;
            segEnd[2 * size_t(g) + 1] = c;
        }
        ri.colOff = int(colStart.size());
        ri.nSCol = greedy_colours(ri.nS, res, 3, ri.nP + ri.nF, order, start);
        for (int k = 0; k < ri.nS; ++k) {
            const int src = ri.sStart + order[k], dst = ri.sStart + k;
            sIdx[dst] = make_int4(desc->stretch_p0[src], desc->stretch_p1[src], desc->stretch_seg[src], order[k]);
            sPar[dst] = make4(desc->stretch_rest[src], desc->stretch_compliance[3 * size_t(src)],
                                    desc->stretch_compliance[3 * size_t(src) + 1],
                                    desc->stretch_compliance[3 * size_t(src) + 2]);
        }
        colStart.insert(colStart.end(), start.begin(), start.end());
        // Bend.
        res.resize(ri.nB);
        for (int k = 0; k < ri.nB; ++k) {
            const int g = ri.bStart + k;
            const int a = desc->bend_a[g], c = desc->bend_b[g];
            if (a < 0 || a >= ri.nF || c < 0 || c >= ri.nF)
                return fail(RP_ERR_INVALID, "rod " + std::to_string(r) + ": bend constraint " + std::to_string(k) +
                                                " index out of range");
            res[k] = {ri.nP + a, ri.nP + c, ri.nP + c};
        }
        ri.nBCol = greedy_colours(ri.nB, res, 2, ri.nP + ri.nF, order, start);
        for (int k = 0; k < ri.nB; ++k) {
            const int src = ri.bStart + order[k], dst = ri.bStart + k;
            bIdx[dst] = make_int4(desc->bend_a[src], desc->bend_b[src], order[k], 0);
            bPar[dst] = make4(desc->bend_rest[src], desc->bend_compliance[3 * size_t(src)],
                                    desc->bend_compliance[3 * size_t(src) + 1],
                                    desc->bend_compliance[3 * size_t(src) + 2]);
            bRest4[dst] = make4(desc->bend_rest_darboux[3 * size_t(src)], desc->bend_rest_darboux[3 * size_t(src) + 1],
                                desc->bend_rest_darboux[3 * size_t(src) + 2], Real(0));
        }
        colStart.insert(colStart.end(), start.begin(), start.end());
        if (ri.nSCol > kMaxColours || ri.nBCol > kMaxColours)
            return fail(RP_ERR_CAPACITY, "rod " + std::to_string(r) + " needs more than " +
                                             std::to_string(kMaxColours) + " colours");
        if (ri.nSCol > kTiledColours || ri.nBCol > kTiledColours) out.colourFit[r] = 0;
        // Contacts.
        if (w >= 0) {
            const auto& wd = desc->worlds[w];
            ri.nPrim = wd.num_prims;
            ri.primOff = primOffOfWorld[w];
            // contactMarginRadii * radius, in the working precision as the CPU
            // path forms it.
            ri.margin = wd.contact_margin_radii * ri.radius;
            ri.self = wd.self_collision ? 1 : 0;
            ri.selfFriction = wd.self_friction;
            ri.table = wd.self_collision ? wd.hash_table_size : 1;
            if (ri.self) {
                // selfCollisionIndexGap (src/core/collision.cpp:67-74), in the
                // working precision.
                Real lMin = ri.nS > 0 ? desc->stretch_rest[ri.sStart] : Real(1);
                for (int k = 0; k < ri.nS; ++k) lMin = std::min(lMin, desc->stretch_rest[ri.sStart + k]);
                const int geometric = int(std::floor(Real(2) * ri.radius / lMin)) + 2;
                ri.selfGap = std::max(wd.self_skip, geometric);
                // Bit 0: consecutive segments (index order) share a particle.
                // Bit 1: segment k joins particles k and k + 1 for every k, so
                // the kernels form segment ends without the segEnd table.
                int chain = 1, identity = 2;
                for (int k = 0; k < ri.nS; ++k) {
                    const int g = ri.sStart + k;
                    const int a = desc->stretch_p0[g], b = desc->stretch_p1[g];
                    if (a != k || b != k + 1) identity = 0;
                    if (k + 1 < ri.nS) {
                        const int c = desc->stretch_p0[g + 1], e = desc->stretch_p1[g + 1];
                        if (a != c && a != e && b != c && b != e) chain = 0;
                    }
                }
                ri.selfChain = chain | identity;
                if (wd.self_capacity > 0)
                    ri.selfCap = wd.self_capacity;
                else if (wd.self_capacity_per_segment > 0)
                    ri.selfCap = std::max(16, int(std::ceil(double(wd.self_capacity_per_segment) * ri.nS)));
                else
                    ri.selfCap = std::max(16, ri.nS / 2);  // src/gpu/gpu_batch.cpp:169
                ri.selfOrder = wd.self_order;
                // The coloured order keeps a colour mask per particle in the
                // broadphase's per-segment arrays (4 ints per segment).
                if (ri.selfOrder && ri.nP > 4 * ri.nS)
                    return fail(RP_ERR_INVALID, "rod " + std::to_string(r) +
                                                    ": the coloured self order needs at most 4 particles per segment");
            }
        } else {
            ri.table = 1;
        }
        maxPlainLds =
            std::max(maxPlainLds, Prec<Real>::plain_lds_bytes(ri.nP, ri.nF, ri.nS, ri.nB, ri.nPrim, ri.self, ri.table, ri.selfCap));
    }
    for (int r = 0; r < R; ++r) {
        out.anySelf = out.anySelf || out.rods[r].self != 0;
        out.anyColoured = out.anyColoured || (out.rods[r].self != 0 && out.rods[r].selfOrder != 0);
    }
    out.maxPlainLds = maxPlainLds;
    out.invMass.assign(desc->inv_mass, desc->inv_mass + P);
    out.iI4.resize(F);
    out.I4.resize(F);
    for (int j = 0; j < F; ++j) {
        out.iI4[j] = make4(desc->inv_inertia[3 * size_t(j)], desc->inv_inertia[3 * size_t(j) + 1],
                           desc->inv_inertia[3 * size_t(j) + 2], Real(0));
        out.I4[j] = make4(desc->inertia[3 * size_t(j)], desc->inertia[3 * size_t(j) + 1],
                          desc->inertia[3 * size_t(j) + 2], Real(0));
    }
    return RP_OK;
}

#define RP_HOST_INSTANTIATE(Real)                                                                              \
    template bool tile_config_available<Real>(int);                                                            \
    template int group_lanes<Real>(const RodInfoT<Real>&, const std::vector<int>&, int, int);                  \
    template int rod_tiled_lds<Real>(const RodInfoT<Real>&);                                                   \
    template int pack_rods<Real>(const std::vector<RodInfoT<Real>>&, const std::vector<int>&,                  \
                                 const std::vector<int>&, int, int, Packing&);                                 \
    template StepArgsT<Real> step_args<Real>(const Prec<Real>::StepParams&, int);                              \
    template int build_records<Real>(const Prec<Real>::BatchDesc*, RecordsT<Real>&);
RP_HOST_INSTANTIATE(float)
RP_HOST_INSTANTIATE(double)
#undef RP_HOST_INSTANTIATE

}  // namespace host
}  // namespace rp

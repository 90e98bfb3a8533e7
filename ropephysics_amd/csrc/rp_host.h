// Host-side pieces shared by the batch (rp_batch.cpp) and the serving loop
// (rp_server.cpp): error reporting, the tile-shape table, validation and
// colouring of a batch description into the records the kernels read, LDS
// limits, tile packing and step arguments.
#pragma once

#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "../../include/ropephysics.h"
#include "rp_layout.h"

namespace rp {

// Kernels and their launch wrappers (rp_kernels.hip), compiled once per
// precision: rp::f32 (rp_kernels.hip) and rp::f64 (rp_kernels64.hip).
#define RP_KERNEL_API(R)                                                                                       \
    int plain_lds_bytes(int nP, int nF, int nS, int nB, int nPrim, int self, int table, int cap);              \
    int tiled_lds_bytes(int nP, int nF, int nS, int nPrim, int self, int table, int cap);                      \
    hipError_t launch_plain(const DeviceArraysT<R>& d, const StepArgsT<R>& sa, size_t lds, hipStream_t s);      \
    hipError_t launch_global(const DeviceArraysT<R>& d, const StepArgsT<R>& sa, int threads, size_t xqLds,     \
                             hipStream_t s);                                                                   \
    int global_max_threads();                                                                                  \
    hipError_t allow_dynamic_lds(int bytes, int* granted);                                                     \
    hipError_t launch_tiled(int tp, int tc, int self, const DeviceArraysT<R>& d, const StepArgsT<R>& sa,      \
                            const Tile* tiles, const TileSlot* slots, int numTiles, int threads, size_t lds,   \
                            hipStream_t s);                                                                    \
    bool tiled_shape_available(int tp, int tc, int threads);                                                   \
    int tiled_resident(int tp, int tc, int self, int threads, size_t lds);                                     \
    hipError_t launch_fixed_frames(R* q, const R* src, const R* invInertia4, int f0, int count, hipStream_t s);
namespace f32 {
RP_KERNEL_API(float)
// Device address and size of the current device's fixed GLOBAL working-array
// region (rp_batch_step_noalloc); looked up once per device, allocates nothing.
hipError_t noalloc_scratch(char** out, size_t* bytes);
}
namespace f64 {
RP_KERNEL_API(double)
}
#undef RP_KERNEL_API

// One precision's kernel entry points and ABI types, for host code written
// once over the scalar.
template <typename Real>
struct Prec;

template <>
struct Prec<float> {
    static constexpr int bits = 32;
    using Vec4 = float4;
    using BatchDesc = rp_batch_desc;
    using WorldDesc = rp_world_desc;
    using Primitive = rp_primitive;
    using StepParams = rp_step_params;
    static constexpr auto plain_lds_bytes = &f32::plain_lds_bytes;
    static constexpr auto tiled_lds_bytes = &f32::tiled_lds_bytes;
    static constexpr auto allow_dynamic_lds = &f32::allow_dynamic_lds;
    static constexpr auto launch_plain = &f32::launch_plain;
    static constexpr auto launch_global = &f32::launch_global;
    static constexpr auto global_max_threads = &f32::global_max_threads;
    static constexpr auto launch_tiled = &f32::launch_tiled;
    static constexpr auto tiled_shape_available = &f32::tiled_shape_available;
    static constexpr auto tiled_resident = &f32::tiled_resident;
    static constexpr auto launch_fixed_frames = &f32::launch_fixed_frames;
    static Vec4 make4(float a, float b, float c, float d) { return make_float4(a, b, c, d); }
};

template <>
struct Prec<double> {
    static constexpr int bits = 64;
    using Vec4 = double4;
    using BatchDesc = rp_batch_desc64;
    using WorldDesc = rp_world_desc64;
    using Primitive = rp_primitive64;
    using StepParams = rp_step_params64;
    static constexpr auto plain_lds_bytes = &f64::plain_lds_bytes;
    static constexpr auto tiled_lds_bytes = &f64::tiled_lds_bytes;
    static constexpr auto allow_dynamic_lds = &f64::allow_dynamic_lds;
    static constexpr auto launch_plain = &f64::launch_plain;
    static constexpr auto launch_global = &f64::launch_global;
    static constexpr auto global_max_threads = &f64::global_max_threads;
    static constexpr auto launch_tiled = &f64::launch_tiled;
    static constexpr auto tiled_shape_available = &f64::tiled_shape_available;
    static constexpr auto tiled_resident = &f64::tiled_resident;
    static constexpr auto launch_fixed_frames = &f64::launch_fixed_frames;
    static Vec4 make4(double a, double b, double c, double d) { return make_double4(a, b, c, d); }
};

namespace host {

// Records `msg` as the calling thread's last error (rp_last_error) and
// returns `code`.
int fail(int code, const std::string& msg);
const char* last_error();

#define RP_HIP(call)                                                                                        \
    do {                                                                                                    \
        const hipError_t e_ = (call);                                                                       \
        if (e_ != hipSuccess)                                                                               \
            return ::rp::host::fail(RP_ERR_DEVICE, std::string(#call) + ": " + hipGetErrorString(e_));     \
    } while (0)

// A TILED launch shape: particles per lane (tp), constraint trips per colour
// (tc) and wavefronts per workgroup.
struct TileConfig {
    int tp, tc, waves;
    const char* name;
};
int num_tile_configs();
const TileConfig& tile_config(int id);  // id in 1 .. num_tile_configs()

// Whether tile shape `id` runs in precision Real (FP64 has no two-trip and
// no 512-lane shapes).
template <typename Real>
bool tile_config_available(int id);

// A batch description validated, coloured and laid out as the kernels read
// it: rod records with offsets from 0, constraints in colour order, 4-vector
// inertia, primitives of every world one after another.
template <typename Real>
struct RecordsT {
    using Vec4 = typename Prec<Real>::Vec4;
    int R = 0, P = 0, F = 0, S = 0, B = 0;
    std::vector<int> pStart, fStart, sStart, bStart;  // CSR, R + 1 entries
    std::vector<RodInfoT<Real>> rods;
    std::vector<unsigned char> colourFit;  // at most two colours per family
    std::vector<int4> sIdx, bIdx;
    std::vector<Vec4> sPar, bPar, bRest4, iI4, I4;
    std::vector<Real> invMass;
    std::vector<int> segEnd, colStart;
    std::vector<PrimT<Real>> prims;
    int maxPlainLds = 0;
    bool anySelf = false;
    bool anyColoured = false;  // some self-colliding rod asks for the coloured self order
};
using Records = RecordsT<float>;

template <typename Real>
int build_records(const typename Prec<Real>::BatchDesc* desc, RecordsT<Real>& out);

// LDS a workgroup may take (dynamic part) and the LDS of a compute unit;
// raises the stepping kernels' dynamic-LDS limit (both precisions) on first
// use.
struct LdsLimits {
    size_t perWorkgroup = 65536;
    size_t perCU = 65536;
    int computeUnits = 0;
};
int lds_limits(LdsLimits& out);
// Dynamic LDS budget of a workgroup when `perCU` of them share a CU.
int lds_budget(const LdsLimits& l, int perCU);
// Workgroups of a launch of `numTiles` pair up on `units` compute units when
// there are more than one and at most two per unit and at least two fit on a
// unit at once (`resident`, from the kernel's registers and LDS): dispatch
// gives every unit one workgroup, then a second, so b and b + units share a
// unit. Returns that distance, or 0. With one resident workgroup per unit the
// launch runs in two rounds instead, and nothing pairs.
inline int pair_units(int numTiles, int units, int resident) {
    return resident >= 2 && units > 0 && numTiles > units && numTiles <= 2 * units ? units : 0;
}

template <typename Real>
int group_lanes(const RodInfoT<Real>& r, const std::vector<int>& colStart, int tp, int tc);
template <typename Real>
int rod_tiled_lds(const RodInfoT<Real>& r);

// Rods `ids` of `rods` packed into workgroups of shape `configId`, each within
// `ldsCap` bytes of LDS (best fit, widest rods first).
struct Packing {
    int config = 0;
    int pairUnits = 0;  // pair_units of the packing's launch
    std::vector<Tile> tiles;
    std::vector<TileSlot> slots;
    int maxLds = 0;
    int lanesBusy = 0;
};
template <typename Real>
int pack_rods(const std::vector<RodInfoT<Real>>& rods, const std::vector<int>& ids, const std::vector<int>& colStart,
              int configId, int ldsCap, Packing& out);
// Reorders the tiles of a packing whose launch pairs workgroups b and b + units
// on a compute unit (`units` from pair_units): the costliest take the units
// they hold alone, the rest pair costliest with cheapest. No-op for units 0.
void place_tiles(Packing& p, int units);

template <typename Real>
StepArgsT<Real> step_args(const typename Prec<Real>::StepParams& p, int steps);

}  // namespace host
}  // namespace rp

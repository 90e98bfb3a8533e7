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
    hipError_t allow_dynamic_lds(int bytes, int* granted);                                                     \
    hipError_t launch_tiled(int tp, int tc, bool self, const DeviceArraysT<R>& d, const StepArgsT<R>& sa,      \
                            const Tile* tiles, const TileSlot* slots, int numTiles, int threads, size_t lds,   \
                            hipStream_t s);                                                                    \
    hipError_t launch_fixed_frames(R* q, const R* src, const R* invInertia4, int f0, int count, hipStream_t s);
namespace f32 {
RP_KERNEL_API(float)
}
namespace f64 {
RP_KERNEL_API(double)
}
#undef RP_KERNEL_API

// The host side below works on FP32 records.
using f32::DeviceArrays;
using f32::Prim;
using f32::RodInfo;
using f32::StepArgs;
using f32::allow_dynamic_lds;
using f32::launch_fixed_frames;
using f32::launch_plain;
using f32::launch_tiled;
using f32::plain_lds_bytes;
using f32::tiled_lds_bytes;

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

// A batch description validated, coloured and laid out as the kernels read
// it: rod records with offsets from 0, constraints in colour order, float4
// inertia, primitives of every world one after another.
struct Records {
    int R = 0, P = 0, F = 0, S = 0, B = 0;
    std::vector<int> pStart, fStart, sStart, bStart;  // CSR, R + 1 entries
    std::vector<RodInfo> rods;
    std::vector<unsigned char> colourFit;  // at most two colours per family
    std::vector<int4> sIdx, bIdx;
    std::vector<float4> sPar, bPar, bRest4, iI4, I4;
    std::vector<float> invMass;
    std::vector<int> segEnd, colStart;
    std::vector<Prim> prims;
    int maxPlainLds = 0;
    bool anySelf = false;
};
int build_records(const rp_batch_desc* desc, Records& out);

// LDS a workgroup may take (dynamic part) and the LDS of a compute unit;
// raises the stepping kernels' dynamic-LDS limit on first use.
struct LdsLimits {
    size_t perWorkgroup = 65536;
    size_t perCU = 65536;
};
int lds_limits(LdsLimits& out);
// Dynamic LDS budget of a workgroup when `perCU` of them share a CU.
int lds_budget(const LdsLimits& l, int perCU);

int group_lanes(const RodInfo& r, const std::vector<int>& colStart, int tp, int tc);
int rod_tiled_lds(const RodInfo& r);

// Rods `ids` of `rods` packed into workgroups of shape `configId`, each within
// `ldsCap` bytes of LDS (best fit, widest rods first).
struct Packing {
    int config = 0;
    std::vector<Tile> tiles;
    std::vector<TileSlot> slots;
    int maxLds = 0;
    int lanesBusy = 0;
};
int pack_rods(const std::vector<RodInfo>& rods, const std::vector<int>& ids, const std::vector<int>& colStart,
              int configId, int ldsCap, Packing& out);

StepArgs step_args(const rp_step_params& p, int steps);

}  // namespace host
}  // namespace rp

// Device-memory layout shared by the host side (rp_batch.cpp) and the
// kernels (rp_kernels.hip). Plain structs only.
//
// HBM holds every per-element array rod after rod (ragged CSR order):
//   state   x[P*3] v[P*3] q[F*4] omega[F*3]         (read and written)
//   loads   force[P*3] kinVel[P*3] torque[F*3]      (read)
//   consts  invMass[P] invInertia4[F] inertia4[F]   (read; float4, w unused)
//   stretch sIdx[S] (p0, p1, seg, -) and sPar[S] (rest, compliance xyz),
//           stored in each rod's colour order, colour after colour
//   bend    bIdx[B] (a, b, -, -), bPar[B] (rest, compliance xyz) and
//           bRest4[B] (rest Darboux xyz, -), same ordering
//   segEnd[S] (p0, p1) in the rod's original constraint order, for the
//           self-collision broadphase, which visits segments in index order
#pragma once

namespace rp {

constexpr int kMaxColours = 8;     // per family, PLAIN path
constexpr int kTiledColours = 2;   // per family, TILED path
constexpr int kMaxPrims = 8;       // per world
constexpr int kPrimFloats = 24;    // one primitive staged in LDS
constexpr int kContactFloats = 8;  // one world-contact slot
constexpr int kSelfFloats = 16;    // one self-contact slot

template <typename R>
struct RodInfoT {
    int pStart, nP, fStart, nF;
    int sStart, nS, bStart, nB;
    int world;     // index into worlds, -1 for none
    int nPrim;     // primitives of that world
    int self;      // self-collision on
    int selfGap;   // minimum segment index gap
    int selfCap;   // pool capacity
    int table;     // hash buckets
    int colOff;    // into colStart: stretch colours then bend colours
    int nSCol, nBCol;
    int primOff;   // into prims
    R radius;
    R margin;  // contact margin (margin radii * radius)
    R selfFriction;
    int selfChain;  // bit 0: consecutive segments (stretch constraints in index order) share a particle;
                    // bit 1: segment k joins particles k and k + 1 for every k
    int selfOrder;  // 0: self contacts projected in list order; 1: coloured order
    int pad;
};

template <typename R>
struct PrimT {
    int type;
    R a[3], b[3], normal[3], half[3], rot[4];
    R radius, friction;
};

// A lane group of a TILED workgroup: `lanes` consecutive lanes starting at
// `lane0` step rod `rod`, whose LDS region starts at byte `lds`.
struct TileSlot {
    int rod;
    int lane0;
    int lanes;
    int lds;
};

struct Tile {
    int firstSlot;
    int numSlots;
    int ldsBytes;
    int pad;
};

constexpr int kMaxWaypoints = 16;

// A kinematically driven particle: the gripper path of one rod. Positions are
// a smoothstep blend between consecutive waypoints over each leg's duration,
// held at the last waypoint afterwards; the velocity applied during a step is
// the average over its control period (reference HarnessTask::gripperAt and
// gripperVelocity, src/apps/harness_routing.cpp:69-92). Evaluated in FP64.
struct PathRec {
    int particle;  // rod-local index, -1: no path
    int nWay;
    int stepsPerControl;
    int pad;
    double height;  // z of the gripper
    double period;  // control period in seconds
    double sx, sy;  // start
    double legs[kMaxWaypoints];
    double wx[kMaxWaypoints];
    double wy[kMaxWaypoints];
};

// Largest rise of one particle's height above z0, sampled every `every` steps
// of the rod's clock (reference HangOutcome, src/apps/cable_hanging.h:68-77,
// observed as heldOnGpu does, src/validation/application_cases.cpp:55-65).
struct ObsRec {
    int particle;  // rod-local index, -1: no observer
    int every;
    double z0;
    double maxRise;
};

// Trajectory recording of one rod: after each step that brings the rod's
// clock to origin + k * every (k = 1..maxFrames), its positions are stored in
// FP32 into frame k - 1 of the batch's frame buffer, rod-major at the rod's
// particle start. This is the frame of the reference's trajectory writer
// (TrajectoryWriter::frame on downloaded positions,
// src/validation/scenes.cpp:45-68), written by the stepping kernel itself.
struct RecRec {
    int origin;  // clock at which recording started
    int every;   // steps per frame; 0: not recording
    int maxFrames;
    int pad;
};

template <typename R>
struct DeviceArraysT {
    // state
    R* x;
    R* v;
    R* q;
    R* omega;
    // loads
    const R* force;
    const R* kinVel;
    const R* torque;
    // constants
    const R* invMass;
    const R* invInertia4;
    const R* inertia4;
    const int* sIdx;  // int4
    const R* sPar;    // 4 reals
    const int* bIdx;  // int4
    const R* bPar;    // 4 reals
    const R* bRest4;  // 4 reals
    const int* segEnd;  // int2
    const int* colStart;
    const RodInfoT<R>* rods;
    const PrimT<R>* prims;
    R* complianceScale;
    R* frictionScale;
    int* selfOverflow;
    int* clock;            // steps taken, per rod
    const PathRec* paths;  // per rod, or null
    ObsRec* obs;           // per rod, or null
    const int* stepLimit;  // per rod: no step once the clock reaches it; null for none
    const RecRec* rec;     // per rod, or null
    float* frames;         // maxFrames frames of framePitch floats (3 per particle)
    long long framePitch;
    // GLOBAL strategy: rod r's working arrays (the PLAIN kernel's LDS layout)
    // at scratch + scratchOff[r] in HBM; null otherwise.
    char* scratch;
    const long long* scratchOff;
    int numRods;
};

template <typename R>
struct StepArgsT {
    R h;
    R invH2;
    R gx, gy, gz;
    R linDecay, angDecay;
    int substeps;
    int iterations;
    int contactInterval;
    int steps;  // steps in this launch
    // When positive, a workgroup starts no new step once this many device
    // wall-clock ticks have passed since it began (time-sliced launches; each
    // workgroup still takes at least one step).
    long long sliceTicks;
    // TILED launches whose workgroups pair up on the compute units: workgroups
    // b and b + pairUnits share one (0: no pairing). progress holds one step
    // count per workgroup of the launch.
    int* progress;
    int pairUnits;
};

// The two arithmetic precisions. Device code is compiled once per precision
// inside its namespace (f32: the reference GPU path's FP32; f64: the
// reference CPU path's FP64), with these names for the records.
namespace f32 {
using real = float;
using RodInfo = RodInfoT<float>;
using Prim = PrimT<float>;
using DeviceArrays = DeviceArraysT<float>;
using StepArgs = StepArgsT<float>;
}  // namespace f32
namespace f64 {
using real = double;
using RodInfo = RodInfoT<double>;
using Prim = PrimT<double>;
using DeviceArrays = DeviceArraysT<double>;
using StepArgs = StepArgsT<double>;
}  // namespace f64

}  // namespace rp

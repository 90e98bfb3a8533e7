/* ropephysics.h -- C ABI of the MI355X (gfx950) batched Cosserat-rod stepper.
 *
 * This is the drop-in boundary for RopePhysics' batched GPU step. The
 * reference exposes it as the C++ class crs::gpu::Batch
 * (src/gpu/gpu_solver.h:66-145) over an nvcc-only inner ABI
 * (src/gpu/gpu_kernels.h:147-172). Here the same operations are plain
 * extern "C" functions over pointers and sizes; the C++ facade in
 * ropephysics/batch.hpp maps the reference's Batch method set onto them.
 *
 * Conventions
 *   - Every call returns an int status (RP_OK == 0) unless stated; on failure
 *     rp_last_error() (thread-local) says why. Nothing fails silently: wrong
 *     sizes, out-of-range rods and device errors are reported.
 *   - The library owns all device memory. Host buffers are only read or
 *     written during the call that receives them.
 *   - A batch is used from one host thread at a time. Batches are
 *     independent; each has its own HIP stream, so work on two batches
 *     overlaps on the device. rp_batch_sync waits for one batch only.
 *   - Arrays are FP32, except in a batch made by rp_batch_create64, whose
 *     values are FP64 throughout (the *64 calls below). Vectors are xyz
 *     triples, quaternions wxyz quadruples (body -> world). Per-element
 *     arrays of a rod range are packed in rod order, each rod contributing
 *     its own element count (ragged CSR).
 */
#ifndef ROPEPHYSICS_H
#define ROPEPHYSICS_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define RP_ABI_VERSION 1

enum rp_status {
    RP_OK = 0,
    RP_ERR_INVALID = -1,  /* bad argument, size or index */
    RP_ERR_DEVICE = -2,   /* no usable device, or a HIP call failed */
    RP_ERR_MEMORY = -3,   /* device or host allocation failed */
    RP_ERR_CAPACITY = -4, /* a rod does not fit the chosen strategy's LDS budget */
    RP_ERR_STATE = -5     /* call not valid in the batch's current state */
};

/* Execution strategy.
 *   PLAIN   one rod per 128-thread workgroup, one launch per step, state
 *           re-read from and written to HBM every step: the reference's fused
 *           launch shape (src/gpu/gpu_solver.cu:531-681), kept as the baseline.
 *   TILED   rods of any sizes packed side by side into wavefront lane groups,
 *           state resident in LDS and VGPRs across every substep of many steps
 *           per launch; launches replayed from a HIP graph.
 *   GLOBAL  for rods whose working arrays exceed a workgroup's LDS, the
 *           counterpart of the reference's multi-kernel strategy
 *           (src/gpu/gpu_solver.cu:417-529,691-730; GpuSim falls back to it,
 *           src/gpu/gpu_suite.cpp:125-130): one rod per workgroup of up to 512
 *           lanes with its working arrays in HBM (allocated at the first step;
 *           positions and orientations held in LDS when every rod's fit two
 *           workgroups per compute unit), many steps per launch. Any rod length;
 *           same arithmetic and results as PLAIN, bit for bit.
 *   AUTO    TILED when every rod fits it, else PLAIN when every rod fits a
 *           workgroup's LDS, else GLOBAL. */
enum rp_strategy { RP_STRATEGY_PLAIN = 0, RP_STRATEGY_TILED = 1, RP_STRATEGY_AUTO = 2, RP_STRATEGY_GLOBAL = 3 };

enum rp_prim_type { RP_PLANE = 0, RP_SPHERE = 1, RP_CAPSULE = 2, RP_BOX = 3 };

/* Same fields, order and meaning as the reference's TPrimitive<float>
 * (src/core/geometry.h:17-65). */
typedef struct rp_primitive {
    int type;
    float a[3];           /* plane: point; sphere: centre; capsule: end A; box: centre */
    float b[3];           /* capsule: end B */
    float normal[3];      /* plane: unit outward normal */
    float half_extents[3];/* box */
    float rotation[4];    /* box orientation, wxyz */
    float radius;         /* sphere / capsule */
    float friction;
} rp_primitive;

/* One collision world (reference CollisionWorld, src/core/collision.h:83-106),
 * with the values the reference derives on the host already resolved. */
typedef struct rp_world_desc {
    int num_prims;
    const rp_primitive* prims;
    float contact_margin_radii; /* contactMarginRadii: a particle within this many rod radii of a
                                   surface gets a contact (reference default 0.25) */
    int self_collision;    /* 0 / 1 */
    float self_friction;
    int self_skip;         /* selfSkip; the gap used is max(self_skip, floor(2r / l_min) + 2) */
    int hash_table_size;   /* spatial-hash buckets (reference default 4096) */
    int self_capacity;     /* self contacts stored per rod, 0 for the rule below; more are
                              counted as overflow (and dropped, where the CPU path keeps them) */
    float self_capacity_per_segment; /* with self_capacity 0: max(16, ceil(this * segments)) slots
                                        when positive, else the reference's max(16, segments / 2) */
    int self_order;        /* 0: self contacts projected one at a time in list order, the
                              reference's order (projectContacts, src/core/solver.cpp:163-216);
                              1: coloured order, an opt-in of this library with no counterpart
                              in the reference -- each regeneration colours the list greedily in
                              list order (the lowest of 32 colours no earlier contact sharing a
                              particle holds; a tail when all are taken) and projects colour by
                              colour, a colour's contacts at once, then the tail in list order.
                              Results then differ from the reference's, within the solver's own
                              ordering freedom; they equal a one-at-a-time projection in that
                              order bit for bit. */
} rp_world_desc;

/* A ragged batch: rods may differ in every size. Element arrays are indexed
 * globally (rod r owns [x_start[r], x_start[r+1])); constraint endpoint
 * indices are rod-local. Constraints may be given in any order; the library
 * computes the colouring (greedy, in index order) itself. */
typedef struct rp_batch_desc {
    int num_rods;
    const int* particle_start; /* [num_rods + 1] */
    const int* frame_start;    /* [num_rods + 1]: segment frames then fixed ghost frames */
    const int* stretch_start;  /* [num_rods + 1] */
    const int* bend_start;     /* [num_rods + 1] */
    const float* inv_mass;     /* [P] */
    const float* inv_inertia;  /* [F * 3] body-frame diagonal */
    const float* inertia;      /* [F * 3] */
    const int* stretch_p0;     /* [S] */
    const int* stretch_p1;     /* [S] */
    const int* stretch_seg;    /* [S] */
    const float* stretch_rest; /* [S] */
    const float* stretch_compliance; /* [S * 3] */
    const int* bend_a;         /* [B] */
    const int* bend_b;         /* [B] */
    const float* bend_rest;    /* [B] */
    const float* bend_compliance;    /* [B * 3] */
    const float* bend_rest_darboux;  /* [B * 3] */
    const float* radius;       /* [num_rods] */
    int num_worlds;            /* 0: no contacts at all */
    const rp_world_desc* worlds;
    const int* world_of_rod;   /* [num_rods], or NULL: every rod uses world 0 */
    int strategy;              /* rp_strategy */
    int tile_config;           /* 0: autotune; else an id from rp_tile_config_name */
} rp_batch_desc;

/* BatchParams (src/gpu/gpu_solver.h:52-60). */
typedef struct rp_step_params {
    float dt;
    int substeps;
    int iterations;
    float gravity[3];
    float linear_damping;
    float angular_damping;
    int contact_interval;
} rp_step_params;

typedef struct rp_device_info {
    char name[256];
    char arch[64];
    int compute_units;
    int wavefront_size;
    size_t lds_per_workgroup;
    size_t total_global_mem;
    int clock_khz;
    int max_tiled_segments;  /* longest chain rod the TILED path accepts (no contacts) */
} rp_device_info;

typedef struct rp_counters {
    long long self_contact_overflow; /* summed over rods and steps */
    int strategy;                    /* resolved strategy */
    int tile_config;                 /* resolved tile configuration id */
    int workgroups;                  /* per launch */
    int lanes_busy;                  /* lanes assigned to rods, over all workgroups */
    int lanes_total;
    size_t lds_bytes_per_workgroup;  /* largest */
    int launches_per_step;           /* PLAIN: 1; TILED: 0 (a call's steps share launches; see
                                        last_step_launches) */
    long long steps_taken;
    size_t device_bytes;             /* device memory held by the batch */
    float last_step_gpu_ms;          /* with timing on: HIP-event time of the last rp_batch_step */
    int last_step_launches;          /* kernel launches in the last rp_batch_step */
    double motion_margin;            /* the reference's float motion margin (floatMotionMargin,
                                        src/apps/cable_hanging.cpp:75-81): gravity's motion per substep
                                        |g| h^2 over the spacing of the batch's precision at the largest
                                        coordinate ever uploaded; the smallest over the steps taken, 0
                                        before any step with gravity. Below about 1 a resting rod cannot
                                        start to slide: rounding acts as static friction. */
    long long low_margin_steps;      /* steps taken with motion_margin below 2, the reference's bar
                                        (src/validation/application_cases.cpp:196-197) */
    size_t scratch_bytes;            /* GLOBAL: the HBM working arrays (0 until the first step) */
} rp_counters;

/* cudaAvailable / deviceName / diagnostic / maxFusedSegments
 * (src/gpu/gpu_solver.h:44-48,150). */
int rp_device_query(rp_device_info* out);

/* Batch::create (src/gpu/gpu_batch.cpp:150-316): allocate, colour, pack and
 * upload topology; state starts zeroed (upload it with rp_batch_set_state). */
int rp_batch_create(const rp_batch_desc* desc, void** out_batch);

/* Batch::~Batch / destroy (src/gpu/gpu_batch.cpp:132-138). */
void rp_batch_destroy(void* batch);

/* Batch::upload / uploadRod (src/gpu/gpu_batch.cpp:318-370), in bulk for rods
 * [first_rod, first_rod + num_rods). x, v: particles * 3; q: frames * 4;
 * omega: frames * 3. NULL leaves a field unchanged. */
int rp_batch_set_state(void* batch, int first_rod, int num_rods, const float* x, const float* v, const float* q,
                       const float* omega);

/* Batch::setLoads / setKinematicVelocities (src/gpu/gpu_batch.cpp:372-391,
 * 418-432): external force and kinematic velocity per particle, external
 * torque (world frame) per frame, and -- when fixed_q is given, frames * 4 --
 * the orientation of every zero-inertia (fixed) frame. NULL leaves a field
 * unchanged. */
int rp_batch_set_loads(void* batch, int first_rod, int num_rods, const float* force, const float* kin_vel,
                       const float* torque, const float* fixed_q);

/* Batch::setMaterialScales (src/gpu/gpu_batch.cpp:434-443): per-rod Young's
 * modulus scale (every compliance is divided by it) and friction scale, each
 * [num_rods] or NULL to leave unchanged. Non-positive youngs scales are an
 * error. */
int rp_batch_set_scales(void* batch, const float* youngs_scale, const float* friction_scale);

/* Batch::step (src/gpu/gpu_batch.cpp:457-473), num_steps times, asynchronous
 * on the batch's stream. TILED runs the call's steps in as few launches as
 * possible: at most 8192 steps each, and about 4 s of device time as estimated
 * from the autotuner's trial (never fewer than 64 steps); RP_STEPS_PER_LAUNCH
 * in the environment at creation replaces both bounds. PLAIN launches once per
 * step. GLOBAL runs about 2^22 element-substeps of its longest rod per launch.
 * Results do not depend on how steps are split into calls or launches. */
int rp_batch_step(void* batch, const rp_step_params* params, int num_steps);

/* rp_batch_step without device allocation. No counterpart in the reference:
 * for callers that cannot call hipMalloc -- code under HIP graph capture,
 * tight serving loops, harnesses that forbid it.
 *   - Steps an FP32 batch of the GLOBAL strategy exactly as rp_batch_step does
 *     (the same kernel, the same results bit for bit), but with the rods'
 *     working arrays in a fixed region of RP_NOALLOC_SCRATCH_BYTES that the
 *     library holds on each device from the first call made with that device
 *     current until the process ends, instead of the batch's own (which this
 *     call never allocates). Launches go straight onto the batch's stream; no
 *     graph is captured or instantiated; nothing is allocated or freed.
 *   - Refused: batches whose working arrays need more than the region
 *     (RP_ERR_CAPACITY; rp_counters cannot tell before the first
 *     rp_batch_step, so the message gives the size), other strategies, FP64
 *     batches, and a current device other than the batch's (RP_ERR_STATE).
 *   - Every call on a device shares that device's region. Calls on one
 *     batch's stream are ordered; calls whose work overlaps on one device
 *     (different batches' streams, several host threads) race on the region
 *     and must be serialised by the caller. The kernel writes every word of
 *     the region it reads within each launch, so the region needs no clearing
 *     between calls. */
#define RP_NOALLOC_SCRATCH_BYTES (32u << 20)
int rp_batch_step_noalloc(void* batch, const rp_step_params* params, int num_steps);

/* Batch::synchronize (src/gpu/gpu_batch.cpp:475), scoped to this batch. */
int rp_batch_sync(void* batch);

/* Batch::download (src/gpu/gpu_batch.cpp:393-416), in bulk; waits for the
 * batch's pending steps. NULL skips a field. */
int rp_batch_get_state(void* batch, int first_rod, int num_rods, float* x, float* v, float* q, float* omega);

/* Batch::selfContactOverflow, launchesPerStep, sharedBytesPerBlock
 * (src/gpu/gpu_batch.cpp:140-148,477-484). */
int rp_batch_counters(void* batch, rp_counters* out);

/* A kinematically driven particle per rod (ragged extension of
 * setKinematicVelocities, src/gpu/gpu_batch.cpp:418-432, with the gripper
 * schedule of HarnessTask::gripperAt/gripperVelocity,
 * src/apps/harness_routing.cpp:69-92, evaluated on the device in FP64):
 * the gripper starts at `start`, blends to waypoint k over leg_seconds[k]
 * with a smoothstep, and holds at the last waypoint; during the rod's step s
 * the particle's kinematic velocity is (P(t1) - P(t0)) / control_period with
 * t0 = floor(s / steps_per_control) * control_period and t1 = t0 +
 * control_period, P at height `height`. The particle must be pinned. */
typedef struct rp_path_desc {
    int particle;            /* rod-local particle index; -1 clears the rod's path */
    int num_waypoints;       /* 1..16 */
    const double* waypoints; /* num_waypoints * 2 (x, y) */
    const double* leg_seconds; /* num_waypoints */
    double start[2];
    double height;
    double control_period;   /* seconds; the reference uses dt * steps_per_control */
    int steps_per_control;
} rp_path_desc;

/* Install paths for rods [first_rod, first_rod + num_rods); paths[i] belongs
 * to rod first_rod + i. */
int rp_batch_set_paths(void* batch, int first_rod, int num_rods, const rp_path_desc* paths);

/* Per-rod step clock (steps taken since it was last set), which paths read.
 * clock == NULL resets to 0. */
int rp_batch_set_clock(void* batch, int first_rod, int num_rods, const int* clock);
int rp_batch_get_clock(void* batch, int first_rod, int num_rods, int* clock);

/* A hold/slide observer per rod (HangOutcome, src/apps/cable_hanging.h:68-77,
 * sampled as heldOnGpu does, src/validation/application_cases.cpp:55-65):
 * whenever the rod's clock reaches a multiple of `every`, the observer takes
 * max_rise = max(max_rise, z(particle) - z0) in FP64. Setting an observer
 * resets its max_rise to 0. */
typedef struct rp_observer_desc {
    int particle;  /* rod-local particle index; -1 clears */
    int every;     /* steps between samples, >= 1 */
    double z0;     /* reference height */
} rp_observer_desc;
int rp_batch_set_observers(void* batch, int first_rod, int num_rods, const rp_observer_desc* observers);
int rp_batch_get_observations(void* batch, int first_rod, int num_rods, double* max_rise);

/* Trajectory recording on the device. It replaces the reference's frame loop,
 * which downloads every rod's positions after each frame's steps
 * (Batch::downloadPositions, src/gpu/gpu_batch.cpp:445-455) and writes them
 * with TrajectoryWriter::frame (src/validation/scenes.cpp:45-68): here the
 * stepping kernels store the frames themselves, into a device buffer.
 *   - rp_batch_record starts recording at every rod's current clock. After
 *     each step that brings a rod's clock to origin + k * every, for
 *     k = 1..max_frames, that rod's positions go into frame k - 1.
 *   - A frame is FP32 x, y, z per particle, rod-major in the batch's particle
 *     order (rod r at particle_start[r]), for FP32 and FP64 batches alike.
 *   - The buffer (max_frames frames) is zeroed on each call. every = 0 stops
 *     recording and frees it.
 *   - rp_batch_set_clock does not move the origin; call rp_batch_record again
 *     after it.
 * rp_batch_recorded_frames gives the number of frames that every rod has
 * completed. rp_batch_get_frames copies frames [first_frame, first_frame +
 * num_frames) of the buffer to the host, num_frames * 3 * particles floats. */
int rp_batch_record(void* batch, int every, int max_frames);
int rp_batch_recorded_frames(void* batch, int* frames);
int rp_batch_get_frames(void* batch, int first_frame, int num_frames, float* positions);

/* Record HIP events around every rp_batch_step on the batch's stream
 * (enable = 1) so rp_counters reports its device time. Off by default. */
int rp_batch_set_timing(void* batch, int enable);

/* The autotuner's trials: for each tile configuration it timed, its id and
 * the device time of an 8-step trial launch (best of two), in order. Returns
 * the number of trials (which may exceed capacity), or a negative status. */
int rp_batch_tune_report(void* batch, int* ids, float* ms, int capacity);

/* Name of a tile configuration id (for logs and the autotuner's report);
 * NULL for an unknown id. */
const char* rp_tile_config_name(int id);

/* Number of tile configuration ids the TILED strategy knows (ids 1..n). */
int rp_tile_config_count(void);

/* Thread-local message for the last failed call on this thread. */
const char* rp_last_error(void);

/* ------------------------------------------------------------------ FP64 batches
 *
 * The same batch in double precision: every real array, parameter and state
 * value is FP64 on the host and on the device, and the kernels perform the
 * reference CPU solver's arithmetic in that solver's own default precision
 * (crs::Real = double, src/core/math3.h:14-22) -- where FP32's spacing at
 * metre scale (about 6e-8 m) is coarser than one substep's gravity drop, a
 * hold/slide or routed/not-routed verdict can differ between FP32 and FP64;
 * an FP64 batch gives the CPU path's verdicts, computed on the GPU. Each rod's
 * results equal the reference CPU step's in colour order, bit for bit.
 *
 * The structs mirror their FP32 counterparts field for field. Calls not
 * listed here (destroy, sync, counters, paths, clock, observers, timing, tune
 * report) take either kind of batch. FP32 state and step calls refuse an FP64
 * batch, and these refuse an FP32 one (RP_ERR_STATE). TILED FP64 batches use
 * the shapes with one constraint trip per colour and at most 256 lanes. */
typedef struct rp_primitive64 {
    int type;
    double a[3];
    double b[3];
    double normal[3];
    double half_extents[3];
    double rotation[4];
    double radius;
    double friction;
} rp_primitive64;

typedef struct rp_world_desc64 {
    int num_prims;
    const rp_primitive64* prims;
    double contact_margin_radii;
    int self_collision;
    double self_friction;
    int self_skip;
    int hash_table_size;
    int self_capacity;
    double self_capacity_per_segment;
    int self_order;
} rp_world_desc64;

typedef struct rp_batch_desc64 {
    int num_rods;
    const int* particle_start;
    const int* frame_start;
    const int* stretch_start;
    const int* bend_start;
    const double* inv_mass;
    const double* inv_inertia;
    const double* inertia;
    const int* stretch_p0;
    const int* stretch_p1;
    const int* stretch_seg;
    const double* stretch_rest;
    const double* stretch_compliance;
    const int* bend_a;
    const int* bend_b;
    const double* bend_rest;
    const double* bend_compliance;
    const double* bend_rest_darboux;
    const double* radius;
    int num_worlds;
    const rp_world_desc64* worlds;
    const int* world_of_rod;
    int strategy;
    int tile_config;
} rp_batch_desc64;

typedef struct rp_step_params64 {
    double dt;
    int substeps;
    int iterations;
    double gravity[3];
    double linear_damping;
    double angular_damping;
    int contact_interval;
} rp_step_params64;

/* As rp_batch_create, rp_batch_set_state, rp_batch_set_loads,
 * rp_batch_set_scales, rp_batch_step and rp_batch_get_state, in FP64. */
int rp_batch_create64(const rp_batch_desc64* desc, void** out_batch);
int rp_batch_set_state64(void* batch, int first_rod, int num_rods, const double* x, const double* v,
                         const double* q, const double* omega);
int rp_batch_set_loads64(void* batch, int first_rod, int num_rods, const double* force, const double* kin_vel,
                         const double* torque, const double* fixed_q);
int rp_batch_set_scales64(void* batch, const double* youngs_scale, const double* friction_scale);
int rp_batch_step64(void* batch, const rp_step_params64* params, int num_steps);
int rp_batch_get_state64(void* batch, int first_rod, int num_rods, double* x, double* v, double* q, double* omega);

/* Bits of a batch's arithmetic: 32 or 64; a negative status for NULL. */
int rp_batch_precision(void* batch);

/* ------------------------------------------------------------------ serving
 *
 * A continuous-batching server. No counterpart in the reference: its callers
 * create one Batch per workload. Requests (each a set of rods with its own
 * collision world, described exactly as for rp_batch_create, plus an initial
 * state and a horizon in steps) are admitted while the device memory reserved
 * for them (an HBM budget fixed at creation) has room, stay resident between
 * calls, and advance together: each rp_server_step launches steps_per_launch
 * steps over the rods of every resident unfinished request (rods without
 * self-collision in one TILED launch, rods with it in another on a second
 * stream, the two sharing the device), every rod stopping at its own
 * request's horizon. One step definition
 * (rp_step_params) serves every request. A request's result is bitwise the
 * one it gets alone: the per-rod arithmetic does not depend on its neighbours.
 * Every server of a process uses the same two streams per device, created
 * once with a full CU mask so that each has a hardware queue of its own.
 *
 * Use: rp_server_sync; read and release what rp_server_finished lists; submit
 * what has arrived; rp_server_step; repeat. Not thread-safe. */

typedef struct rp_server_desc {
    long long hbm_budget_bytes;  /* device memory for resident requests' state, constants and records */
    int max_rods;                /* rod records (0: one per 16 segments of the budget) */
    rp_step_params params;       /* the step every request takes */
    int tile_config;             /* 0: tp3-tc1-w4 for rods without self-collision, tp1-tc1-w8 for
                                    rods with it (tp2-tc1-w4 when time-sliced); else this shape
                                    for both */
    int steps_per_launch;        /* steps per rp_server_step (0: 64) */
    int self_tile_config;        /* when positive: the shape for rods with self-collision */
    int time_slice_us;           /* when positive: time-sliced steps (below) */
} rp_server_desc;

/* A request's initial state and loads, packed rod after rod like the
 * request's CSR; only x and q are required (the rest default to zero). The
 * scales are per rod and may be NULL (1). */
typedef struct rp_request_state {
    const float* x;
    const float* v;
    const float* q;
    const float* omega;
    const float* force;
    const float* kin_vel;
    const float* torque;
    const float* youngs_scale;
    const float* friction_scale;
} rp_request_state;

typedef struct rp_request_info {
    int num_rods, num_particles, num_frames;
    int steps;       /* horizon */
    int steps_done;  /* steps issued so far (all of them once finished) */
    long long device_bytes;
} rp_request_info;

typedef struct rp_server_stats {
    int resident_requests;  /* admitted and not released */
    int running_requests;   /* of those, not yet at their horizon */
    int running_rods;
    long long bytes_in_use;  /* device bytes held by resident requests */
    long long device_bytes;  /* allocated at creation */
    long long capacity_particles, used_particles;
    long long capacity_rods, used_rods;
    long long launches;
    int workgroups;  /* of the last packing, both launches */
    int tile_config; /* shape of the rods without self-collision */
    int steps_per_launch;
} rp_server_stats;

int rp_server_create(const rp_server_desc* desc, void** out_server);
void rp_server_destroy(void* server);

/* Admit a request of `steps` steps. The host arrays are copied into pinned
 * staging memory before the call returns; their upload is queued behind the
 * launch in flight, not waited for. RP_ERR_CAPACITY: no room now (release
 * finished requests and retry) or never (larger than the budget, or a rod no
 * tile shape takes; rp_last_error says which). */
int rp_server_submit(void* server, const rp_batch_desc* request, const rp_request_state* state, int steps,
                     int* out_id);

/* Launch one chunk of steps for every running request (asynchronous); sets
 * *num_finished to the number of requests that reach their horizon in it.
 * With time_slice_us > 0 the chunk is a time slice instead: every workgroup
 * runs as many steps as it starts within the slice (at least one, at most
 * 8192 and up to its rods' horizons), so light rods are not held to the pace
 * of the slowest. Progress is then known only once the launch completes:
 * *num_finished is 0, and rp_server_sync, rp_server_finished and the other
 * queries read the rods' step clocks back first. Results are the same bit for
 * bit whichever mode, slice or chunk length. */
int rp_server_step(void* server, int* num_finished);

/* Wait for every launched chunk. */
int rp_server_sync(void* server);

/* Ids of finished, unreleased requests (up to capacity); returns how many
 * there are. */
int rp_server_finished(void* server, int* ids, int capacity);

int rp_server_request_info(void* server, int id, rp_request_info* out);

/* Copy a finished request's state out (any pointer may be NULL); waits for
 * its last chunk. Sizes as its CSR: x, v particles x 3; q frames x 4 (w, x,
 * y, z); omega frames x 3. */
int rp_server_result(void* server, int id, float* x, float* q, float* v, float* omega);

/* Free a request's device ranges (finished or not). */
int rp_server_release(void* server, int id);

int rp_server_get_stats(void* server, rp_server_stats* out);

#ifdef __cplusplus
}
#endif
#endif /* ROPEPHYSICS_H */

/* C interface of the CPU oracle -- TEST INFRASTRUCTURE.
 *
 * Two libraries export exactly this interface:
 *   oracle/liboracle.so        the restatement in rp_oracle.hpp (FP64 and FP32);
 *   oracle/_ref/libref64.so,   the reference's own CPU sources, compiled from the
 *   oracle/_ref/libref32.so    RopePhysics checkout by oracle/Makefile (FP64, and
 *                              FP32 with CRS_REAL_FLOAT): they pin the restatement,
 *                              and libref64.so is the CPU baseline bench.py times.
 *                              The built libraries travel with the tree; the
 *                              reference's sources never do.
 * Only tests/, the smoke check and bench.py's CPU-baseline and verdict-check legs
 * load them, always as the checker or the baseline, never as the product.
 *
 * Every array crosses as double; a precision-32 rod stores float internally and
 * rounds on the way in, exactly as the reference's Batch does (toF).
 */
#ifndef RP_ORACLE_CAPI_H
#define RP_ORACLE_CAPI_H

#ifdef __cplusplus
extern "C" {
#endif

/* Caller-owned views onto one rod's arrays. Sizes come from orc_rod_sizes.
 * Vectors are xyz triples, quaternions wxyz quadruples. A NULL pointer skips
 * that field in both directions. */
typedef struct orc_rod_view {
    int nP, nF, nStretch, nBend;
    double *x, *xPrev, *v, *invMass, *mass, *extForce, *kinVel;
    double *q, *qPrev, *omega, *invInertia, *inertia, *extTorque;
    int *sP0, *sP1, *sSeg;
    double *sRest, *sComp, *sLam;
    int *bA, *bB;
    double *bRest, *bRestD, *bComp, *bLam;
    double material[4]; /* youngs, poisson, density, radius */
} orc_rod_view;

typedef struct orc_prim {
    int type; /* 0 plane, 1 sphere, 2 capsule, 3 box */
    double a[3], b[3], normal[3], half[3], rot[4];
    double radius, friction;
} orc_prim;

typedef struct orc_world {
    int numPrims;
    const orc_prim* prims;
    int selfCollision;
    double selfFriction;
    int selfSkip;
    double marginRadii;
    int hashTableSize;
    int selfOrder; /* 0: self contacts projected in list order (the reference's); 1: coloured order */
} orc_world;

typedef struct orc_params {
    double dt;
    int substeps, iterations;
    double gravity[3];
    double linDamp, angDamp;
    int contactInterval;
    int coloured;            /* 1: colour-ordered sweep (the GPU's order) */
    double complianceScale;  /* per-rod alpha multiplier (batch semantics) */
    double frictionScale;    /* per-rod friction multiplier */
} orc_params;

/* Rods. precision is 64 or 32. */
void* orc_rod_straight(int precision, int n, double length, const double material[4], const double origin[3],
                       const double dir[3]);
void* orc_rod_centerline(int precision, int numPoints, const double* points, const double material[4],
                         const double seed[3], double twistPerLength);
int orc_rod_clamp_end(void* rod, int tip);
void orc_rod_pin(void* rod, int particle);
void orc_rod_set_rest_from_current(void* rod);
void orc_rod_sizes(void* rod, int out[4]);
void orc_rod_get(void* rod, orc_rod_view* out);
int orc_rod_set(void* rod, const orc_rod_view* in);
void orc_rod_free(void* rod);

/* Step `steps` full steps. world may be NULL (contact-free step). The contact
 * list and spatial hash persist in the rod handle between calls. Returns 0. */
int orc_step(void* rod, const orc_params* p, const orc_world* world, int steps);
/* Contacts generated at the last regeneration: world (1-particle) and self. */
void orc_contact_counts(void* rod, int* worldContacts, int* selfContacts);
/* Step many rods on `threads` std::threads, static partition over rods. */
int orc_step_many(void** rods, int count, const orc_params* p, const orc_world* const* worlds, int steps,
                  int threads);
/* As orc_step_many, and before step s of rod r the kinematic velocity of its
 * particle driven[r] (when >= 0) is set to kin_vel[(r * steps + s) * 3 ..+3]:
 * a gripper schedule evaluated by the caller. */
int orc_step_many_driven(void** rods, int count, const orc_params* p, const orc_world* const* worlds, int steps,
                         int threads, const int* driven, const double* kin_vel);

/* Pieces of the arithmetic, for pinning against textbook routines. */
int orc_chol_solve3(int precision, const double A[9], const double b[3], double y[3]);
double orc_signed_distance(int precision, const orc_prim* prim, const double p[3], double n[3]);
void orc_seg_seg(int precision, const double p1[3], const double q1[3], const double p2[3], const double q2[3],
                 double st[2]);
unsigned orc_hash_cell(int ix, int iy, int iz, int table);

/* The self contacts of the last contact generation, in list order: their four
 * particle indices, and (when the world asked for the coloured order) each
 * one's colour (32: the sequential tail) and its position in the projection
 * order among the self contacts; -1 otherwise. Returns the number of self
 * contacts (which may exceed capacity; only capacity are written). */
int orc_self_contacts(void* rod, int* idx4, int* colour, int* rank, int capacity);
int orc_cell_coord(int precision, double v, double cell);
int orc_self_gap(void* rod, const orc_world* world);
/* Step a copy of an FP64 rod with the operation-counting scalar and report
 * {adds (incl. subtracts), multiplies, divides, square roots, compares}.
 * The rod itself is not advanced. Returns 0, or -1 for an FP32 rod. */
int orc_count_ops(void* rod, const orc_params* p, const orc_world* world, int steps, long long out[5]);
/* Greedy colouring of a rod's stretch (family 0) or bend (family 1)
 * constraints: writes order[count] and start[colours + 1] (start capacity
 * maxColours + 1); returns the colour count. */
int orc_colour(void* rod, int family, int* order, int* start, int maxColours);

#ifdef __cplusplus
}
#endif
#endif

/*
 * melon3d_ext.h - extension entry points of the MI355X backend (m3dx_*).
 *
 * None of these exist in the reference C API; they sit beside the unchanged
 * m3d_* functions of melon3d.h and serve four callers:
 *   - parity tests: full-state snapshot/restore so one step can be compared
 *     against the reference engine from identical state, plus read-only
 *     views of per-stage outputs (grid tiers, islands, colors);
 *   - the benchmark: device/stage counters for the HBM-traffic model;
 *   - device selection and error reporting.
 * Plain C types only; no torch types cross this boundary.
 */
#pragma once

#ifndef M3D_VERSION_MAJOR /* melon3d.h already seen (possibly from another include path) */
#include "melon3d.h"
#endif

#ifdef __cplusplus
extern "C"
{
#endif

/* ------------------------------------------------------------ snapshot */

/* One body slot (in use or free), mirroring the reference Body
 * (src/m3d_internal.h:77-119) plus its broadphase sidecars
 * (src/m3d_internal.h:627-639). All fields are 4 bytes wide. */
typedef struct m3dx_body_state
{
	float position[3];
	float rotation[4];
	float linearVelocity[3];
	float angularVelocity[3];
	float prevPosition[3];
	float prevRotation[4];
	float prevLinearVelocity[3];
	float prevAngularVelocity[3];
	float invMass;
	float invInertiaLocal[3];
	float invInertiaWorld[9]; /* columns cx, cy, cz */
	int32_t shapeType;
	float radius;
	float halfHeight;
	float halfExtents[3];
	float friction;
	float restitution;
	float rollingResistance;
	float mass;
	float force[3];
	float torque[3];
	uint32_t categoryBits;
	uint32_t maskBits;
	int32_t groupIndex;
	int32_t isSensor;
	int32_t type;
	float linearDamping;
	float angularDamping;
	float gravityScale;
	float sleepTime;
	float smoothedEnergy;
	int32_t enableSleep;
	int32_t isAwake;
	int32_t inUse;
	uint32_t generation;
	uint32_t userDataLo;
	uint32_t userDataHi;
	/* broadphase sidecars */
	float fatAABB[6]; /* lower xyz, upper xyz */
	int32_t hasFatAABB;
	int32_t inStableTier; /* migrated into the stable tier at its last build */
	int32_t inStableCells; /* has entries in the stable tier (not oversized) */
} m3dx_body_state;

/* One manifold point (src/m3d_internal.h:154-187). */
typedef struct m3dx_point_state
{
	float localAnchorA[3];
	float localAnchorB[3];
	float frictionAnchorA[3];
	float frictionAnchorB[3];
	float separation;
	uint32_t id;
	float lambdaN;
	float lambdaT;
	float separationOffset;
	float slipDir[3];
	float rA[3];
	float rB[3];
	float wN;
	float angJA[3];
	float angJB[3];
} m3dx_point_state;

/* One contact (src/m3d_internal.h:189-229). */
typedef struct m3dx_contact_state
{
	int32_t bodyA;
	int32_t bodyB;
	float friction;
	float restitution;
	float rollingResistance;
	int32_t touching;
	int32_t isSensor;
	int32_t tunnelGuard;
	int32_t curved;
	int32_t hasCache;
	float cachePosA[3];
	float cachePosB[3];
	float cacheRotA[4];
	float cacheRotB[4];
	float normal[3];
	int32_t pointCount;
	m3dx_point_state points[4];
} m3dx_contact_state;

/* One joint slot (src/m3d_internal.h:313-349). type: 0 distance, 1 ball,
 * 2 hinge, 3 weld. */
typedef struct m3dx_joint_state
{
	int32_t type;
	int32_t bodyA;
	int32_t bodyB;
	float localAnchorA[3];
	float localAnchorB[3];
	float length;
	float hertz;
	float dampingRatio;
	float localAxisA[3];
	float localAxisB[3];
	float localRefA[3];
	float localRefB[3];
	int32_t enableLimit;
	float lowerAngle;
	float upperAngle;
	int32_t enableMotor;
	float motorSpeed;
	float maxMotorTorque;
	float relRot0[4];
	float compliance;
	float lambda;
	float motorLambda;
	int32_t inUse;
	uint32_t generation;
} m3dx_joint_state;

/* Broadphase grid entry: packed cell key (21 bits per axis) and body. */
typedef struct m3dx_grid_entry
{
	uint64_t key;
	int32_t body;
	int32_t pad;
} m3dx_grid_entry;

typedef struct m3dx_world_header
{
	int32_t bodyCapacity;
	int32_t contactCount;
	int32_t jointCapacity;
	int32_t freeBodyCount;
	int32_t freeJointCount;
	int32_t stableEntryCount;
	int32_t stableOversizedCount;
	int32_t gridDirty;
	int32_t stableDirty;
	int32_t stableForce;
	int32_t stepsSinceStableBuild;
	int32_t anyForcesApplied;
	int32_t jointedPairsDirty;
	float gravity[3];
	int32_t workerCount;
	int32_t positionIterations;
	float gridCellSize;
	int32_t enableSleep;
} m3dx_world_header;

/* Caller-owned arrays sized from header (fill the header with a first call
 * whose array pointers are NULL). */
typedef struct m3dx_snapshot
{
	m3dx_world_header header;
	m3dx_body_state* bodies;          /* bodyCapacity */
	m3dx_contact_state* contacts;     /* contactCount */
	m3dx_joint_state* joints;         /* jointCapacity */
	int32_t* freeBodies;              /* freeBodyCount, in free-list order */
	int32_t* freeJoints;              /* freeJointCount */
	m3dx_grid_entry* stableEntries;   /* stableEntryCount, sorted */
	int32_t* stableOversized;         /* stableOversizedCount */
} m3dx_snapshot;

/* Fills snapshot->header always and every non-NULL array. Returns 0 on
 * success, -1 on error (see m3dx_last_error). */
M3D_API int m3dx_world_export(m3d_world* world, m3dx_snapshot* snapshot);

/* Creates a new world whose full state equals the snapshot. NULL on error. */
M3D_API m3d_world* m3dx_world_import(const m3dx_snapshot* snapshot);

/* ------------------------------------------------- per-stage views */

/* Sorted grid tiers as of the last broadphase refresh. tier 0 = active,
 * 1 = stable. Returns the entry count; copies at most cap entries. */
M3D_API int m3dx_debug_grid_entries(m3d_world* world, int tier, m3dx_grid_entry* out, int cap);
/* Oversized (coarse) list of a tier, in list order. */
M3D_API int m3dx_debug_oversized(m3d_world* world, int tier, int32_t* out, int cap);
/* Island of every body slot at the last solve (-1: not a dynamic body, or no
 * awake dynamic body that step). Returns the island count. islandAwake gets
 * one flag per island when non-NULL. */
M3D_API int m3dx_debug_islands(m3d_world* world, int32_t* islandOfBody, int bodyCap, uint8_t* islandAwake,
							   int islandCap);
/* Graph color of every contact at the last solve: -1 not in a colored
 * island, 0..7 a color, 8 the serial overflow pass. Returns contact count. */
M3D_API int m3dx_debug_contact_colors(m3d_world* world, int32_t* colorOfContact, int cap);

/* ------------------------------------------------------ counters */

typedef struct m3dx_step_stats
{
	int32_t bodyCapacity;
	int32_t awakeDynamicBodies;   /* bodies integrated per substep */
	int32_t contacts;
	int32_t solvedContacts;       /* touching, non-sensor, in awake islands */
	int32_t solvedPoints;         /* manifold points of solved contacts */
	int32_t boxBoxPoints;         /* points of solved box-box contacts */
	int32_t regeneratedManifolds; /* narrowphase regenerations this step */
	int32_t activeEntries;
	int32_t stableEntries;
	int32_t refattedBodies;
	int32_t islands;
	int32_t coloredIslands;
	int32_t sweeps;               /* grid-wide Gauss-Seidel sweeps per iteration (giant islands) */
	int32_t colorRounds;          /* rounds of the giant-island coloring and level fixpoints */
} m3dx_step_stats;

M3D_API void m3dx_world_stats(m3d_world* world, m3dx_step_stats* out);

/* ------------------------------------------------------- kernel timing */

/* Kernels timed with HIP events on the world's stream when timing is on. */
typedef enum m3dx_kernel
{
	M3DX_K_GIANT_NORMALS = 0, /* giant islands: normal projection passes */
	M3DX_K_GIANT_FRICTION,    /* giant islands: stiction pass */
	M3DX_K_GIANT_VELOCITY,    /* giant islands: velocity pass */
	M3DX_K_GIANT_ORDER,       /* giant islands: coloring and level fixpoints */
	M3DX_K_JOINT_ISLANDS,     /* islands with joints (one lane each) */
	M3DX_K_NARROW,            /* narrowphase */
	M3DX_K_PAIRS,             /* pair counting and emission */
	M3DX_K_ADJOINT,           /* differentiable worlds: a step's forward replay and reverse */
	M3DX_K_ADJ_FORWARD,       /* differentiable worlds: forward replays inside a checkpoint segment */
	M3DX_K_GIANT_PREPARE,     /* solver: per-substep contact preparation */
	M3DX_K_GIANT_BODIES,      /* solver: body integration and velocity reconstruction */
	M3DX_K_SORT,              /* radix sorts of every stage (histogram, scan and scatter passes) */
	M3DX_K_COUNT
} m3dx_kernel;

typedef struct m3dx_kernel_time
{
	int32_t launches;   /* launches since timing was (re)enabled */
	float ms;           /* summed event time */
	double bytes;       /* summed algorithmic bytes (0: no byte model) */
} m3dx_kernel_time;

/* enable != 0 starts (and resets) timing, 0 stops it. Returns 0 or -1. */
M3D_API int m3dx_set_kernel_timing(m3d_world* world, int enable);
/* Workgroups of the giant-island solver passes: 0 (default) lets the
   world's tuner choose them by measurement, per size class of the step's
   solver entries; > 0 forces a count. Results do not depend on it.
   Returns 0 or -1. */
M3D_API int m3dx_set_pass_blocks(m3d_world* world, int blocks);
/* Workgroups the last step's solver passes ran with (0: none ran). */
M3D_API int m3dx_pass_blocks(m3d_world* world);
/* Steps of this world that stopped on an error (m3dx_last_error holds the
   last message; m3d_world_step has no return value). */
M3D_API int32_t m3dx_step_failures(m3d_world* world);
/* A batch runs its island-solve substep launches (integration, prepare,
   records, the cooperative passes) as one hipGraph once the passes'
   workgroup count is settled; enable = 0 launches them directly. Results do
   not depend on it. Returns 0 or -1. */
M3D_API int m3dx_set_step_graphs(m3d_world* world, int enable);
/* Steps whose substep loop ran as a graph. */
M3D_API int64_t m3dx_step_graph_launches(m3d_world* world);
/* Grid barriers of the cooperative solver passes in XCD mode (default on
   where the device's XCC_ID register names its 8 XCDs, checked once per
   device; M3D_XCD_BARRIER=0 turns it off for the process): the workgroups of
   a launch are grouped by the XCD they run on, and each barrier writes back
   and invalidates each XCD's L2 once instead of once per workgroup. enable =
   0 fences every workgroup at the agent scope. Results do not depend on it.
   Returns 0 or -1. */
M3D_API int m3dx_set_xcd_barrier(m3d_world* world, int enable);
/* Launches of the grid-barrier kernels (the coloring fixpoint and the
   solver passes). cooperative = 0 (the default): plain launches, sized to
   what every CU's occupancy holds at once so the workgroups are co-resident;
   1: cooperative launches (M3D_COOP_LAUNCH=1 sets this for every world).
   waitTicks > 0 bounds every grid-barrier wait, in ticks of the device's
   100 MHz wall clock (default 2^28, about 2.7 s); a wait beyond it fails the
   step (m3d_world_step counts the failure and the error names it; the step's
   solve is void) and turns the world to cooperative launches. 0 keeps the
   bound. Returns 0 or -1. */
M3D_API int m3dx_set_grid_launch(m3d_world* world, int cooperative, uint32_t waitTicks);
/* 1 when the world launches its grid kernels cooperatively, 0 when not, -1
   for no world. */
M3D_API int m3dx_grid_launch_mode(m3d_world* world);
/* Copies the M3DX_K_COUNT accumulators. Returns M3DX_K_COUNT. */
M3D_API int m3dx_kernel_timing(m3d_world* world, m3dx_kernel_time* out);

/* ------------------------------------------------------ batched worlds */

/* A batch: one device world holding `count` independent sub-worlds that share
   the world definition (no counterpart in the reference, which steps one world
   per call). m3d_world_step steps every sub-world; each steps exactly as a
   world of its own with the same bodies would (same pairs, contact order,
   islands and results). Bodies are created with m3d_body_create into the
   selected sub-world, sub-worlds filled in ascending order; body ids and
   every m3d_body_* getter and setter work as for a plain world.
   m3d_body_destroy destroys in the body's sub-world as in a world of its
   own (its contacts leave the sub-world's run in the reference's order, the
   slot goes to the sub-world's free list, which creates into that sub-world
   reuse while it is the selected one). Contact events hold those of every
   sub-world: the events of one sub-world's bodies, in the batch's order,
   are the events its own world would report, in its order. Not supported
   on a batch: joints between sub-worlds (creation returns the null id) and
   m3dx_body_respawn's destroys. A batch exports
   its full state (contacts grouped by sub-world, see
   m3dx_batch_contact_counts) for inspection; importing it gives a plain
   world. Queries see the bodies of all sub-worlds and use the broadphase of
   the last step.
   count in [1, 262144]; NULL on error. */
M3D_API m3d_world* m3dx_batch_create(const m3d_world_def* def, int32_t count);
/* Sub-world that m3d_body_create adds to; never below the current one.
   Returns 0, or -1 (see m3dx_last_error). */
M3D_API int m3dx_batch_select(m3d_world* world, int32_t sub);
/* Number of sub-worlds (0 for a plain world). */
M3D_API int32_t m3dx_batch_size(m3d_world* world);
/* Body index range [first, first + count) of a sub-world. Returns 0 or -1. */
M3D_API int m3dx_batch_body_range(m3d_world* world, int32_t sub, int32_t* first, int32_t* count);
/* Live contacts of each sub-world (up to cap); returns the sub-world count. */
M3D_API int m3dx_batch_contact_counts(m3d_world* world, int32_t* out, int32_t cap);

/* ---------------------------------------------- differentiable worlds */

/* Adjoint of the XPBD solve (no counterpart in the reference). While
   recording, every step stores its solve inputs: body state, mass and
   friction, and the solver contacts with their manifolds (the fixed contact
   set of that step). The adjoint is that of the step's solve (the four
   substeps of integrate, prepare, normal projection, stiction, velocity
   reconstruction and velocity pass, in the order the engine ran them: the
   serial order of src/xpbd.cpp:898-974, or for an island of 300 or more
   solver contacts the colored schedule of src/xpbd.cpp:1167-1369 with its
   packet arithmetic and overflow pass) with the contact set, manifolds and
   coloring held fixed; a recorded step's output state is the next step's
   input. Requires a world created with enableSleep = false; bodies must not
   be created or destroyed while recording. Plain worlds and batches (one
   lane per sub-world) both work. A step the adjoint does not cover (a world
   with joints) is refused: that step runs as usual but is not recorded,
   recording stops, and m3dx_diff_record(world, 0) and every m3dx_diff_*
   call on the records return -1 with the reason in m3dx_last_error. */
/* enable != 0 starts recording (dropping earlier records), 0 stops it.
   Returns 0 or -1. */
M3D_API int m3dx_diff_record(m3d_world* world, int enable);
/* Number of recorded steps. */
M3D_API int32_t m3dx_diff_steps(m3d_world* world);
/* State recorded at the start of a step (13 floats per body: x q v w).
   Returns the step's solver contact count, or -1. */
M3D_API int m3dx_diff_state(m3d_world* world, int step, float* out);
/* Replays recorded steps first .. first + count - 1 in sequence from the
   state recorded at `first` (or stateIn: 13 values per body, x q v w),
   with per-body mass and friction (null: as recorded). precision 0 is the
   engine's float arithmetic, 1 double; arrays are host memory of that type.
   stateOut receives the final state (13 per body). Returns 0 or -1. */
M3D_API int m3dx_diff_replay(m3d_world* world, int first, int count, int precision, const void* stateIn,
                             const void* mass, const void* friction, void* stateOut);
/* Vector-Jacobian product of all recorded steps: seed is the adjoint of the
   final state (13 floats per body: x q v w); returns the adjoint of the first
   step's input state (13 per body; its v part is the gradient with respect to
   the initial linear velocities) and the gradients with respect to each
   body's mass and friction. Output pointers may be null. Returns 0 or -1.
   The parallel adjoint: per world and step, the solve's ops in reverse level
   order, each op's vector-Jacobian product from forward tangents; states are
   recomputed from a checkpoint every m3dx_diff_set_checkpoint steps. Worlds
   of up to 3072 solver contacts and 2048 bodies. */
M3D_API int m3dx_diff_backward(m3d_world* world, const float* seed, float* stateAdj, float* massGrad,
                               float* frictionGrad);
/* m3dx_diff_backward with double arrays: the adjoint is evaluated in double
   precision along the engine's float path (float values, double
   derivatives); the float entry point rounds its results. */
M3D_API int m3dx_diff_backward_f64(m3d_world* world, const double* seed, double* stateAdj, double* massGrad,
                                   double* frictionGrad);
/* The same product by a reverse-mode tape in float walked by one lane per
   world (a second, independent evaluation, for checking). */
M3D_API int m3dx_diff_backward_tape(m3d_world* world, const float* seed, float* stateAdj, float* massGrad,
                                    float* frictionGrad);
/* Steps between body-state checkpoints of the backward pass (default 10).
   Returns 0 or -1. */
M3D_API int m3dx_diff_set_checkpoint(m3d_world* world, int every);
/* Recorded steps first .. first + count - 1 replayed by the parallel
   adjoint's forward kernels (float, the engine's arithmetic) from the state
   recorded at `first`; stateOut receives the final state (13 floats per
   body). Returns 0 or -1. */
M3D_API int m3dx_diff_forward(m3d_world* world, int first, int count, float* stateOut);
/* Forward tangents of recorded steps first .. first + count - 1 on the host
   CPU along one direction (dStateIn, dMass, dFriction) from stateIn with
   per-body mass and friction (null: the recorded values); derivatives in
   double. primal 0: values in float, the engine's own path; 1: values in
   double (the FP64 replay). stateOut and dStateOut receive the final state
   and its directional derivative (13 doubles per body). Host arrays.
   Returns 0 or -1. */
M3D_API int m3dx_diff_jvp_cpu(m3d_world* world, int first, int count, int primal, const double* stateIn,
                              const double* dStateIn, const double* mass, const double* dMass,
                              const double* friction, const double* dFriction, double* stateOut, double* dStateOut);

/* ---------------------------------------------- scene queries */

/* Batched ray casts on the device (no counterpart in the reference, which
   casts one ray per call): out[k] = m3d_world_ray_cast_filtered(world,
   origins[k], translations[k], maskBits ? maskBits[k] : 0xFFFFFFFF),
   evaluated as the reference's ray cast (include/melon3d.h:510-514,
   src/world.cpp:1081-1200: coarse lists, then the grid cells along the
   segment) over the device grid, one lane per ray. Plain worlds only.
   Returns 0 or -1. */
M3D_API int m3dx_world_ray_cast_batch(m3d_world* world, const m3d_vec3* origins, const m3d_vec3* translations,
                                      const uint32_t* maskBits, int32_t count, m3d_ray_result* out);
/* m3dx_world_ray_cast_batch that allocates nothing (no counterpart in the
   reference; for callers that cannot allocate device memory: code under
   HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
   Device staging lives in a fixed per-GPU array of the library (4 MiB,
   held until the process ends), host staging in fixed chunks on the stack.
   At most 65536 rays per call (more is an error). It never refreshes the
   broadphase (a refresh may grow device buffers), so it refuses (-1) where
   the reference's ray cast would refresh it with an effect: after any body
   or API change since the last step or query, and after a step that left
   the stable tier dirty until a refreshing query (m3d_world_ray_cast,
   m3dx_world_ray_cast_batch) has run. Calls that overlap on one GPU share
   the array and race. Same results as m3dx_world_ray_cast_batch. Returns 0
   or -1. */
M3D_API int m3dx_world_ray_cast_batch_noalloc(m3d_world* world, const m3d_vec3* origins,
                                              const m3d_vec3* translations, const uint32_t* maskBits,
                                              int32_t count, m3d_ray_result* out);
/* Batched sphere casts on the device: out[k] = m3d_world_sphere_cast(world,
   origins[k], radii[k], translations[k], maskBits ? maskBits[k] :
   0xFFFFFFFF) (include/melon3d.h:515-518, src/world.cpp:1286-1392), with
   candidates from the device grid. Plain worlds only. Returns 0 or -1. */
M3D_API int m3dx_world_sphere_cast_batch(m3d_world* world, const m3d_vec3* origins, const float* radii,
                                         const m3d_vec3* translations, const uint32_t* maskBits, int32_t count,
                                         m3d_ray_result* out);

/* Transforms of many bodies in one call (no counterpart in the reference,
   which returns one per call): out[k] = m3d_body_transform(world, ids[k])
   (include/melon3d.h:470, src/world.cpp:634-645), gathered on the device (a frame of a scene dump,
   see melon3d_amd/scene_dump.py). Returns 0 or -1. */
M3D_API int m3dx_world_transforms(m3d_world* world, const m3d_body_id* ids, int32_t count, m3d_transform* out);

/* ------------------------------------------------- bulk spawn/despawn */

/* Bulk despawn and spawn (no counterpart in the reference, which creates and
   destroys one body per call, src/world.cpp:512-615): the same result as the
   calls, in this order,
     for k in 0 .. count-1:
       m3d_body_destroy(world, destroy[k])                        (destroy != NULL)
       ids[k] = m3d_body_create(world, &bodyDefs[k], &shapeDefs[k]) (bodyDefs != NULL)
   - the same ids and slot reuse (LIFO free list), contact removal order,
   end events, wakes and joint removals - with the contacts of the destroyed
   bodies found on the device instead of through a host copy of every
   contact. Destroys are not supported on batches. Returns 0 or -1. */
M3D_API int m3dx_body_respawn(m3d_world* world, const m3d_body_id* destroy, const m3d_body_def* bodyDefs,
                              const m3d_shape_def* shapeDefs, int32_t count, m3d_body_id* ids);

/* Many AABB overlap queries in one call on the device (no counterpart in
   the reference, which runs one query per call with a callback,
   include/melon3d.h:519-522, src/world.cpp:1394-1416): the bodies whose
   tight box overlaps boxes[k], in ascending body index (the order the
   reference's callback sees them), are out[offsets[k] .. offsets[k + 1]).
   offsets (count + 1 entries) is always filled; at most cap ids are written
   (out may be NULL with cap 0, to size a second call). Returns the total
   number of results, or -1. */
M3D_API int m3dx_world_overlap_aabb_batch(m3d_world* world, const m3d_aabb* boxes, int32_t count,
                                          int32_t* offsets, m3d_body_id* out, int32_t cap);

/* Stable sort of n (key, value) pairs by the low `bits` bits of the keys
   (keys below 2^bits), ascending: the device radix sort every stage of the
   step uses, on host arrays, on the device m3dx_set_device names. No
   counterpart in the reference (its stages sort on the host); a hook for
   testing the sort. Returns 0 or -1. */
M3D_API int m3dx_sort_pairs(const uint64_t* keys, const uint32_t* vals, int64_t n, int bits, uint64_t* keysOut,
                            uint32_t* valsOut);

/* Last error message of this thread ("" when none). */
M3D_API const char* m3dx_last_error(void);

/* Device the next m3d_world_create binds to (default 0). */
M3D_API int m3dx_set_device(int device);

#ifdef __cplusplus
}
#endif

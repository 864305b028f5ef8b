/*
 * clothsim.h — C-ABI of the MI355X-native cloth step (libclothsim.so).
 *
 * This is the drop-in boundary for the reference's `--gpu` per-step hot path.
 * The reference has no FFI of its own: the hot path sits behind C++ methods
 * selected by the global `gpu` flag.  Each entry point below names the
 * reference interface it replaces (paths relative to the reference root).
 *
 * Conventions
 *  - Plain C types only: pointers, sizes, POD structs.  No torch types.
 *  - Every call returns cs_status; nothing calls exit()/abort().
 *  - The caller owns every host buffer; the context owns every device buffer
 *    and one HIP stream.  Calls on one context are serialised on that stream
 *    and return after the stream completes.  A context is not thread-safe;
 *    distinct contexts may be used from distinct threads.
 *  - Indices are 0-based.  Global node/face/edge numbering is: all cloths in
 *    creation order, then all obstacles in creation order (the order in which
 *    the reference's MemoryPool hands out records, src/MemoryPool.cu:79-116),
 *    and within a mesh: nodes and faces in OBJ order, edges sorted by their
 *    (min node, max node) pair (src/Mesh.cu:405-447).
 */
#ifndef CLOTHSIM_H
#define CLOTHSIM_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum cs_status {
    CS_OK = 0,
    CS_E_ARG = 1,              /* invalid argument / size */
    CS_E_HIP = 2,              /* HIP runtime error */
    CS_E_OOM = 3,              /* device allocation failed */
    CS_E_NOT_CONVERGED = 4,    /* PCG hit pcg_max_iters (the step still completes) */
    CS_E_COLLISION_FAILED = 5, /* reference prints "Collision step failed!" and exit(1)s: src/Simulator.cu:668-671 */
    CS_E_FORMAT = 6,           /* malformed mesh / material input */
    CS_E_STATE = 7,            /* call not valid in the context's current state */
    CS_E_CAPACITY = 8,         /* caller buffer or internal fixed limit too small */
    CS_E_NO_DEVICE = 9         /* no usable gfx950 device */
} cs_status;

typedef struct cs_ctx cs_ctx; /* opaque context: device buffers + one HIP stream */

/* PCG preconditioners.  BLOCK_JACOBI: inverse 3x3 diagonal blocks (the reference uses the
 * scalar diagonal, src/Cloth.cu:766-809).  AMG: one V(1,1) cycle of aggregation multigrid
 * with damped block-Jacobi smoothing.  LEARNED: M^-1 = G^T G with G emitted per solve by a
 * message-passing network from the assembled system (weights set with
 * cs_set_learned_weights; CS_E_STATE without them).  AMG_LEARNED: the AMG cycle with the
 * learned G^T G, damped by a power-iteration estimate of its spectrum, as the finest level's
 * smoother (cloths too small for a hierarchy fall back to block Jacobi).  None but
 * BLOCK_JACOBI has a counterpart in the reference; they change the iteration count, not the
 * system solved. */
enum { CS_PRECOND_BLOCK_JACOBI = 0, CS_PRECOND_LEARNED = 1, CS_PRECOND_AMG = 2, CS_PRECOND_AMG_LEARNED = 3 };
enum { CS_MESH_CLOTH = 0, CS_MESH_OBSTACLE = 1 };

/* cs_params.flags */
enum {
    CS_FLAG_NO_PROXIMITY = 1, /* skip repulsion/friction proximities (testing only; the reference always runs them) */
    CS_FLAG_NO_COLLISION = 2, /* cs_step skips the collision step (testing only) */
    CS_FLAG_TIME_KERNELS = 4, /* sample HIP-event durations of the hot kernels (see cs_kernel_time) */
    CS_FLAG_INSTANCES = 8,    /* independent instances: cloth i collides only with itself and with obstacle i
                                 (the reference traverses every cloth pair, src/Simulator.cu:179-196, which finds
                                 nothing between instances placed apart), and with block-Jacobi PCG or
                                 CS_PRECOND_AMG all cloths' solves advance together in batched launches, each
                                 cloth with its own multigrid hierarchy, one launch per level for all of them
                                 (no counterpart) */
    CS_FLAG_FIXED_TUNING = 16, /* batched solves keep their default launch shape instead of autotuning it
                                 (bit-reproducible runs; no counterpart) */
    CS_FLAG_GRAPH_CYCLES = 32, /* multigrid cycles replayed from HIP graphs captured once per solve instead of
                                 launched kernel by kernel (the same kernels in the same order; no counterpart) */
    CS_FLAG_PERSISTENT_BATCH = 64 /* the batched block-Jacobi solve (CS_FLAG_INSTANCES) runs the iterations between
                                 two host polls as one cooperative launch, a grid barrier between the two phases
                                 of an iteration, instead of two launches per iteration (the same results bit
                                 for bit; no counterpart) */
};

/* kernel ids for cs_kernel_time.  CS_KERNEL_EVENT_PAIR times an empty event pair recorded on
 * the stream just before each sampled SpMV: the marker latency included in every sample, which
 * a caller subtracts to get the kernel's own duration. */
enum {
    CS_KERNEL_PCG_SPMV = 0,
    CS_KERNEL_ASSEMBLE_ROWS = 1,
    CS_KERNEL_FACE_RECORDS = 2,
    CS_KERNEL_EVENT_PAIR = 3,
    CS_KERNEL_COUNT = 4
};

/* Scene-level parameters.  Defaults (cs_default_params) are the reference's:
 * Magic ctor src/Magic.cu:5-17, Wind ctor src/Wind.cu:3-7, Simulator ctor
 * src/Simulator.cu:19-27, MAX_COLLISION_ITERATION src/CollisionHelper.cuh:21. */
typedef struct cs_params {
    float dt;
    float gravity[3];
    float wind_density, wind_drag, wind_velocity[3];
    float handle_stiffness, collision_stiffness;
    float repulsion_thickness, collision_thickness;
    float friction, obstacle_friction;
    float pcg_rel_tol;     /* stop when ||b - A dv|| <= pcg_rel_tol * ||b||   (default 1e-5): the loop stops on
                              its recursive residual, and the stop stands only if the true residual (row sums
                              in FP64) confirms it; otherwise the loop continues from the true residual, at most
                              4 times, and the step reports CS_E_NOT_CONVERGED if it still misses */
    int pcg_max_iters;     /* default 4000 */
    int precond;           /* CS_PRECOND_* */
    int max_collision_iters; /* default 100 */
    int flags;
} cs_params;

/* One triangle mesh, already placed in world space (the reference applies the
 * JSON "transform" while reading the OBJ, src/Mesh.cu:320-338).
 * x: 3*n_nodes, v: 3*n_nodes or NULL (zero), uv: 2*n_verts (material space, metres),
 * face_nodes / face_verts: 3*n_faces. */
typedef struct cs_mesh_in {
    int n_nodes;
    const float* x;
    const float* v;
    int n_verts;
    const float* uv;
    int n_faces;
    const int* face_nodes;
    const int* face_verts;
} cs_mesh_in;

/* Material JSON content plus the cloth's "materials" multipliers
 * (src/Material.cu:3-41). */
typedef struct cs_material_in {
    float density;
    float stretching[6][4];
    float bending[3][5];
    float thicken, density_mult, stretching_mult, bending_mult;
} cs_material_in;

/* Per-call statistics. */
typedef struct cs_stats {
    int pcg_iters;            /* summed over cloths */
    float pcg_rel_residual;   /* max over cloths of ||r||/||b|| at exit (recursive residual) */
    int pcg_converged;        /* 1 if every cloth reached pcg_rel_tol */
    int n_proximities;
    int n_contact_blocks;     /* dynamic 3x3 blocks merged into the system */
    int collision_iters;      /* impact-response iterations of the last collision step */
    int collision_deform;     /* 0 or 1: the deform pass that succeeded */
    int n_impacts;            /* impacts resolved in the last collision step */
    /* host wall time (ms) of the stages: tree builds of both stages (refits inside the collision
     * loop count to ms_collision), proximity search, assembly, solve, position/geometry update,
     * collision loop.  Every step entry point zeroes the whole struct first. */
    float ms_bvh, ms_proximity, ms_assemble, ms_solve, ms_update, ms_collision;
} cs_stats;

/* One repulsion/friction record (src/Proximity.cuh:10-21), nodes are global ids. */
typedef struct cs_proximity_rec {
    int nodes[4];
    float n[3];
    float mu, stiffness;
    float w[4];
} cs_proximity_rec;

/* One CCD impact (src/Impact.cuh:7-15), computed in FP64.  type: 0 vertex-face
 * (nodes[0] is the vertex), 1 edge-edge. */
typedef struct cs_impact_rec {
    int type;
    int nodes[4];
    double t;
    double w[4];
    double n[3];
} cs_impact_rec;

/* ---- context lifetime (replaces Simulator::Simulator / ~Simulator, src/Simulator.cu:3-135) ---- */
const char* cs_version(void);
void cs_default_params(cs_params* p);
cs_status cs_create(const cs_params* params, cs_ctx** ctx_out);
void cs_destroy(cs_ctx* ctx);
const char* cs_last_error(const cs_ctx* ctx);
cs_status cs_set_params(cs_ctx* ctx, const cs_params* params);

/* ---- scene construction ---- */
/* replaces Cloth::Cloth(json, pool) (src/Cloth.cu:3-23) + Mesh::load (src/Mesh.cu:306-452) */
cs_status cs_add_cloth(cs_ctx* ctx, const cs_mesh_in* mesh, const cs_material_in* material, int* cloth_id);
/* replaces Obstacle::Obstacle(json, transformations, pool) (src/Obstacle.cu:3-11); static until
 * cs_set_obstacle_motion gives it a motion */
cs_status cs_add_obstacle(cs_ctx* ctx, const cs_mesh_in* mesh, int* obstacle_id);
/* replaces Cloth::addHandles(nodeIndices, motionIndex=-1, ...) (src/Cloth.cu:649-673); anchors = current x */
cs_status cs_add_handles(cs_ctx* ctx, int cloth_id, const int* nodes, int n);

/* ---- motions (replaces Motion, src/Motion.cu:3-72, and its users: Obstacle::transform/step,
 * src/Obstacle.cu:44-85, src/ObstacleHelper.cu:10-26; handle anchors, src/ClothHelper.cu:3-14,
 * src/PhysicsHelper.cu:379-396; Simulator::updateTransformations, src/Simulator.cu:565-569) ---- */
#define CS_MAX_MOTIONS 64
/* Transformation (src/Transformation.cuh:12-16): p -> scale * rotate(p) + translate with the
 * quaternion (s, x, y, z) = rotate[0..3] applied as x (s^2 - |v|^2) + 2 v (v.x) + 2 s v x x
 * (src/Quaternion.cu:73-75; not renormalised, as in the reference). */
typedef struct cs_transform {
    float scale;
    float translate[3];
    float rotate[4];
} cs_transform;
/* Motion(json) (src/Motion.cu:3-20): n >= 1 keyframes at increasing times; a first key after
 * t = 0 gets an identity key at t = 0 in front.  The transformation at time t is the cubic
 * Hermite of each component between the bracketing keys, with tangents
 * (x[i+1] - x[i-1]) / (t[i+1] - t[i-1]) and zero at both ends, and the last key after the last
 * time (src/Motion.cu:45-72).  Fails with CS_E_CAPACITY beyond CS_MAX_MOTIONS. */
cs_status cs_add_motion(cs_ctx* ctx, const float* times, const cs_transform* keys, int n, int* motion_id);
/* Motion::computeTransformation for the keyframes (times, keys, n) at `time`, without a
 * context (host code only; the steps evaluate their motions with the same routine). */
cs_status cs_motion_eval(const float* times, const cs_transform* keys, int n, float time, cs_transform* out);
/* Obstacle(json) with "motion" (src/Obstacle.cu:3-10,44-64): the obstacle's current positions
 * become its base shape, and x = x0 = T(t)(base) now, T = the motion at the context's time
 * t = step count * dt.  Every physics step then sets v = (T(t)(base) - x) / dt on its nodes
 * before the proximity search (src/Obstacle.cu:66-85).  Before the first step only. */
cs_status cs_set_obstacle_motion(cs_ctx* ctx, int obstacle_id, int motion_id);
/* Cloth::addHandles with a motion (src/Cloth.cu:653-673): anchor = T(t)^-1 (x) now; every
 * physics step pulls the node towards T(t)(anchor).  Before the first step only. */
cs_status cs_add_moving_handles(cs_ctx* ctx, int cloth_id, const int* nodes, int n, int motion_id);
/* Simulator::nSteps: each cs_physics_step first increments it and evaluates the motions at
 * nSteps * dt (src/Simulator.cu:838-843,565-569).  Setting it also places every moving
 * obstacle at T(nSteps * dt)(base), x = x0, as Simulator::resume does (src/Simulator.cu:1045-1053). */
cs_status cs_set_step_count(cs_ctx* ctx, int steps);
cs_status cs_get_step_count(const cs_ctx* ctx, int* steps);

/* ---- stepping ---- */
/* replaces Simulator::physicsStep (src/Simulator.cu:571-600) incl. Cloth::physicsStep GPU (src/Cloth.cu:695-813) */
cs_status cs_physics_step(cs_ctx* ctx, cs_stats* stats);
/* replaces Simulator::collisionStep (src/Simulator.cu:602-684) */
cs_status cs_collision_step(cs_ctx* ctx, cs_stats* stats);
/* physics then collision: the per-step part of Simulator::simulateStep (src/Simulator.cu:838-897) */
cs_status cs_step(cs_ctx* ctx, cs_stats* stats);
/* run `steps` steps back to back with no host readback in between (benchmarks) */
cs_status cs_run(cs_ctx* ctx, int steps, cs_stats* stats_accum);

/* ---- state exchange (replaces Mesh::save/load node I/O, src/Mesh.cu:454-507) ---- */
cs_status cs_get_counts(cs_ctx* ctx, int kind, int id, int* n_nodes, int* n_faces, int* n_edges);
cs_status cs_get_state(cs_ctx* ctx, int kind, int id, float* x, float* x0, float* v);
cs_status cs_set_state(cs_ctx* ctx, int kind, int id, const float* x, const float* x0, const float* v);
/* The faces and material-space vertices Mesh::save writes (src/Mesh.cu:454-507): face node ids
 * (3 nf) and face uv-vertex ids (3 nf), both mesh-local, and the uv vertices (2 n_verts).  NULL
 * buffers query n_verts only.  Host data: needs no device. */
cs_status cs_get_topology(cs_ctx* ctx, int kind, int id, int* face_nodes, int* face_verts, float* uv, int* n_verts);
/* Asynchronous frame download, double-buffered (no counterpart: Simulator::save copies every
 * cloth's nodes synchronously, src/Simulator.cu:950-957 through src/Cloth.cu:865-887 and
 * src/Mesh.cu:454-507).  cs_frame_begin snapshots the positions and velocities of every cloth
 * node into device buffer `slot` (0 or 1) on the context's stream, starts their download into
 * pinned host memory on a copy stream, and returns without waiting: the steps that follow run
 * while the download is in flight.  Beginning a slot again first waits (on the device) for its
 * previous download.  cs_frame_end waits for slot's download and copies cloth `cloth`'s nodes
 * of the snapshot (3 floats per node into x and v, either may be NULL; n_nodes receives the
 * count, which a remesh after cs_frame_begin does not change).  cs_frame_end touches only the
 * slot, so it may run on another host thread while the context steps, unless that slot is being
 * begun at the same time.  CS_E_STATE: the slot was never begun. */
cs_status cs_frame_begin(cs_ctx* ctx, int slot);
cs_status cs_frame_end(cs_ctx* ctx, int slot, int cloth, float* x, float* v, int* n_nodes);
/* node normal (3n), area (n), mass (n) as maintained by Mesh::updateNodeGeometries (src/Mesh.cu:182-214) */
cs_status cs_get_geometry(cs_ctx* ctx, int kind, int id, float* normals, float* area, float* mass);

/* ---- sub-stage entry points (parity tests run them on the context's current state) ---- */
/* Material::Material stretching table (30^3 x 4) and bending table (3x5) (src/Material.cu:3-85) */
cs_status cs_material_table(const cs_material_in* material, float* table, float* density, float* bending);
/* BVH(mesh, ccd) + Simulator::traverse (src/BVH.cu:3-51, src/Simulator.cu:179-196): candidate face pairs
 * (global face ids, first = query face) over cloth self, cloth x earlier cloth, cloth x obstacle.
 * mode: CS_LBVH_CCD builds the trees over the swept boxes of x0 -> x; with CS_LBVH_REFIT the trees of
 * the previous call (same CS_LBVH_CCD bit) are refit to the current positions instead of rebuilt
 * (BVH::update, src/BVH.cu:167-177; CS_E_STATE when there are none). */
enum { CS_LBVH_CCD = 1, CS_LBVH_REFIT = 2 };
cs_status cs_lbvh_pairs(cs_ctx* ctx, int mode, float thickness, int* pairs, int cap, int* n_out);
/* Simulator::findProximitiesGpu (src/Simulator.cu:300-349) */
cs_status cs_find_proximities(cs_ctx* ctx, cs_proximity_rec* out, int cap, int* n_out);
/* Assembled implicit-Euler system of one cloth after proximity detection (src/Cloth.cu:695-764):
 * A as BSR (3x3 row-major blocks, rows sorted, columns sorted and unique), b (3n).
 * Pass NULL buffers to query nnzb only. */
cs_status cs_assemble(cs_ctx* ctx, int cloth_id, int* row_ptr, int* col_idx, float* vals, int cap_blocks,
                      int* nnzb_out, float* b);
/* Solve the last assembled system of one cloth (src/Cloth.cu:766-809): dv (3n). */
cs_status cs_solve(cs_ctx* ctx, int cloth_id, float* dv, cs_stats* stats);
/* ---- solve without allocation.  No counterpart in the reference: for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  The block-Jacobi PCG of cs_solve (x = 0, same iteration, stop when
 * ||r|| <= rel_tol ||b|| or after max_iters) on a block-CSR system in the caller's DEVICE memory:
 * row_ptr (n_nodes + 1), col_idx (nnzb), vals (9 per block, 3x3 row-major; every row holds its
 * diagonal block), b (3 n_nodes) in, x (3 n_nodes) out; iters and rel_residual (the recursive
 * ||r|| / ||b||) are host pointers and may be NULL.  Runs on `stream` (a hipStream_t, NULL for
 * the default stream) and returns after it has completed.  Allocates nothing: every temporary
 * lives in one fixed device array of the library (about 5.5 MB per GPU, held until the process
 * ends), shared by every call on that GPU, so calls that overlap on one GPU race.
 * CS_E_CAPACITY above CS_NOALLOC_MAX_NODES nodes; CS_E_ARG for a row without its diagonal
 * block or with a column outside [0, n_nodes); CS_E_NOT_CONVERGED when max_iters ends the
 * solve. */
#define CS_NOALLOC_MAX_NODES 65536
cs_status cs_solve_bsr_noalloc(void* stream, int n_nodes, const int* row_ptr, const int* col_idx, const float* vals,
                               const float* b, float* x, float rel_tol, int max_iters, int* iters,
                               float* rel_residual);
/* The CS_PRECOND_AMG, CS_PRECOND_LEARNED or CS_PRECOND_AMG_LEARNED preconditioner of the last
 * assembled system (no counterpart in the reference, whose PCG uses the Jacobi diagonal): the
 * level sizes of the multigrid hierarchy (n_levels = 0 when the cloth cannot use it; always 0
 * for LEARNED) and, when r != NULL, z = M^-1 r (3n each).  CS_E_ARG for BLOCK_JACOBI. */
cs_status cs_precondition(cs_ctx* ctx, int cloth_id, const float* r, float* z, int* level_sizes, int cap,
                          int* n_levels);
/* ---- learned preconditioner (replaces the preconditioner step z = M^-1 r of the reference's
 * PCG, src/PhysicsHelper.cu:494-499 inside src/Cloth.cu:766-809; the network itself has no
 * counterpart in the reference) ----
 * The weights are one float vector of cs_lpc_weight_count() entries in the order the
 * package's weight file and oracle/lpc.py list them.  Setting weights resets the optimiser. */
int cs_lpc_weight_count(void);
cs_status cs_set_learned_weights(cs_ctx* ctx, const float* w, int n);
cs_status cs_get_learned_weights(cs_ctx* ctx, float* w, int n);
/* The factor G of one cloth's last assembled system: diag (9 nn, row-major 3x3 G_ii) and lower
 * (9 per static-pattern entry (i, j < i), rows ascending then columns ascending).  NULL
 * buffers query n_lower only. */
cs_status cs_lpc_factor(cs_ctx* ctx, int cloth_id, float* diag, float* lower, int cap_lower, int* n_lower);
/* One training step on one cloth's last assembled system: the Kaporin loss of the network's
 * factor (plus sol_weight (1 - cos^2) of the energy-norm angle between M^-1 b and the solution,
 * when the system has been solved by cs_solve or a step), its gradient (grad,
 * cs_lpc_weight_count() floats, may be NULL) and, when lr > 0, one Adam step (beta 0.9 / 0.999)
 * on the context's weights. */
cs_status cs_lpc_train_step(cs_ctx* ctx, int cloth_id, float lr, float sol_weight, float* loss, float* grad);
/* cs_lpc_train_step plus smooth_weight |(I - G^T G A) v|_A^2 / |v|_A^2 for the probe v of
 * `seed` (v_i = (+-1, +-1, +-1) from a hash of (i, seed)): the energy one sweep of the learned
 * smoother leaves, the figure a multigrid smoother is judged by (no counterpart in the
 * reference). */
cs_status cs_lpc_train_step_smooth(cs_ctx* ctx, int cloth_id, float lr, float sol_weight, float smooth_weight,
                                   unsigned int seed, float* loss, float* grad);
/* The damping w of the learned smoother (CS_PRECOND_AMG_LEARNED) of one cloth's last assembled
 * system: x += w G^T G (b - A x). */
cs_status cs_lpc_smoother_omega(cs_ctx* ctx, int cloth_id, float* omega);
/* The launch shape of the batched solve of independent instances (CS_FLAG_INSTANCES):
 * rows per workgroup and iterations per host poll in use, and the number of solves its
 * autotuner has timed so far (it settles after 8).  No counterpart in the reference. */
cs_status cs_batch_tuning(const cs_ctx* ctx, int* rows_per_block, int* chunk, int* trials);

/* Simulator::findImpacts (src/Simulator.cu:383-392), deduplicated and sorted by (t, type, nodes). */
cs_status cs_find_impacts(cs_ctx* ctx, float thickness, cs_impact_rec* out, int cap, int* n_out);
/* Simulator::independentImpacts GPU (src/Simulator.cu:422-461) on caller-provided impacts
 * (sorted by t, as cs_find_impacts returns them); writes the selected subset in order. */
cs_status cs_independent_impacts(cs_ctx* ctx, const cs_impact_rec* in, int n, int deform,
                                 cs_impact_rec* out, int* n_out);

/* CollisionOptimization(impacts, collision_thickness, deform, obstacleMass) + Optimization::solve
 * (src/optimization/CollisionOptimization.cu:37-145, src/optimization/Optimization.cu:137-181) on
 * caller-provided impacts: moves the impacts' nodes (free nodes only when deform == 0) from their
 * current positions x, minimising 1/2 invMass sum m |x - x1|^2 (x1: the positions of the last
 * geometry update) subject to sum_j w_j n . x_j >= collision_thickness per impact.  iters (may be
 * NULL): the solve's gradient iterations. */
cs_status cs_impact_response(cs_ctx* ctx, const cs_impact_rec* impacts, int n, int deform, float obstacle_mass,
                             int* iters);

/* ---- instrumentation (no counterpart in the reference, which only prints wall-clock stage
 * times, src/Simulator.cu:845-896) ----
 * With CS_FLAG_TIME_KERNELS set, the context brackets sampled launches of the hot kernels with
 * HIP events on its own stream.  Returns the summed duration (ms) and the number of sampled
 * launches of kernel `kernel_id` (CS_KERNEL_*) since the last reset, and optionally resets them. */
cs_status cs_kernel_time(cs_ctx* ctx, int kernel_id, double* total_ms, int* count, int reset);

/* ---- the reference's scene files (replaces Simulator::Simulator reading the configuration,
 * src/Simulator.cu:3-135, with Cloth ctor src/Cloth.cu:3-40, Obstacle ctor src/Obstacle.cu:3-10,
 * Mesh::load src/Mesh.cu:305-347, Material ctor src/Material.cu:3-41, Magic ctor
 * src/Magic.cu:5-17, Wind ctor src/Wind.cu:3-7) ---- */
typedef struct cs_scene_info {
    float frame_time;      /* "frame_time"; dt = frame_time / frame_steps */
    int frame_steps;       /* "frame_steps" */
    float end_time;        /* "end_time", +inf when absent */
    int end_frame;         /* "end_frame", INT_MAX when absent */
    int n_cloths, n_obstacles, n_handles;
    int n_cloth_nodes, n_cloth_faces;
    int n_motions;           /* "motions": keyframe lists, or the bodies of a "motfile" */
    int n_moving_obstacles;  /* obstacles with a "motion" (every one of an "obstacles": "<pattern>" sequence) */
    int remeshing;           /* 1 unless "disable" lists "remeshing": then every cloth has its "remeshing"
                                block set (cs_set_remeshing) and a frame ends with cs_remesh */
} cs_scene_info;

/* Reads the configuration JSON at config_path; relative mesh and material paths are resolved
 * against root_dir (the reference resolves them against its working directory).  base (or the
 * defaults when NULL) supplies what the file does not hold (PCG settings, flags).  With
 * ctx_out == NULL only parses and fills info.  "motions" may be a list of keyframe lists or a
 * {"motfile", "fps", "transform"} block (src/Simulator.cu:39-79; the .mot file path is resolved
 * against root_dir too); obstacles and handles take a "motion" index; "obstacles":
 * "<printf pattern>" loads <pattern % 0>, <pattern % 1>, ... until a file is missing, obstacle i
 * moving with motion i (src/Simulator.cu:90-101).  Unreadable or malformed files give
 * CS_E_ARG / CS_E_FORMAT.  The message of the last failure on this thread: cs_scene_error(). */
cs_status cs_load_scene(const char* config_path, const char* root_dir, const cs_params* base, cs_ctx** ctx_out,
                        cs_scene_info* info);
/* Simulator::save's configuration file (src/Simulator.cu:950-957, src/Cloth.cu:865-872): the
 * configuration at config_path with cloth i's "mesh" set to cloth_meshes[i] and its "transform"
 * removed, written as JSON to out_path (comments are not kept).  A later cs_load_scene of out_path
 * resumes from those meshes; handle node ids stay valid for cloths that were not remeshed (cs_remesh
 * renumbers nodes, as the reference's remeshing does). */
cs_status cs_write_config(const char* config_path, const char* out_path, const char* const* cloth_meshes,
                          int n_cloths);
/* The transformation of motion `motion_id` of a configuration at `time` (host only: parses the
 * file's motions, no meshes and no context). */
cs_status cs_scene_motion(const char* config_path, const char* root_dir, int motion_id, float time,
                          cs_transform* out);
const char* cs_scene_error(void);

/* ---- remeshing: Cloth::remeshingStep (src/Cloth.cu:815-833), run once per frame by
 * Simulator::remeshingStep (src/Simulator.cu:686-695) when the scene does not disable it ---- */
typedef struct cs_remeshing { /* a cloth's "remeshing" block (src/Remeshing.cu:3-19) */
    float refine_angle, refine_velocity, refine_compression, rib_stiffening;
    float size_min, size_max, aspect_min, flip_threshold;
} cs_remeshing;
/* The block's defaults: refine_* = +inf, rib_stiffening 1, size (-inf, +inf), aspect_min -inf,
 * flip_threshold 1e-2. */
void cs_default_remeshing(cs_remeshing* r);
/* The edge operators of the remeshing step (flip, split, collapse; src/Cloth.cu:268-646,
 * src/Operator.cu:7-217) on one mesh held in host arrays, with the sizing given per face
 * (face_sizing: 4 floats per face, a row-major 2x2 tensor).  Nodes: x, x0, v (3 each), preserve
 * (1: never collapsed away: boundary, seam and handle nodes); uv vertices with their node;
 * faces as uv-vertex triples; edges as node pairs, in list order (a freshly loaded mesh lists its
 * distinct (min, max) node pairs ascending).  flip_count carries the size of the last flip pass
 * between calls (the reference keeps it in a function-level static, src/Cloth.cu:321); start it
 * at 0.  Host only: needs no device.  The result is read with cs_remesh_result_* and freed with
 * cs_remesh_result_free.  CS_E_FORMAT when a face edge is missing from edge_nodes, an edge side
 * has two faces (a non-manifold edge, or neighbours with opposite orientations) or an index is
 * out of range. */
typedef struct cs_remesh_result cs_remesh_result;
cs_status cs_remesh_mesh(const cs_remeshing* r, int n_nodes, const float* x, const float* x0, const float* v,
                         const unsigned char* preserve, int n_verts, const float* uv, const int* vert_node,
                         int n_faces, const int* face_verts, int n_edges, const int* edge_nodes,
                         const float* face_sizing, int* flip_count, cs_remesh_result** out);
cs_status cs_remesh_result_counts(const cs_remesh_result* res, int* n_nodes, int* n_verts, int* n_faces,
                                  int* n_edges);
/* Any pointer may be NULL.  vert_sizing: 4 per vertex; node_origin: the input node of each
 * output node, -1 for a node a split made. */
cs_status cs_remesh_result_get(const cs_remesh_result* res, float* x, float* x0, float* v, unsigned char* preserve,
                               float* uv, int* vert_node, int* face_verts, int* edge_nodes, float* vert_sizing,
                               int* node_origin);
void cs_remesh_result_free(cs_remesh_result* res);
/* Sets (r != NULL) or clears one cloth's remeshing parameters (the Cloth ctor's Remeshing,
 * src/Cloth.cu:3-17).  cs_load_scene sets them for every cloth unless the scene disables
 * remeshing. */
cs_status cs_set_remeshing(cs_ctx* ctx, int cloth_id, const cs_remeshing* r);
/* Simulator::remeshingStep (src/Simulator.cu:686-695): every cloth with remeshing parameters is
 * remeshed at its current state -- sizing on the device from the nearest obstacle planes within
 * 10 repulsion thicknesses, the curvature, normal, velocity and compression terms; flips, splits
 * and collapses on the host -- and the context is rebuilt for the new topology (node, face and
 * edge counts change; handles keep their nodes, under new ids; cs_get_counts / cs_get_topology
 * give the new meshes).  New nodes take the mean position and velocity of their edge.  Then the
 * separation step (Simulator::separationStep, src/Simulator.cu:745-799): face pairs that cross
 * after remeshing are pulled apart along the direction that separated their points before it
 * (an augmented-Lagrangian solve like the impact response's, first with the obstacles fixed,
 * then deformable), and the cloth velocities become (x - x0) / dt.  CS_E_COLLISION_FAILED when
 * crossings remain after 100 solves in both passes (the reference prints and exits).  A no-op
 * without remeshed cloths. */
cs_status cs_remesh(cs_ctx* ctx);
/* Wall-clock split of the last cs_remesh (ms): the sizing field on the device (with its
 * download), the operators on the host, the rebuild of the context, the separation step, and the
 * separation step's solves (instrumentation; no counterpart in the reference).  Any pointer may
 * be NULL. */
cs_status cs_remesh_times(const cs_ctx* ctx, float* sizing_ms, float* operators_ms, float* rebuild_ms,
                          float* separation_ms, int* separation_solves);
/* The separation step alone on the current state (the second half of cs_remesh), with the
 * frame's faces before remeshing given per cloth: faces_per_cloth[c] faces of cloth c, each
 * with its corners' material-space coordinates (backup_uv, 6 floats) and world positions
 * (backup_x, 9 floats), concatenated over the cloths in order (Mesh::backupFaces,
 * src/Mesh.cu:116-128).  For tests and callers that edit meshes themselves; same status as
 * cs_remesh. */
cs_status cs_separation_step(cs_ctx* ctx, const int* faces_per_cloth, const float* backup_uv, const float* backup_x);
/* The sizing the remeshing step would use for one cloth at the current state, with parameters r:
 * face_sizing (4 per face, a row-major 2x2 tensor, faces in cs_get_topology order) and, when not
 * NULL, planes (8 per node: the nearest obstacle point p, 0, the unit normal n, 0; n = 0 where
 * no obstacle is within 10 repulsion thicknesses).  No counterpart as an entry point in the
 * reference: the remeshing step computes it internally (src/Cloth.cu:200-265). */
cs_status cs_remesh_sizing(cs_ctx* ctx, int cloth_id, const cs_remeshing* r, float* face_sizing, float* planes);

#ifdef __cplusplus
}
#endif

#endif /* CLOTHSIM_H */

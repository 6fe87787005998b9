/*
 * rkhip.h -- C ABI of librkhip.so, the MI355X (gfx950) batched Levenberg-Marquardt library.
 *
 * Plain C types only (pointers + sizes); host arrays are float32 / int32, row-major, and are
 * copied during the call (borrowed only for its duration).  Every call returns an int status
 * (RK_OK = 0); rk_last_error() gives the calling thread's last message.  Calls on one context
 * are serialised by an internal mutex; contexts are independent.  Each context owns one HIP
 * stream (default priority) on its device.
 *
 * The Python package robokit_amd binds these with ctypes; the entry points replace the
 * reference's Warp launches on the LM path (citations are files of the reference repository).
 */
#ifndef RKHIP_H
#define RKHIP_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define RK_OK 0
#define RK_EINVAL 1     /* bad argument / shape mismatch  (reference raises ValueError) */
#define RK_EHIP 2       /* HIP runtime error */
#define RK_ECAPACITY 3  /* batch larger than capacity (optimizer.py:321-322 ValueError) */
#define RK_ENOMEM 4     /* device memory budget exceeded */
#define RK_EBUSY 5      /* no free serving slot / request not finished */

typedef struct rk_context rk_context;
typedef struct rk_batch rk_batch;
typedef struct rk_group rk_group;
typedef struct rk_server rk_server;

/* thread-local message of the last failing call */
const char* rk_last_error(void);
int32_t rk_abi_version(void);

/* one context per device ordinal; owns a stream and the uploaded robots */
int rk_context_create(int32_t device, rk_context** out);
int rk_context_destroy(rk_context* ctx);
int rk_context_sync(rk_context* ctx);
/* bytes of device memory currently allocated by the context (batches, groups, servers, robots) */
int rk_context_mem(rk_context* ctx, int64_t* bytes);
/* timing marks on the context stream (no counterpart in the reference; for benchmarks):
 * rk_context_mark records HIP event `slot` (0..7) on the stream; rk_context_elapsed waits for
 * mark b and gives the milliseconds from mark a to mark b */
int rk_context_mark(rk_context* ctx, int32_t slot);
int rk_context_elapsed(rk_context* ctx, int32_t a, int32_t b, float* ms);

/*
 * Robot topology (robo/robot_spec.py:45-138 fields).  Copied to a resident device model.
 * joint_type FIXED 0 / PRISMATIC 1 / REVOLUTE 2; parent_T [nj][7] xyz-wxyz;
 * limits [na][2] float32-narrowed (robo/robot_spec_tensors.py:17-31).
 * Limits: nj <= 64, nl <= 64, na <= 32.
 */
typedef struct {
    int32_t nj, nl, na;
    const int32_t* joint_type;
    const float* joint_axis;
    const float* parent_T;
    const int32_t* parent_joint;
    const int32_t* act_idx;
    const int32_t* mimic_idx;
    const float* mimic_mul;
    const float* mimic_off;
    const int32_t* link_parent_joint;
    const float* limits;
} rk_robot_desc;

/* replaces Robot.load -> RobotSpec.get_tensors (robo/robot_spec.py:1219-1250): resident device copy */
int rk_robot_upload(rk_context* ctx, const rk_robot_desc* desc, int32_t* robot_id);

/* replaces Robot.forward_kinematics (robo/robot.py:131-271): q [B][na] -> T_link [B][nl][7] */
int rk_fk(rk_context* ctx, int32_t robot, int32_t B, const float* q, float* T_link);

/* replaces RobotState.get_link_jacobian (robo/robot_state.py:192-224 -> robot_kernels.py:667-737):
 * frame 0 spatial / 1 body; J [B][6][na] */
int rk_link_jacobian(rk_context* ctx, int32_t robot, int32_t B, const float* q, int32_t link, int32_t frame,
                     float* J);

/* the same with the robot's base at T_world_base [B][7] (xyz-wxyz; NULL = identity):
 * Robot.forward_kinematics(q, T_world_base) (robo/robot.py:131-271, accumulating FK
 * robot_kernels.py:327-377) -> T_link [B][nl][7]; and the floating-base link Jacobian
 * (robot_kernels.py:667-737, has_floating_base) -> J [B][6][6 + na], the six base columns first:
 * Ad(T_world_base) (spatial) or Ad(T_link^-1 T_world_base) (body), then the joint columns */
int rk_fk_base(rk_context* ctx, int32_t robot, int32_t B, const float* q, const float* T_world_base, float* T_link);
int rk_link_jacobian_base(rk_context* ctx, int32_t robot, int32_t B, const float* q, const float* T_world_base,
                          int32_t link, int32_t frame, float* J);

/* measurement utility (no reference counterpart): two launches of a streaming copy of `bytes`
 * (read once, written once), the second timed; calibrates the FETCH_SIZE / WRITE_SIZE counters */
int rk_copy_probe(rk_context* ctx, int64_t bytes, float* kernel_ms);

/* device SE(3) ops for tests (lie/se3.py:38-290 host API): op 0 compose(a,b) 1 inverse(a) 2 log(a)
 * 3 jlog(a) [36] 4 adjoint(a) [36] 5 apply(a, b[3]) ; a,b [n][7]; out [n][w] */
int rk_se3_op(rk_context* ctx, int32_t op, int32_t n, const float* a, const float* b, float* out);

/*
 * Scene geometry beyond primitives, resident in the context until it is destroyed and referenced
 * by id from the collision terms (rk_terms_desc.geoms, rk_traj_desc.geoms) and rk_scene_query.
 * Scene order: obstacle spheres, boxes, then the geometries in the order given; the first strict
 * minimum decides the distance and the normal (the reference chains one query per attached
 * geometry through a per-point minimum, geom/scene.py:132-256).
 *
 * Volume (VolumeGeom / SdfVolume, geom/geoms.py:604-660, geom/sdf_volume.py:80-155): signed
 * distances at the grid points origin + (i, j, k) * voxel_size, values [nz][ny][nx] (x fastest),
 * nx, ny, nz >= 2.  Inside [aabb_min, aabb_max] the distance is the trilinear sample and the
 * normal the normalised gradient of the interpolant ((0, 0, 1) below 1e-8); outside, the distance
 * to the box plus `padding` and the normal the direction from the box (_sample_volume_sdf_func,
 * geom/sdf_kernels.py:346-380).  The box must lie inside the grid.
 * Mesh (MeshGeom, geom/geoms.py:502-600): triangles faces [nf][3] over vertices [nv][3].  The
 * distance is the exact distance to the nearest triangle, negative inside a closed, consistently
 * oriented (counter-clockwise seen from outside) mesh: the sign is that of the angle-weighted
 * pseudonormal of the nearest feature (face, edge or vertex); the normal is the nearest
 * triangle's face normal (closest_on_meshes, geom/sdf_kernels.py:151-222).  The mesh's bounding
 * box skips it wherever something nearer is already known.
 * Limits: <= 64 geometries per context, <= 2^20 triangles, <= 2^26 grid values.
 */
typedef struct {
    const float* values;
    int32_t nx, ny, nz;
    float origin[3];
    float voxel_size;
    float aabb_min[3], aabb_max[3];
    float padding;
} rk_volume_desc;
typedef struct {
    const float* vertices;  /* [n_vertices][3] */
    int32_t n_vertices;
    const int32_t* faces;   /* [n_faces][3] vertex indices */
    int32_t n_faces;
} rk_mesh_desc;
int rk_geom_upload_volume(rk_context* ctx, const rk_volume_desc* desc, int32_t* geom_id);
int rk_geom_upload_mesh(rk_context* ctx, const rk_mesh_desc* desc, int32_t* geom_id);
/* WarpScene.query_sdf (geom/scene.py:132-256) of the scene of primitives and geometries at
 * points [n][3]: signed distance [n] and the normal of the deciding element [n][3] (nullable);
 * 1e6 and (1, 0, 0) where nothing is nearer.  Also bakes a mesh into a volume: query the grid
 * points (SdfVolume.from_mesh, geom/sdf_volume.py:80-155). */
int rk_scene_query(rk_context* ctx, int32_t n_obs_spheres, const float* obs_spheres, int32_t n_obs_boxes,
                   const float* obs_boxes, int32_t n_geoms, const int32_t* geoms, int32_t n, const float* points,
                   float* dist, float* normal);

/*
 * Residual-term structure of a batch (uniform over its problems), in this row order:
 *   FrameTask log mode   (terms/dense/frame_task.py:505-621)      6 rows per pose target
 *   PositionTask         (terms/dense/position_task.py:188-272)   3 rows per position target
 *   RotationTask         (terms/dense/rotation_task.py:323-401)   3 rows per rotation target
 *   PositionLimit "abs"  (terms/dense/position_limit.py:508-573)  na rows
 *   SmoothnessTask       (terms/dense/smoothness_task.py:350-422) na rows
 *   SceneCollisionTask   (terms/dense/scene_collision_task.py:276-908) n_coll rows
 * Limits: n_pose <= 4, n_pos <= 12, n_pose + n_rot <= 4.  Batches with rotation targets run the
 * LDS-staged fused kernels (threads_per_problem 4..64), or, for a serial arm with one position and
 * one rotation target on links of one parent joint (the IK presets' term pair), the one-lane-per-
 * problem register kernel (threads_per_problem 1).
 */
typedef struct {
    int32_t robot;
    int32_t n_pose;
    const int32_t* pose_links;
    const float* pose_weights; /* [n_pose][6] */
    int32_t n_pos;
    const int32_t* pos_links;
    const float* pos_weights;  /* [n_pos][3] */
    int32_t limit;
    const float* limit_weights; /* [na] */
    int32_t smooth;
    const float* smooth_weights; /* [na] */
    int32_t n_rot;               /* RotationTask (terms/dense/rotation_task.py:323-401), rows after the position rows */
    const int32_t* rot_links;
    const float* rot_weights;    /* [n_rot][3] */
    /* SceneCollisionTask (terms/dense/scene_collision_task.py:276-908), rows after the smoothness
     * rows: one per robot sphere, r = w pen(sd) with sd = the primitive scene's signed distance at
     * the sphere centre minus its radius (geom/sdf_kernels.py:256-295) and the penalty of
     * terms/sphere_collision_kernels.py:19-42: 0 smooth (the reference default), 1 plain, 2 surface
     * distance.  Runs the one-lane-per-problem register kernel of 6- and 7-dof serial arms. */
    int32_t n_coll;              /* robot spheres, <= 64 (0: no collision term) */
    const int32_t* coll_links;   /* [n_coll] */
    const float* coll_centers;   /* [n_coll][3] link frame */
    const float* coll_radii;     /* [n_coll] */
    const float* coll_weights;   /* [n_coll] (NULL: ones) */
    float coll_margin;
    int32_t coll_penalty;
    int32_t n_obs_spheres;       /* <= 64 */
    const float* obs_spheres;    /* [n][4] centre, radius */
    int32_t n_obs_boxes;         /* <= 16 */
    const float* obs_boxes;      /* [n][6] centre, half extents (axis-aligned) */
    int32_t n_geoms;             /* <= 4 volumes / meshes of the context (rk_geom_upload_*), after the boxes */
    const int32_t* geoms;        /* [n_geoms] geometry ids */
} rk_terms_desc;

/* LMOptimizerConfig fields (opt/lm_optimizer.py:502-523) */
typedef struct {
    float lm_lambda, lambda_factor, lambda_min, lambda_max, rho_min, gain_ratio_epsilon;
} rk_lm_params;

/*
 * A batch: device-resident structure-of-arrays buffers for up to `capacity` problems of one
 * term structure (LMOptimizer._build, opt/lm_optimizer.py:544-635: buffers sized once, smaller
 * batches are prefixes, larger ones fail with RK_ECAPACITY).  trace_iters > 0 keeps a per-
 * iteration cost / accept trace of that many iterations.
 */
int rk_batch_create(rk_context* ctx, const rk_terms_desc* terms, int32_t capacity, int32_t trace_iters,
                    rk_batch** out);
int rk_batch_destroy(rk_batch* b);
/* host arrays: q [B][na], pose_targets [B][n_pose][7], pos_targets [B][n_pos][3], q_prev [B][na]
 * (NULL = leave unchanged); sets the live size to B and resets damping to `lm_lambda` at solve */
int rk_batch_set(rk_batch* b, int32_t B, const float* q, const float* pose_targets, const float* pos_targets,
                 const float* q_prev);

#define RK_MODE_FUSED 0 /* one launch: whole solve per problem in LDS/registers */
#define RK_MODE_PLAIN 1 /* reference decomposition: per-term kernels, J in HBM, host loop */
#define RK_FROM_SEED 0x100 /* flag: start from the seeds of the last rk_batch_set (kept on the device) */
#define RK_CONTINUE 0x200  /* flag: continue the last solve (current q and damping, iteration count and
                            * traces carry on) -- Optimizer.step after reset (opt/optimizer.py:280-287) */
#define RK_NOWAIT 0x400    /* flag: return once the solve is enqueued on the context's stream (kernel_ms is
                            * not written; rk_batch_get, rk_batch_select's rows or rk_context_sync wait) */
/* threads_per_problem codes: 1 = one lane per problem, register-resident (serial arms);
 * RK_TPP_WAVE = one wavefront per problem, register-resident rows (trees, <= 64 joints, <= 32 dofs);
 * RK_TPP_HALF = two problems per wavefront, 32 lanes each, register-resident rows (trees,
 *               <= 32 joints, <= 32 dofs);
 * RK_TPP_SUPER = as RK_TPP_HALF, with the damped normal equations solved in block-arrow form:
 *               the chains of the elimination tree (a humanoid's limbs, <= 7 dofs each, <= 8 of
 *               them) factored side by side on their own lanes, then the <= 3-dof root they share
 *               (the waist) from its Schur complement;
 * 4, 8, 16, 32, 64 = that many lanes per problem, LDS-staged (any robot the batch accepts) */
#define RK_TPP_WAVE (-64)
#define RK_TPP_HALF (-32)
#define RK_TPP_SUPER (-33)
#define RK_TPP_QUAD (-4)   /* four lanes per problem, register-resident: 6- and 7-dof serial arms with one
                            * pose target (forward kinematics by a scan over the quad) */
#define RK_TPP_TRAJ (-128) /* a trajectory: two wavefronts per problem (rk_group_kernel_times, autotune) */
#define RK_USE_GRAPH 0x1   /* rk_group_solve flag: capture the launch sequence in a hipGraph and replay it */
#define RK_ASYNC 0x2       /* rk_group_solve flag: return once the solve is enqueued on the context's stream
                            * (kernel_ms is not written; rk_context_sync or a later synchronous call waits) */

/*
 * Optimizer.solve with a fixed iteration count (opt/optimizer.py:216-240, LM step
 * lm_optimizer.py:834-1005).  threads_per_problem (fused mode): 0 = library default, else one
 * of 4, 8, 16 (robots with nj <= 16, na <= 8, n_pose <= 2, n_pos <= 2) or 16, 32, 64.
 * kernel_ms (nullable) = HIP-event time of the solve's kernels on the context stream.
 * iters = 0 only evaluates the cost at the current q (or the seeds): rk_batch_get returns it.
 */
int rk_batch_solve(rk_batch* b, int32_t iters, const rk_lm_params* p, int32_t mode, int32_t threads_per_problem,
                   float* kernel_ms);
/* rotation targets [B][n_rot][4] (w, x, y, z) of the live batch (after rk_batch_set) */
int rk_batch_set_rotations(rk_batch* b, int32_t B, const float* rot_targets);
/* Population select (PopulationSolver._select, opt/population_solver.py:436-449, with the order
 * of _select_best_k_kernel, :24-44): src holds n_inst problems x width_in candidates (row
 * i * width_in + j = candidate j of problem i); for each problem keep the width_out candidates of
 * lowest cost (the cost at the current iterate; NaN last; equal costs in candidate order) and
 * write them best first into dst rows i * width_out + r: q (also as dst's seeds), targets,
 * q_prev and the cost (rk_batch_get reads it).  dst takes n_inst * width_out problems, must share src's robot and term structure, and
 * its next solve starts afresh (lambda = lm_lambda).  1 <= width_out <= width_in <= 64.
 * src_rows [n_inst * width_out] (nullable, host): the source row of each dst row. */
int rk_batch_select(rk_batch* src, int32_t width_in, rk_batch* dst, int32_t width_out, int32_t* src_rows);
/* the same select ranking by the costs of `scores` (a batch of src's size whose cost holds each
 * candidate's score: a zero-iteration solve of the score terms at src's iterates) instead of src's
 * own costs; dst's costs receive the kept rows' scores -- IKSolverConfig.score_terms
 * (helpers/ik/config.py:57-59), ranked at every select (PopulationSolver._score / _select,
 * opt/population_solver.py:395-449) */
int rk_batch_select_scored(rk_batch* src, const rk_batch* scores, int32_t width_in, rk_batch* dst, int32_t width_out,
                           int32_t* src_rows);
/* copies the live problems' q of src into dst (same robot and dofs) and sets dst's size to src's:
 * a score batch evaluates another batch's iterates */
int rk_batch_copy_q(rk_batch* src, rk_batch* dst);
/* warm start for a sequence of frames (helpers/retarget: the previous frame's solution is the
 * next frame's seed and its SmoothnessTask reference, terms/dense/smoothness_task.py:350-365):
 * on the device, q_prev <- q and seeds <- q; the next rk_batch_set then only brings new targets */
int rk_batch_advance(rk_batch* b);
/* outputs (nullable): q [B][na], cost [B], lambda [B], cost_trace [B][trace_iters+1], accept_trace [B][trace_iters] */
int rk_batch_get(rk_batch* b, float* q, float* cost, float* lambda, float* cost_trace, int32_t* accept_trace);
/* iterates of the last traced solve: q_trace [B][trace_iters][na] = q before step i,
 * lam_trace [B][trace_iters] = damping step i used (the state an iteration_callback sees,
 * opt/optimizer.py:230-236) */
int rk_batch_get_iterates(rk_batch* b, float* q_trace, float* lam_trace);
/* _ik_metrics_kernel (helpers/ik/ik.py:59-82) over the batch's pose targets at its current q */
int rk_batch_metrics(rk_batch* b, float pos_thr, float ang_thr, float* pos_err, float* ang_err, int32_t* success);
/* No counterpart in the reference: for callers that cannot allocate device memory (tight serving
 * loops, code around HIP graph capture).  The same results as rk_batch_metrics, bit for bit, but
 * nothing is allocated: the per-problem results are staged in a static device array of the library
 * (one per GPU, held until the process ends) for at most RK_NOALLOC_MAX_PROBLEMS problems (larger
 * batches fail with RK_EINVAL).  Calls that overlap on one GPU (several threads or contexts on the
 * same device at once) share that array and race; calls on one context are serialised. */
#define RK_NOALLOC_MAX_PROBLEMS 65536
int rk_batch_metrics_noalloc(rk_batch* b, float pos_thr, float ang_thr, float* pos_err, float* ang_err,
                             int32_t* success);

/*
 * Uneven batch: several batches (different robots / sizes / terms; at most 32) solved by ONE
 * persistent kernel that pulls problem tiles from a device work queue; each batch runs with its
 * own threads-per-problem.  Trajectory batches (rk_traj_batch_create) of any horizons join the
 * group too: those of one arm class (dofs; bound rows or scene geometries) share one launch of
 * consecutive workgroup ranges, on a stream forked from the context stream, so they overlap the
 * IK kernel (parallel branches of the captured graph).  No reference counterpart (the reference
 * batches one robot per optimizer).
 */
int rk_group_create(rk_context* ctx, int32_t nbatches, rk_batch* const* batches, rk_group** out);
int rk_group_destroy(rk_group* g);
/* tpp [nbatches] (0 = default), workgroups per CU of the persistent grid (0 = occupancy limit),
 * iterations per launch (0 = all in one launch) */
int rk_group_set_tuning(rk_group* g, const int32_t* tpp, int32_t wg_per_cu, int32_t iters_per_launch);
/* time every legal threads-per-problem for each batch on a short run and keep the fastest;
 * chosen_tpp [nbatches] (nullable) receives the choice */
int rk_group_autotune(rk_group* g, int32_t iters, const rk_lm_params* p, int32_t* chosen_tpp);
/* solve all batches; flags: RK_USE_GRAPH (capture once per shape, then replay), RK_FROM_SEED, RK_ASYNC.
 * A graph or async solve of register-resident variants (threads-per-problem 1 up to 7 dofs,
 * RK_TPP_WAVE, RK_TPP_HALF) runs as one launch; otherwise one kernel per variant on forked
 * streams.  Results are the same bits either way. */
int rk_group_solve(rk_group* g, int32_t iters, const rk_lm_params* p, int32_t flags, float* kernel_ms);
/* per persistent kernel of the last eager (not graph-replayed) solve -- one per kernel variant,
 * launched concurrently on forked streams: its threads-per-problem code, HIP-event time on its
 * stream and the batches it ran (bit i = batch i of the group, first 32); *n = kernel count
 * (0 after a graph replay) */
int rk_group_kernel_times(rk_group* g, int32_t cap, int32_t* tpp, float* ms, int32_t* batch_mask, int32_t* n);

/*
 * Collision-aware trajectory LM (512 trajectories x 64 frames in the benchmark): T frames of a
 * serial arm per problem, residuals in this row order:
 *   goal       TrajectoryTask(FrameTask log) on the last frame (terms/trajectory_task.py:65-295)
 *   velocity   TrajectorySmoothnessTask order 1, weight / dt (trajectory_smoothness_task.py:506-557)
 *   accel.     TrajectorySmoothnessTask order 2, weight / dt^2 (first window: no history, weight 0)
 *   collision  TrajectoryCollisionTask, primitive scene, penalty "plain", no sweep
 *              (terms/sparse/trajectory_collision_task.py:176-293): w * max(0, margin - sd)
 *   bounds     TrajectoryLimitTask, forward stencils of orders 0..2 (optional, see below)
 * Banded normal equations (half-bandwidth 2 * dofs) solved per problem
 * (opt/lm_optimizer.py:81-136).  The batch then takes rk_batch_set (q [B][T][na], pose_targets =
 * goal [B][7]), rk_batch_solve (RK_MODE_FUSED), rk_batch_get, rk_batch_advance.
 * Limits: serial arms with 6 or 7 dofs, T <= 64, <= 64 robot spheres, <= 64 obstacle spheres,
 * <= 16 axis-aligned obstacle boxes.
 */
typedef struct {
    int32_t robot;
    int32_t frames;
    int32_t goal_link;
    const float* goal_weights;    /* [6] (NULL: ones) */
    float vel_weight, acc_weight, dt;
    int32_t n_spheres;
    const int32_t* sphere_links;  /* [n] */
    const float* sphere_centers;  /* [n][3] link frame */
    const float* sphere_radii;    /* [n] */
    float coll_weight, coll_margin;
    int32_t n_obs_spheres;
    const float* obs_spheres;     /* [n][4] centre, radius */
    int32_t n_obs_boxes;
    const float* obs_boxes;       /* [n][6] centre, half extents */
    /* TrajectoryLimitTask, forward stencils (terms/sparse/trajectory_limit_task.py:29-537): bit o of
     * limit_orders adds one bound row per window of o + 1 frames and dof, order 0 position,
     * 1 velocity (q[t+1] - q[t]) / dt, 2 acceleration (q[t] - 2 q[t+1] + q[t+2]) / dt^2;
     * r = w max(lower - v, v - upper, 0).  limit_weights / limit_lower / limit_upper [3][na] (rows of
     * absent orders unused; NULL when limit_orders = 0).  Rows after the collision rows. */
    int32_t limit_orders;
    const float* limit_weights;
    const float* limit_lower;
    const float* limit_upper;
    /* linear solver of each LM step: 0 the band Cholesky; n > 0 n iterations of conjugate
     * gradients on (J^T J + lam I) dq = -g from dq = 0 (LMOptimizerConfig(linear_solver="cg",
     * cg_max_iter=n), opt/lm_optimizer.py:781-832), always the full budget: the reference's CG
     * early exit is a host check of the whole batch, which it also skips under graph capture
     * and with early_stopping_interval = 0.  Predicted reduction -g.dq - dq^T J^T J dq / 2. */
    int32_t cg_max_iter;
    int32_t n_geoms;              /* <= 4 volumes / meshes of the context (rk_geom_upload_*), after the boxes */
    const int32_t* geoms;         /* [n_geoms] geometry ids */
} rk_traj_desc;

int rk_traj_batch_create(rk_context* ctx, const rk_traj_desc* desc, int32_t capacity, int32_t trace_iters,
                         rk_batch** out);
/* goal position / angle error at the last frame and the smallest sphere-to-scene signed distance
 * over all frames and spheres, per problem [B] */
int rk_traj_metrics(rk_batch* b, float* pos_err, float* ang_err, float* min_sd);

/*
 * Request server: continuous batching of solves under a device-memory budget.  No reference
 * counterpart (the reference solves one caller's batch at a time); it serves many callers of the
 * LM path of opt/optimizer.py:216-240 concurrently.  A worker thread packs every queued request of
 * one term structure and iteration count into one launch whenever one of `streams` launch streams
 * is free and the `mem_budget`-byte device arena has room; launches on different streams overlap.
 * Once trajectory terms are registered (and streams > 1), stream 0 carries trajectory launches and
 * the others IK launches, and a trajectory launch is enqueued a few iterations at a time, so that an
 * IK launch never waits behind a whole trajectory solve; `streams` is best kept within the
 * process's hardware queues (4 by default).
 * Robots and term structures stay resident; a session keeps its last solution resident as the
 * next request's seeds and smoothness reference.  success uses pos_thr / ang_thr
 * (helpers/ik/ik.py:59-82) and needs pose targets.
 */
typedef struct {
    int32_t terms;               /* id from rk_server_add_terms */
    int32_t n;                   /* problems, 1..max_launch_problems */
    int32_t iters;               /* fixed LM iterations */
    rk_lm_params params;
    const float* q;              /* [n][na] seeds; NULL: the session's last solution */
    const float* pose_targets;   /* [n][n_pose][7] */
    const float* pos_targets;    /* [n][n_pos][3] */
    const float* q_prev;         /* [n][na] smoothness reference; NULL: the session's last solution, else q */
    int64_t session;             /* 0: none */
} rk_request;

typedef struct {
    int64_t submitted, completed, failed, queued, launches, problems, converged;
    double p50_ms, p99_ms, mean_ms, max_ms;   /* submit -> results ready, completed requests */
    double mean_problems_per_launch;
    int64_t budget_bytes, peak_bytes, bytes_in_use;
    double span_s;                            /* first submit -> last completion */
} rk_serve_stats;

int rk_server_create(rk_context* ctx, int64_t mem_budget, int32_t streams, int32_t max_launch_problems,
                     float pos_thr, float ang_thr, rk_server** out);
/* register a term structure (resident; its kernel variant is chosen once) */
int rk_server_add_terms(rk_server* s, const rk_terms_desc* terms, int32_t* term_id);
/* register a trajectory problem set (rk_traj_desc); its requests carry q [n][frames][na] and
 * pose_targets = goals [n][7]; success = goal within the thresholds and no sphere penetrating */
int rk_server_add_traj(rk_server* s, const rk_traj_desc* desc, int32_t* term_id);
/* queue a request (inputs copied); returns at once with a ticket */
int rk_server_submit(rk_server* s, const rk_request* req, int64_t* ticket);
/* results of a ticket: wait_ms 0 = poll (RK_EBUSY while running), < 0 = block; outputs nullable:
 * q [n][na], cost [n], success [n], pos_err / ang_err [n][n_pose], latency_ms; collects the ticket */
int rk_server_poll(rk_server* s, int64_t ticket, int32_t wait_ms, float* q, float* cost, int32_t* success,
                   float* pos_err, float* ang_err, float* latency_ms);
/* release a session's resident state (RK_EBUSY while it has requests queued or running) */
int rk_server_end_session(rk_server* s, int64_t session);
int rk_server_stats(rk_server* s, rk_serve_stats* out);
/* drains the queue, waits for the device, frees everything */
int rk_server_destroy(rk_server* s);

#ifdef __cplusplus
}
#endif
#endif /* RKHIP_H */

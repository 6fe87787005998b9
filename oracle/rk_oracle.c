/*
 * rk_oracle.c -- TEST INFRASTRUCTURE ONLY (the CPU oracle that results are checked against).
 *
 * A plain-C restatement of RoboKit's Warp-CPU Levenberg-Marquardt path, used by
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg to check and
 * time the HIP product path.  Nothing in robokit_amd/ may link or call this.
 *
 * Built twice from this one source (see oracle/Makefile):
 *   liboracle_f32.so  REAL=float   -- follows the reference's FP32 arithmetic
 *   liboracle_f64.so  REAL=double  -- the FP64 twin, measures intrinsic sensitivity
 *
 * Every function cites the reference file:line whose algorithm it restates
 * (paths relative to the reference repository root).  Operation order follows
 * those formulas; Warp's own tile_matmul / tile_cholesky rounding order is not
 * available (Warp is absent), so sums here run in ascending index order.
 *
 * Layout conventions (same as the reference): pose = [x,y,z,qw,qx,qy,qz],
 * twist = [v; w], joint types FIXED 0 / PRISMATIC 1 / REVOLUTE 2.
 *
 * Pinned (tests/test_oracle.py) before it is trusted:
 *   ork_se3_*            golden constants of tests/lie/test_se3.py:39-80, homogeneous-matrix
 *                        algebra, exp(log T) = T, Ad(T^-1) Ad(T) = I, jlog = finite-difference
 *                        derivative of log(T exp(xi))
 *   ork_fk               tests/robo/fk_ground_truth.json (panda, ur10; with/without base),
 *                        doctest literals of robo/robot.py:139-154
 *   ork_link_jacobian    tests/robo/jacobian_ground_truth.json (joint + base columns),
 *                        doctest literals of robo/robot.py:362-376, central differences of FK
 *   ork_residual_jacobian  central differences of the stacked residual for every term kind,
 *                        closed forms of the limit / smoothness rows, H = J^T J, g = J^T r
 *   ork_damped_solve     numpy.linalg.solve(H + lam I, -g); pred = quadratic-model reduction
 *   ork_accept           truth table of tests/opt/test_optimizer_acceptance.py:41-66 + clamps
 *   ork_lm_solve/_step   one step = solve(J^T J + lam I, -J^T r) (tests/opt/test_optimizer.py:
 *                        366-429); step == first solve iteration; IK convergence from near seeds
 *   ork_ik_metrics       errors of targets displaced by known translations / rotations
 * Parity unpinned: the exact FP32 rounding sequence of Warp's CPU tile_matmul / tile_cholesky
 * (Warp is not importable here); only its tolerance-level effect is bounded.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#ifndef REAL
#define REAL float
#endif
typedef REAL real;

#if REAL_IS_DOUBLE
#define RSIN sin
#define RCOS cos
#define RSQRT sqrt
#define RATAN2 atan2
#define RASIN asin
#define RFABS fabs
#else
#define RSIN sinf
#define RCOS cosf
#define RSQRT sqrtf
#define RATAN2 atan2f
#define RASIN asinf
#define RFABS fabsf
#endif

#define EXPORT __attribute__((visibility("default")))

/* ------------------------------------------------------------------------- */
/* quaternion helpers: xform/warp/rotation_conversions.py                     */
/* ------------------------------------------------------------------------- */

/* quaternion_multiply_func + standardize_quaternion_func, rotation_conversions.py:259-279, 249-256 */
static void quat_mul(const real *a, const real *b, real *out) {
    real w1 = a[0], x1 = a[1], y1 = a[2], z1 = a[3];
    real w2 = b[0], x2 = b[1], y2 = b[2], z2 = b[3];
    real w = w1 * w2 - x1 * x2 - y1 * y2 - z1 * z2;
    real x = w1 * x2 + x1 * w2 + y1 * z2 - z1 * y2;
    real y = w1 * y2 - x1 * z2 + y1 * w2 + z1 * x2;
    real z = w1 * z2 + x1 * y2 - y1 * x2 + z1 * w2;
    if (w < (real)0) { w = -w; x = -x; y = -y; z = -z; }
    out[0] = w; out[1] = x; out[2] = y; out[3] = z;
}

static void cross3(const real *a, const real *b, real *o) {
    real x = a[1] * b[2] - a[2] * b[1];
    real y = a[2] * b[0] - a[0] * b[2];
    real z = a[0] * b[1] - a[1] * b[0];
    o[0] = x; o[1] = y; o[2] = z;
}

/* quaternion_apply_func, rotation_conversions.py:193-200: v + 2(w (u x v) + u x (u x v)) */
static void quat_apply(const real *q, const real *v, real *out) {
    real w = q[0];
    real u[3] = {q[1], q[2], q[3]};
    real uv[3], uuv[3];
    cross3(u, v, uv);
    cross3(u, uv, uuv);
    for (int i = 0; i < 3; ++i) out[i] = v[i] + (real)2 * (w * uv[i] + uuv[i]);
}

/* quaternion_to_matrix_func, rotation_conversions.py:174-187 (row-major 3x3) */
static void quat_to_mat(const real *q, real *R) {
    real w = q[0], x = q[1], y = q[2], z = q[3];
    real two_s = (real)2 / (w * w + x * x + y * y + z * z);
    R[0] = (real)1 - two_s * (y * y + z * z);
    R[1] = two_s * (x * y - z * w);
    R[2] = two_s * (x * z + y * w);
    R[3] = two_s * (x * y + z * w);
    R[4] = (real)1 - two_s * (x * x + z * z);
    R[5] = two_s * (y * z - x * w);
    R[6] = two_s * (x * z - y * w);
    R[7] = two_s * (y * z + x * w);
    R[8] = (real)1 - two_s * (x * x + y * y);
}

/* quaternion_to_axis_angle_func, rotation_conversions.py:226-246 */
static void quat_to_axis_angle(const real *q, real *aa) {
    real w = q[0];
    real norm = RSQRT(q[1] * q[1] + q[2] * q[2] + q[3] * q[3]);
    real half_angle = RATAN2(norm, w);
    real angle = (real)2 * half_angle;
    real s;
    if (RFABS(angle) < (real)1e-6) {
        s = (real)0.5 - (angle * angle) / (real)48;
    } else {
        s = RSIN(half_angle) / angle;
    }
    aa[0] = q[1] / s; aa[1] = q[2] / s; aa[2] = q[3] / s;
}

/* ------------------------------------------------------------------------- */
/* SE(3): lie/se3_kernels.py, lie/so3_kernels.py                             */
/* ------------------------------------------------------------------------- */

/* se3_compose_func, se3_kernels.py:87-112 */
static void se3_compose(const real *a, const real *b, real *out) {
    real q[4], t[3];
    quat_mul(a + 3, b + 3, q);
    quat_apply(a + 3, b, t);
    out[0] = a[0] + t[0]; out[1] = a[1] + t[1]; out[2] = a[2] + t[2];
    out[3] = q[0]; out[4] = q[1]; out[5] = q[2]; out[6] = q[3];
}

/* se3_inverse_func, se3_kernels.py:69-78 */
static void se3_inverse(const real *a, real *out) {
    real qi[4] = {a[3], -a[4], -a[5], -a[6]};
    real t[3];
    quat_apply(qi, a, t);
    out[0] = -t[0]; out[1] = -t[1]; out[2] = -t[2];
    out[3] = qi[0]; out[4] = qi[1]; out[5] = qi[2]; out[6] = qi[3];
}

static void hat(const real *v, real *K) {
    K[0] = 0; K[1] = -v[2]; K[2] = v[1];
    K[3] = v[2]; K[4] = 0; K[5] = -v[0];
    K[6] = -v[1]; K[7] = v[0]; K[8] = 0;
}

static void mat3_mul(const real *A, const real *B, real *C) {
    real T[9];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            real s = A[i * 3 + 0] * B[0 * 3 + j];
            s += A[i * 3 + 1] * B[1 * 3 + j];
            s += A[i * 3 + 2] * B[2 * 3 + j];
            T[i * 3 + j] = s;
        }
    memcpy(C, T, sizeof(T));
}

static void mat3_vec(const real *A, const real *v, real *o) {
    real t[3];
    for (int i = 0; i < 3; ++i) t[i] = A[i * 3 + 0] * v[0] + A[i * 3 + 1] * v[1] + A[i * 3 + 2] * v[2];
    o[0] = t[0]; o[1] = t[1]; o[2] = t[2];
}

/* se3_exp_map_func, se3_kernels.py:125-167 (eps-regularised) */
static void se3_exp(const real *xi, real eps, real *out) {
    const real *w = xi + 3;
    real angle_sq = w[0] * w[0] + w[1] * w[1] + w[2] * w[2];
    real angle_sq_reg = angle_sq + eps * eps;
    real angle_reg = RSQRT(angle_sq_reg);
    real half = angle_reg * (real)0.5;
    real sh = RSIN(half), ch = RCOS(half);
    real k = sh / angle_reg;
    real q[4] = {ch, k * w[0], k * w[1], k * w[2]};
    real qn = RSQRT(q[0] * q[0] + q[1] * q[1] + q[2] * q[2] + q[3] * q[3]);
    for (int i = 0; i < 4; ++i) q[i] = q[i] / qn;
    real sa = RSIN(angle_reg), ca = RCOS(angle_reg);
    real A = sa / angle_reg;
    real c1 = ((real)1 - ca) / angle_sq_reg;
    real c2 = ((real)1 - A) / angle_sq_reg;
    real K[9], K2[9], V[9];
    hat(w, K);
    mat3_mul(K, K, K2);
    for (int i = 0; i < 9; ++i) V[i] = ((i % 4) == 0 ? (real)1 : (real)0) + c1 * K[i] + c2 * K2[i];
    real t[3];
    mat3_vec(V, xi, t);
    out[0] = t[0]; out[1] = t[1]; out[2] = t[2];
    out[3] = q[0]; out[4] = q[1]; out[5] = q[2]; out[6] = q[3];
}

/* se3_V_matrix_inverse_func, se3_kernels.py:230-282 */
static void se3_V_inv(const real *w, real eps, real *Vi) {
    real angle = RSQRT(w[0] * w[0] + w[1] * w[1] + w[2] * w[2]);
    real K[9], K2[9];
    hat(w, K);
    mat3_mul(K, K, K2);
    real c2;
    if (angle < eps) {
        c2 = (real)1 / (real)12;
    } else {
        real sa = RSIN(angle), ca = RCOS(angle);
        real omc = (real)1 - ca;
        if (RFABS(omc) < eps)
            c2 = (real)1 / (real)12;
        else
            c2 = ((real)1 / (angle * angle)) * ((real)1 - (angle * sa) / ((real)2 * omc));
    }
    for (int i = 0; i < 9; ++i) Vi[i] = ((i % 4) == 0 ? (real)1 : (real)0) - (real)0.5 * K[i] + c2 * K2[i];
}

/* se3_log_map_func, se3_kernels.py:285-309 */
static void se3_log(const real *T, real eps, real *out) {
    real w[3], Vi[9], v[3];
    quat_to_axis_angle(T + 3, w);
    se3_V_inv(w, eps, Vi);
    mat3_vec(Vi, T, v);
    out[0] = v[0]; out[1] = v[1]; out[2] = v[2];
    out[3] = w[0]; out[4] = w[1]; out[5] = w[2];
}

/* se3_adjoint_func, se3_kernels.py:323-346: [[R, t^R],[0,R]] row-major 6x6 */
static void se3_adjoint(const real *T, real *Ad) {
    real R[9], tx[9], tR[9];
    quat_to_mat(T + 3, R);
    hat(T, tx);
    mat3_mul(tx, R, tR);
    memset(Ad, 0, 36 * sizeof(real));
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            Ad[i * 6 + j] = R[i * 3 + j];
            Ad[i * 6 + 3 + j] = tR[i * 3 + j];
            Ad[(i + 3) * 6 + 3 + j] = R[i * 3 + j];
        }
}

/* so3_jac_left_inv_func, so3_kernels.py:13-40 */
static void so3_jl_inv(const real *th, real eps, real *out) {
    real t2 = th[0] * th[0] + th[1] * th[1] + th[2] * th[2];
    int taylor = t2 < eps;
    real t2s = taylor ? (real)1 : t2;
    real ts = RSQRT(t2s);
    real hts = ts / (real)2;
    real K[9], K2[9];
    hat(th, K);
    mat3_mul(K, K, K2);
    if (taylor) {
        for (int i = 0; i < 9; ++i)
            out[i] = ((i % 4) == 0 ? (real)1 : (real)0) - (real)0.5 * K[i] + t2s * K2[i] / (real)12;
    } else {
        real ch = RCOS(hts), sh = RSIN(hts);
        real coeff = ((real)1 - ts * ch / ((real)2 * sh)) / t2s;
        for (int i = 0; i < 9; ++i) out[i] = ((i % 4) == 0 ? (real)1 : (real)0) - (real)0.5 * K[i] + coeff * K2[i];
    }
}

/* se3_jlog_func, se3_kernels.py:358-428 -> row-major 6x6 */
static void se3_jlog(const real *T, real eps, real *J) {
    const real *p = T;
    real w[3];
    quat_to_axis_angle(T + 3, w);
    real t2 = w[0] * w[0] + w[1] * w[1] + w[2] * w[2];
    real th = RSQRT(t2);
    int taylor = t2 < eps;
    real jl[9], jlT[9];
    so3_jl_inv(w, eps, jl);
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) jlT[i * 3 + j] = jl[j * 3 + i];
    real th_safe = taylor ? (real)1 : th;
    real th_inv = (real)1 / th_safe;
    real th2_inv = th_inv * th_inv;
    real st = RSIN(th), ct = RCOS(th);
    real omc = (real)1 - ct;
    real inv_2_2ct = taylor ? (real)0.5 : (real)1 / ((real)2 * omc);
    real beta = th2_inv - st * th_inv * inv_2_2ct;
    real bdot = (real)-2 * th2_inv * th2_inv + ((real)1 + st * th_inv) * th2_inv * inv_2_2ct;
    real wTp = w[0] * p[0] + w[1] * p[1] + w[2] * p[2];
    real v3[3];
    for (int i = 0; i < 3; ++i) v3[i] = bdot * wTp * w[i] - (t2 * bdot + (real)2 * beta) * p[i];
    real C[9], sk[9], B[9];
    hat(p, sk);
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
            C[i * 3 + j] = v3[i] * w[j] + beta * (w[i] * p[j]) + ((i == j) ? (wTp * beta) : (real)0);
    for (int i = 0; i < 9; ++i) C[i] = C[i] + (real)0.5 * sk[i];
    mat3_mul(C, jlT, B);
    if (taylor)
        for (int i = 0; i < 9; ++i) B[i] = (real)0.5 * sk[i];
    memset(J, 0, 36 * sizeof(real));
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            J[i * 6 + j] = jlT[i * 3 + j];
            J[i * 6 + 3 + j] = B[i * 3 + j];
            J[(i + 3) * 6 + 3 + j] = jlT[i * 3 + j];
        }
}

static void mat6_vec(const real *M, const real *v, real *o) {
    real t[6];
    for (int i = 0; i < 6; ++i) {
        real s = M[i * 6 + 0] * v[0];
        for (int k = 1; k < 6; ++k) s += M[i * 6 + k] * v[k];
        t[i] = s;
    }
    memcpy(o, t, sizeof(t));
}

/* ------------------------------------------------------------------------- */
/* exported SE(3) ops (batched, for tests against the reference's constants)  */
/* ------------------------------------------------------------------------- */
EXPORT void ork_se3_compose(int n, const real *a, const real *b, real *out) {
    for (int i = 0; i < n; ++i) se3_compose(a + 7 * i, b + 7 * i, out + 7 * i);
}
EXPORT void ork_se3_inverse(int n, const real *a, real *out) {
    for (int i = 0; i < n; ++i) se3_inverse(a + 7 * i, out + 7 * i);
}
EXPORT void ork_se3_apply(int n, const real *T, const real *p, real *out) {
    for (int i = 0; i < n; ++i) {
        real r[3];
        quat_apply(T + 7 * i + 3, p + 3 * i, r);
        for (int k = 0; k < 3; ++k) out[3 * i + k] = r[k] + T[7 * i + k];
    }
}
EXPORT void ork_se3_exp(int n, const real *xi, real eps, real *out) {
    for (int i = 0; i < n; ++i) se3_exp(xi + 6 * i, eps, out + 7 * i);
}
EXPORT void ork_se3_log(int n, const real *T, real eps, real *out) {
    for (int i = 0; i < n; ++i) se3_log(T + 7 * i, eps, out + 6 * i);
}
EXPORT void ork_se3_adjoint(int n, const real *T, real *out) {
    for (int i = 0; i < n; ++i) se3_adjoint(T + 7 * i, out + 36 * i);
}
EXPORT void ork_se3_jlog(int n, const real *T, real eps, real *out) {
    for (int i = 0; i < n; ++i) se3_jlog(T + 7 * i, eps, out + 36 * i);
}

/* ------------------------------------------------------------------------- */
/* robot model (robo/robot_spec.py:45-138, 1043-1142)                         */
/* ------------------------------------------------------------------------- */
typedef struct {
    int nj, nl, na;
    const int32_t *joint_type;        /* [nj] */
    const float *joint_axis;          /* [nj*3] */
    const float *parent_T;            /* [nj*7] parent_joint_transforms */
    const int32_t *parent_joint;      /* [nj] */
    const int32_t *act_idx;           /* [nj] actuated_joint_indices */
    const int32_t *mimic_idx;         /* [nj] mimic_actuated_joint_indices */
    const float *mimic_mul;           /* [nj] */
    const float *mimic_off;           /* [nj] */
    const int32_t *link_parent_joint; /* [nl] */
    const float *limits;              /* [na*2] float32-narrowed lo,hi */
} ork_robot;

/* joint_twists, robot_spec.py:320-331 */
static void joint_twist(const ork_robot *rb, int j, real *tw) {
    for (int k = 0; k < 6; ++k) tw[k] = 0;
    const float *ax = rb->joint_axis + 3 * j;
    if (rb->joint_type[j] == 2) { tw[3] = ax[0]; tw[4] = ax[1]; tw[5] = ax[2]; }
    else if (rb->joint_type[j] == 1) { tw[0] = ax[0]; tw[1] = ax[1]; tw[2] = ax[2]; }
}

/* joints_to_actuated_mapping, robot_spec.py:1120-1132 */
static real j2a(const ork_robot *rb, int j, int a) {
    if (rb->act_idx[j] >= 0) return rb->act_idx[j] == a ? (real)1 : (real)0;
    if (rb->mimic_idx[j] >= 0) return rb->mimic_idx[j] == a ? (real)rb->mimic_mul[j] : (real)0;
    return 0;
}

/* link_ancestor_joints_mask, robot_spec.py:1134-1143 */
static int ancestor(const ork_robot *rb, int link, int joint) {
    int j = rb->link_parent_joint[link];
    while (j != -1) {
        if (j == joint) return 1;
        j = rb->parent_joint[j];
    }
    return 0;
}

/* _compute_joint_local_transform_func, robo/robot_kernels.py:264-297 */
static void joint_local(const ork_robot *rb, const real *q, int j, real *X) {
    real pos;
    if (rb->mimic_idx[j] == -1) {
        pos = rb->act_idx[j] == -1 ? (real)0 : q[rb->act_idx[j]];
    } else {
        pos = q[rb->mimic_idx[j]] * (real)rb->mimic_mul[j] + (real)rb->mimic_off[j];
    }
    real Xj[7] = {0, 0, 0, 1, 0, 0, 0};
    const float *ax = rb->joint_axis + 3 * j;
    if (rb->joint_type[j] == 2) {
        real half = pos * (real)0.5;
        real c = RCOS(half), s = RSIN(half);
        Xj[3] = c; Xj[4] = (real)ax[0] * s; Xj[5] = (real)ax[1] * s; Xj[6] = (real)ax[2] * s;
    } else if (rb->joint_type[j] == 1) {
        Xj[0] = (real)ax[0] * pos; Xj[1] = (real)ax[1] * pos; Xj[2] = (real)ax[2] * pos;
    }
    real P[7];
    for (int k = 0; k < 7; ++k) P[k] = (real)rb->parent_T[7 * j + k];
    se3_compose(P, Xj, X);
}

/*
 * Two-pass FK: compute_forward_kinematics_local_kernel (robot_kernels.py:300-324)
 * then compute_forward_kinematics_accum_kernel (robot_kernels.py:327-377): each
 * link composes its parent chain leaf-to-root, then T_world_base on the left.
 * With S != NULL also the motion-subspace column S_world[:, j] = Ad(T_wj) twist_j
 * (robot_kernels.py:365-368), layout S[6*nj] row-major (row r, joint j) = S[r*nj+j].
 */
static void fk(const ork_robot *rb, const real *q, const real *base, real *Xl, real *Tj, real *Tl, real *S) {
    for (int j = 0; j < rb->nj; ++j) joint_local(rb, q, j, Xl + 7 * j);
    for (int l = 0; l < rb->nl; ++l) {
        int j = rb->link_parent_joint[l];
        real Xw[7];
        for (int k = 0; k < 7; ++k) Xw[k] = base[k];
        if (j != -1) {
            for (int k = 0; k < 7; ++k) Xw[k] = Xl[7 * j + k];
            int p = rb->parent_joint[j];
            while (p != -1) {
                real t[7];
                se3_compose(Xl + 7 * p, Xw, t);
                memcpy(Xw, t, sizeof(t));
                p = rb->parent_joint[p];
            }
            real t[7];
            se3_compose(base, Xw, t);
            memcpy(Xw, t, sizeof(t));
        }
        memcpy(Tl + 7 * l, Xw, sizeof(Xw));
        if (j != -1) {
            memcpy(Tj + 7 * j, Xw, sizeof(Xw));
            if (S) {
                real Ad[36], tw[6], col[6];
                se3_adjoint(Xw, Ad);
                joint_twist(rb, j, tw);
                mat6_vec(Ad, tw, col);
                for (int r = 0; r < 6; ++r) S[r * rb->nj + j] = col[r];
            }
        }
    }
}

static const real IDENT7[7] = {0, 0, 0, 1, 0, 0, 0};

/* Robot.forward_kinematics (robo/robot.py:131-271), CPU two-pass path. base may be NULL. */
EXPORT void ork_fk(const ork_robot *rb, int B, const real *q, const real *base, real *T_link, real *T_joint) {
    real *Xl = (real *)malloc(sizeof(real) * 7 * (rb->nj + 1));
    real *Tj = (real *)malloc(sizeof(real) * 7 * (rb->nj + 1));
    for (int b = 0; b < B; ++b) {
        const real *bs = base ? base + 7 * b : IDENT7;
        fk(rb, q + (size_t)b * rb->na, bs, Xl, Tj, T_link + (size_t)b * 7 * rb->nl, NULL);
        if (T_joint) memcpy(T_joint + (size_t)b * 7 * rb->nj, Tj, sizeof(real) * 7 * rb->nj);
    }
    free(Xl);
    free(Tj);
}

/*
 * RobotState.get_link_jacobian -> compute_link_jacobian_kernel, robot_kernels.py:667-737.
 * frame 0 = spatial, 1 = body.  Output J[6 x (6*fb + na)] row-major per batch element;
 * with base != NULL the floating-base columns come first.
 */
EXPORT void ork_link_jacobian(const ork_robot *rb, int B, const real *q, const real *base, int link, int frame,
                              real *J) {
    int nj = rb->nj, na = rb->na;
    int fb = base ? 6 : 0;
    int cols = fb + na;
    real *Xl = (real *)malloc(sizeof(real) * 7 * (nj + 1));
    real *Tj = (real *)malloc(sizeof(real) * 7 * (nj + 1));
    real *Tl = (real *)malloc(sizeof(real) * 7 * (rb->nl + 1));
    real *S = (real *)calloc(6 * (nj + 1), sizeof(real));
    for (int b = 0; b < B; ++b) {
        const real *bs = base ? base + 7 * b : IDENT7;
        memset(S, 0, sizeof(real) * 6 * nj);
        fk(rb, q + (size_t)b * na, bs, Xl, Tj, Tl, S);
        real *Jb = J + (size_t)b * 6 * cols;
        const real *Tf = Tl + 7 * link;
        for (int c = 0; c < cols; ++c) {
            real col[6];
            if (c < fb) {
                real u[6] = {0, 0, 0, 0, 0, 0};
                u[c] = 1;
                real Ad[36];
                if (frame == 0) {
                    se3_adjoint(bs, Ad);
                } else {
                    real Tfw[7], Tfb[7];
                    se3_inverse(Tf, Tfw);
                    se3_compose(Tfw, bs, Tfb);
                    se3_adjoint(Tfb, Ad);
                }
                mat6_vec(Ad, u, col);
            } else {
                int a = c - fb;
                real sc[6] = {0, 0, 0, 0, 0, 0};
                for (int j = 0; j < nj; ++j) {
                    if (!ancestor(rb, link, j)) continue;
                    real w = j2a(rb, j, a);
                    if (w != (real)0)
                        for (int r = 0; r < 6; ++r) sc[r] += S[r * nj + j] * w;
                }
                if (frame == 1) {
                    real Tfw[7], Ad[36];
                    se3_inverse(Tf, Tfw);
                    se3_adjoint(Tfw, Ad);
                    mat6_vec(Ad, sc, col);
                } else {
                    memcpy(col, sc, sizeof(col));
                }
            }
            for (int r = 0; r < 6; ++r) Jb[r * cols + c] = col[r];
        }
    }
    free(Xl);
    free(Tj);
    free(Tl);
    free(S);
}

/* primitive scene SDF at p (geom/sdf_kernels.py:223-295): the minimum over the obstacle spheres,
 * then the axis-aligned boxes, the first strict minimum kept; n = its gradient direction */
static real prims_sdf(int nobs_s, const float *obs_s, int nobs_b, const float *obs_b, const real *p, real *n) {
    real best = (real)1e6;
    n[0] = 1; n[1] = 0; n[2] = 0;
    for (int i = 0; i < nobs_s; ++i) {
        const float *o = obs_s + 4 * i;
        real d[3] = {p[0] - (real)o[0], p[1] - (real)o[1], p[2] - (real)o[2]};
        real len = RSQRT(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
        real dist = len - (real)o[3];
        if (dist < best) {
            best = dist;
            if (len > (real)1e-10) { n[0] = d[0] / len; n[1] = d[1] / len; n[2] = d[2] / len; }
            else { n[0] = 1; n[1] = 0; n[2] = 0; }
        }
    }
    for (int i = 0; i < nobs_b; ++i) {
        const float *o = obs_b + 6 * i;
        real ql[3], qv[3], h[3];
        for (int k = 0; k < 3; ++k) {
            ql[k] = p[k] - (real)o[k];
            h[k] = (real)o[3 + k];
            qv[k] = RFABS(ql[k]) - h[k];
        }
        real qmax = qv[0] > qv[1] ? qv[0] : qv[1];
        if (qv[2] > qmax) qmax = qv[2];
        real qp[3] = {qv[0] > 0 ? qv[0] : 0, qv[1] > 0 ? qv[1] : 0, qv[2] > 0 ? qv[2] : 0};
        real dist = RSQRT(qp[0] * qp[0] + qp[1] * qp[1] + qp[2] * qp[2]) + (qmax < 0 ? qmax : 0);
        if (dist < best) {
            best = dist;
            real nn[3] = {0, 0, 1};
            if (qmax < 0) {
                if (qv[0] >= qv[1] && qv[0] >= qv[2]) { nn[0] = ql[0] < 0 ? -1 : 1; nn[2] = 0; }
                else if (qv[1] >= qv[2]) { nn[1] = ql[1] < 0 ? -1 : 1; nn[2] = 0; }
                else { nn[2] = ql[2] < 0 ? -1 : 1; }
            } else {
                real df[3];
                for (int k = 0; k < 3; ++k) {
                    real c = ql[k] < -h[k] ? -h[k] : (ql[k] > h[k] ? h[k] : ql[k]);
                    df[k] = ql[k] - c;
                }
                real dl = RSQRT(df[0] * df[0] + df[1] * df[1] + df[2] * df[2]);
                if (dl > (real)1e-10) { nn[0] = df[0] / dl; nn[1] = df[1] / dl; nn[2] = df[2] / dl; }
            }
            n[0] = nn[0]; n[1] = nn[1]; n[2] = nn[2];
        }
    }
    return best;
}

/* Scene volumes and meshes (geom/geoms.py:502-660, geom/sdf_kernels.py:151-222, 346-380). */
typedef struct {
    int kind;                  /* 0 volume, 1 mesh */
    const float *values;       /* volume: [nz][ny][nx] at origin + (i, j, k) * voxel */
    int nx, ny, nz;
    float origin[3], voxel, lo[3], hi[3], pad;
    const float *verts;        /* mesh: [nv][3] */
    int nv;
    const int32_t *faces;      /* [nf][3] */
    int nf;
} ork_geom;

/* volume: inside [lo, hi] the trilinear interpolant written as the weighted sum
 * sum_{ijk} w_i(fx) w_j(fy) w_k(fz) v_ijk (w_0(f) = 1 - f, w_1(f) = f) of the 8 grid values of the
 * cell holding p (cell index clamped to the grid), grad = its derivative in index units;
 * outside, |p - clamp(p)| + pad and the direction from the box (_sample_volume_sdf_func,
 * sdf_kernels.py:346-380) */
static real vol_sdf(const ork_geom *g, const real *p, real *grad) {
    int inside = 1;
    for (int k = 0; k < 3; ++k) inside = inside && p[k] >= (real)g->lo[k] && p[k] <= (real)g->hi[k];
    if (inside) {
        const int n[3] = {g->nx, g->ny, g->nz};
        int i[3];
        real f[3];
        for (int k = 0; k < 3; ++k) {
            real u = (p[k] - (real)g->origin[k]) / (real)g->voxel;
            int b = (int)floor((double)u);
            if (b < 0) b = 0;
            if (b > n[k] - 2) b = n[k] - 2;
            i[k] = b;
            f[k] = u - (real)b;
        }
        real val = 0, gr[3] = {0, 0, 0};
        for (int dz = 0; dz < 2; ++dz)
            for (int dy = 0; dy < 2; ++dy)
                for (int dx = 0; dx < 2; ++dx) {
                    const real v = (real)g->values[((size_t)(i[2] + dz) * g->ny + (i[1] + dy)) * g->nx + (i[0] + dx)];
                    const real wx = dx ? f[0] : 1 - f[0], wy = dy ? f[1] : 1 - f[1], wz = dz ? f[2] : 1 - f[2];
                    val += wx * wy * wz * v;
                    gr[0] += (dx ? 1 : -1) * wy * wz * v;
                    gr[1] += wx * (dy ? 1 : -1) * wz * v;
                    gr[2] += wx * wy * (dz ? 1 : -1) * v;
                }
        if (grad) for (int k = 0; k < 3; ++k) grad[k] = gr[k];
        return val;
    }
    real o[3], l2 = 0;
    for (int k = 0; k < 3; ++k) {
        real c = p[k] < (real)g->lo[k] ? (real)g->lo[k] : (p[k] > (real)g->hi[k] ? (real)g->hi[k] : p[k]);
        o[k] = p[k] - c;
        l2 += o[k] * o[k];
    }
    const real ad = RSQRT(l2);
    if (grad) for (int k = 0; k < 3; ++k) grad[k] = o[k] / ad;
    return ad + (real)g->pad;
}

static void mesh_face_normal(const ork_geom *g, int f, real *n) {
    const int32_t *F = g->faces + 3 * f;
    real e1[3], e2[3];
    for (int k = 0; k < 3; ++k) {
        e1[k] = (real)g->verts[3 * F[1] + k] - (real)g->verts[3 * F[0] + k];
        e2[k] = (real)g->verts[3 * F[2] + k] - (real)g->verts[3 * F[0] + k];
    }
    n[0] = e1[1] * e2[2] - e1[2] * e2[1];
    n[1] = e1[2] * e2[0] - e1[0] * e2[2];
    n[2] = e1[0] * e2[1] - e1[1] * e2[0];
    const real l = RSQRT(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
    if (l > 0) for (int k = 0; k < 3; ++k) n[k] /= l;
}

/* angle-weighted pseudonormal (Baerentzen and Aanaes, IEEE TVCG 11(3), 2005) of vertex v (u < 0)
 * or of the edge (v, u): the incident angles times the face normals, or the sum of the normals of
 * the faces holding the edge, normalised; computed by a scan of every face */
static void mesh_pseudonormal(const ork_geom *g, int v, int u, real *out) {
    out[0] = out[1] = out[2] = 0;
    for (int f = 0; f < g->nf; ++f) {
        const int32_t *F = g->faces + 3 * f;
        int cv = -1, hu = 0;
        for (int c = 0; c < 3; ++c) {
            if (F[c] == v) cv = c;
            if (F[c] == u) hu = 1;
        }
        if (cv < 0 || (u >= 0 && !hu)) continue;
        real n[3];
        mesh_face_normal(g, f, n);
        real wgt = 1;
        if (u < 0) {
            const int v1 = F[(cv + 1) % 3], v2 = F[(cv + 2) % 3];
            real a[3], b[3];
            for (int k = 0; k < 3; ++k) {
                a[k] = (real)g->verts[3 * v1 + k] - (real)g->verts[3 * v + k];
                b[k] = (real)g->verts[3 * v2 + k] - (real)g->verts[3 * v + k];
            }
            const real la = RSQRT(a[0] * a[0] + a[1] * a[1] + a[2] * a[2]);
            const real lb = RSQRT(b[0] * b[0] + b[1] * b[1] + b[2] * b[2]);
            real cs = (la > 0 && lb > 0) ? (a[0] * b[0] + a[1] * b[1] + a[2] * b[2]) / (la * lb) : 1;
            if (cs > 1) cs = 1;
            if (cs < -1) cs = -1;
            wgt = (real)acos((double)cs);
        }
        for (int k = 0; k < 3; ++k) out[k] += wgt * n[k];
    }
    const real l = RSQRT(out[0] * out[0] + out[1] * out[1] + out[2] * out[2]);
    if (l > 0) for (int k = 0; k < 3; ++k) out[k] /= l;
}

/* mesh: the nearest point of each triangle by the Voronoi-region tests of Ericson, Real-Time
 * Collision Detection (2005) section 5.1.5, the first strict minimum over the triangles; signed
 * by the pseudonormal of the nearest feature (vertex, edge or face); face = its triangle
 * (closest_on_meshes, sdf_kernels.py:151-222) */
static real mesh_sdf(const ork_geom *g, const real *p, int *face, real *dir) {
    real best = 1e300, bq[3] = {0, 0, 0};
    int bf = -1, bv = -1, bu = -1, bface_feat = 0;
    for (int f = 0; f < g->nf; ++f) {
        const int32_t *F = g->faces + 3 * f;
        real a[3], b[3], c[3], ab[3], ac[3], ap[3], bp[3], cp[3], q[3];
        for (int k = 0; k < 3; ++k) {
            a[k] = (real)g->verts[3 * F[0] + k];
            b[k] = (real)g->verts[3 * F[1] + k];
            c[k] = (real)g->verts[3 * F[2] + k];
            ab[k] = b[k] - a[k]; ac[k] = c[k] - a[k];
            ap[k] = p[k] - a[k]; bp[k] = p[k] - b[k]; cp[k] = p[k] - c[k];
        }
#define ODOT(x, y) ((x)[0] * (y)[0] + (x)[1] * (y)[1] + (x)[2] * (y)[2])
        int fv = -1, fu = -1, isface = 0;  /* nearest feature: vertex fv, edge (fv, fu), face */
        const real d1 = ODOT(ab, ap), d2 = ODOT(ac, ap);
        const real d3 = ODOT(ab, bp), d4 = ODOT(ac, bp);
        const real d5 = ODOT(ab, cp), d6 = ODOT(ac, cp);
        const real vc = d1 * d4 - d3 * d2, vb = d5 * d2 - d1 * d6, va = d3 * d6 - d5 * d4;
        if (d1 <= 0 && d2 <= 0) {
            for (int k = 0; k < 3; ++k) q[k] = a[k];
            fv = F[0];
        } else if (d3 >= 0 && d4 <= d3) {
            for (int k = 0; k < 3; ++k) q[k] = b[k];
            fv = F[1];
        } else if (vc <= 0 && d1 >= 0 && d3 <= 0) {
            const real v = d1 / (d1 - d3);
            for (int k = 0; k < 3; ++k) q[k] = a[k] + v * ab[k];
            fv = F[0]; fu = F[1];
        } else if (d6 >= 0 && d5 <= d6) {
            for (int k = 0; k < 3; ++k) q[k] = c[k];
            fv = F[2];
        } else if (vb <= 0 && d2 >= 0 && d6 <= 0) {
            const real w = d2 / (d2 - d6);
            for (int k = 0; k < 3; ++k) q[k] = a[k] + w * ac[k];
            fv = F[0]; fu = F[2];
        } else if (va <= 0 && (d4 - d3) >= 0 && (d5 - d6) >= 0) {
            const real w = (d4 - d3) / ((d4 - d3) + (d5 - d6));
            for (int k = 0; k < 3; ++k) q[k] = b[k] + w * (c[k] - b[k]);
            fv = F[1]; fu = F[2];
        } else {
            const real den = 1 / (va + vb + vc), v = vb * den, w = vc * den;
            for (int k = 0; k < 3; ++k) q[k] = a[k] + ab[k] * v + ac[k] * w;
            isface = 1;
        }
        real r[3] = {p[0] - q[0], p[1] - q[1], p[2] - q[2]};
        const real d = ODOT(r, r);
        if (d < best) {
            best = d;
            bf = f; bv = fv; bu = fu; bface_feat = isface;
            for (int k = 0; k < 3; ++k) bq[k] = q[k];
        }
    }
    *face = bf;
    if (bf < 0) return (real)1e30;
    real pn[3];
    if (bface_feat) mesh_face_normal(g, bf, pn);
    else mesh_pseudonormal(g, bv, bu, pn);
    real r[3] = {p[0] - bq[0], p[1] - bq[1], p[2] - bq[2]};
    const real d = RSQRT(best);
    const real sg = ODOT(r, pn) < 0 ? -1 : 1;
    if (dir) {  /* sign * (p - nearest) / |p - nearest|; the face normal on the surface */
        if (d > (real)1e-12) for (int k = 0; k < 3; ++k) dir[k] = sg * r[k] / d;
        else mesh_face_normal(g, bf, dir);
    }
    return sg * d;
#undef ODOT
}

/* the scene: primitives (prims_sdf), then each geometry in order, the first strict minimum; n =
 * its normal: a volume's normalised gradient ((0, 0, 1) below 1e-8, sdf_kernels.py:540-545); a
 * mesh's outward direction from the nearest point = the face normal of the reference
 * (sdf_kernels.py:206-213) wherever that point is inside a face, the distance gradient on edges
 * and vertices, where the reference's face choice follows its BVH search order */
static real scene_sdf_all(int nobs_s, const float *obs_s, int nobs_b, const float *obs_b, int ngeom,
                          const ork_geom *geoms, const real *p, real *n) {
    real best = prims_sdf(nobs_s, obs_s, nobs_b, obs_b, p, n);
    for (int i = 0; i < ngeom; ++i) {
        const ork_geom *g = geoms + i;
        if (g->kind == 0) {
            real gr[3];
            const real d = vol_sdf(g, p, gr);
            if (d < best) {
                best = d;
                const real l = RSQRT(gr[0] * gr[0] + gr[1] * gr[1] + gr[2] * gr[2]);
                if (l > (real)1e-8) { n[0] = gr[0] / l; n[1] = gr[1] / l; n[2] = gr[2] / l; }
                else { n[0] = 0; n[1] = 0; n[2] = 1; }
            }
        } else {
            int f;
            real dir[3];
            const real d = mesh_sdf(g, p, &f, dir);
            if (d < best) {
                best = d;
                n[0] = dir[0]; n[1] = dir[1]; n[2] = dir[2];
            }
        }
    }
    return best;
}

/* WarpScene.query_sdf at points [n][3]: distance [n], normal [n][3] */
EXPORT void ork_scene_query(int nobs_s, const float *obs_s, int nobs_b, const float *obs_b, int ngeom,
                            const ork_geom *geoms, int n, const real *pts, real *dist, real *normal) {
    for (int i = 0; i < n; ++i) dist[i] = scene_sdf_all(nobs_s, obs_s, nobs_b, obs_b, ngeom, geoms, pts + 3 * i,
                                                        normal + 3 * i);
}

/* ------------------------------------------------------------------------- */
/* residual terms                                                            */
/* ------------------------------------------------------------------------- */
enum { ORK_FRAME = 0, ORK_POSITION = 1, ORK_LIMIT = 2, ORK_SMOOTH = 3, ORK_ROTATION = 4, ORK_COLLISION = 5 };

/* SceneCollisionTask data: the robot's spheres and a primitive scene */
typedef struct {
    int nsph;
    const int32_t *sph_link;  /* [nsph] */
    const float *sph_c;       /* [nsph][3] link frame */
    const float *sph_r;       /* [nsph] */
    const float *sph_w;       /* [nsph] row weights */
    float margin;
    int penalty;              /* 0 smooth, 1 plain, 2 surface distance */
    int nobs_s;
    const float *obs_s;       /* [n][4] */
    int nobs_b;
    const float *obs_b;       /* [n][6] */
    int ngeom;                /* then volumes / meshes */
    const ork_geom *geoms;
} ork_scene;

typedef struct {
    int kind;
    int nframes;            /* FRAME / POSITION: number of targets */
    const int32_t *links;   /* [nframes] link indices */
    const float *targets;   /* [B][nframes][7] world targets (one per problem) */
    const float *weights;   /* FRAME: 6*nframes; POSITION: 3*nframes; LIMIT/SMOOTH: na */
    const float *q_prev;    /* SMOOTH: [B][na] */
    const ork_scene *scene; /* COLLISION */
} ork_term;

static int term_rows(const ork_robot *rb, const ork_term *t) {
    switch (t->kind) {
    case ORK_FRAME: return 6 * t->nframes;
    case ORK_POSITION: return 3 * t->nframes;
    case ORK_ROTATION: return 3 * t->nframes;
    case ORK_COLLISION: return t->scene->nsph;
    default: return rb->na;
    }
}

/* _compute_collision_penalty_func (terms/sphere_collision_kernels.py:19-42): the residual of a
 * signed distance and its derivative; smooth (0), plain (1), surface distance (2) */
static void coll_penalty(real sd, real margin, int mode, real *pen, real *dpen) {
    if (mode == 2) {
        *pen = sd < 0 ? -sd : sd;
        *dpen = sd < 0 ? (real)-1 : (real)1;
    } else if (mode == 1) {
        *pen = sd < margin ? margin - sd : (real)0;
        *dpen = sd < margin ? (real)-1 : (real)0;
    } else if (sd < 0) {
        *pen = (real)0.5 * margin - sd;
        *dpen = -1;
    } else if (sd <= margin) {
        const real diff = sd - margin, inv = (real)1 / (margin + (real)1e-6);
        *pen = (real)0.5 * inv * diff * diff;
        *dpen = inv * diff;
    } else {
        *pen = 0;
        *dpen = 0;
    }
}

/* spatial column of link `link` for actuated dof a: sum over ancestor joints of S_j * map[j,a]
 * (frame_task.py:431-437, position_task.py:245-251) */
static void spatial_column(const ork_robot *rb, const real *S, int link, int a, real *sc) {
    for (int r = 0; r < 6; ++r) sc[r] = 0;
    for (int j = 0; j < rb->nj; ++j) {
        if (!ancestor(rb, link, j)) continue;
        real w = j2a(rb, j, a);
        if (w != (real)0)
            for (int r = 0; r < 6; ++r) sc[r] += S[r * rb->nj + j] * w;
    }
}

/*
 * Weighted residual of one term for one problem b (row-major r[R]); with J != NULL also its
 * weighted Jacobian J[R x na] (caller zeroes it).  Follows:
 *   FRAME    compute_(multi_)frame_task_weighted_residual_kernel frame_task.py:505-554,
 *            compute_(multi_)frame_task_weighted_jacobian_kernel frame_task.py:391-456, 557-621
 *   POSITION position_task.py:188-212, 215-272
 *   LIMIT    compute_position_limit_weighted_residual/jacobian_kernel position_limit.py:508-573 ("abs")
 *   SMOOTH   compute_smoothness_task_weighted_residual/jacobian_kernel smoothness_task.py:350-365, 404-422
 *   ROTATION compute_rotation_task_weighted_residual_kernel rotation_task.py:323-349 (axis-angle of
 *            q_link * q_target^-1) and the geometric Jacobian rotation_task.py:352-401 (w * omega column)
 *   COLLISION sphere_collision_residual_apply_kernel sphere_collision_kernels.py:68-97 (w pen(sd),
 *            sd = scene distance of the sphere centre - radius) and _sphere_collision_jacobian_kernel
 *            scene_collision_task.py:33-81 (w pen'(sd) n . (v + w x c), zero beyond the margin
 *            except for the surface distance), the point column of scene_distance_task.py:44-70
 */
static void term_eval(const ork_robot *rb, const ork_term *t, int b, const real *q, const real *Tl, const real *S,
                      real *r, real *J) {
    int na = rb->na;
    if (t->kind == ORK_FRAME) {
        for (int f = 0; f < t->nframes; ++f) {
            const real *Tf = Tl + 7 * t->links[f];
            real tgt[7];
            for (int k = 0; k < 7; ++k) tgt[k] = (real)t->targets[((size_t)b * t->nframes + f) * 7 + k];
            const float *w = t->weights + 6 * f;
            real Tfw[7], Tft[7], e[6];
            se3_inverse(Tf, Tfw);
            se3_compose(Tfw, tgt, Tft);
            se3_log(Tft, (real)1e-4, e);
            for (int i = 0; i < 6; ++i) r[6 * f + i] = (real)w[i] * e[i];
            if (J) {
                real Ttw[7], Ttf[7], jl[36], Ad[36];
                se3_inverse(tgt, Ttw);
                se3_compose(Ttw, Tf, Ttf);
                se3_jlog(Ttf, (real)1e-4, jl);
                se3_adjoint(Tfw, Ad);
                for (int a = 0; a < na; ++a) {
                    real sc[6], bc[6], jc[6];
                    spatial_column(rb, S, t->links[f], a, sc);
                    mat6_vec(Ad, sc, bc);
                    mat6_vec(jl, bc, jc);
                    for (int i = 0; i < 6; ++i) J[(6 * f + i) * na + a] = -(real)w[i] * jc[i];
                }
            }
        }
    } else if (t->kind == ORK_POSITION) {
        for (int f = 0; f < t->nframes; ++f) {
            const real *Tf = Tl + 7 * t->links[f];
            const float *tg = t->targets + ((size_t)b * t->nframes + f) * 7;
            const float *w = t->weights + 3 * f;
            for (int i = 0; i < 3; ++i) r[3 * f + i] = (real)w[i] * ((real)tg[i] - Tf[i]);
            if (J) {
                for (int a = 0; a < na; ++a) {
                    real sc[6], wxp[3];
                    spatial_column(rb, S, t->links[f], a, sc);
                    cross3(sc + 3, Tf, wxp);
                    for (int i = 0; i < 3; ++i) J[(3 * f + i) * na + a] = -(real)w[i] * (sc[i] + wxp[i]);
                }
            }
        }
    } else if (t->kind == ORK_LIMIT) {
        for (int d = 0; d < na; ++d) {
            real lo = (real)rb->limits[2 * d], hi = (real)rb->limits[2 * d + 1];
            real x = q[d];
            real eu = x - hi > (real)0 ? x - hi : (real)0;
            real el = lo - x > (real)0 ? lo - x : (real)0;
            r[d] = (real)t->weights[d] * (eu + el);
            if (J) {
                real sign = 0;
                if (x > hi) sign = 1;
                else if (x < lo) sign = -1;
                J[d * na + d] = (real)t->weights[d] * sign;
            }
        }
    } else if (t->kind == ORK_ROTATION) {
        for (int f = 0; f < t->nframes; ++f) {
            const real *Tf = Tl + 7 * t->links[f];
            const float *tg = t->targets + ((size_t)b * t->nframes + f) * 7;
            const float *w = t->weights + 3 * f;
            real qinv[4] = {(real)tg[3], -(real)tg[4], -(real)tg[5], -(real)tg[6]}, qe[4], aa[3];
            quat_mul(Tf + 3, qinv, qe);
            quat_to_axis_angle(qe, aa);
            for (int i = 0; i < 3; ++i) r[3 * f + i] = (real)w[i] * aa[i];
            if (J) {
                for (int a = 0; a < na; ++a) {
                    real sc[6];
                    spatial_column(rb, S, t->links[f], a, sc);
                    for (int i = 0; i < 3; ++i) J[(3 * f + i) * na + a] = (real)w[i] * sc[3 + i];
                }
            }
        }
    } else if (t->kind == ORK_SMOOTH) {
        const float *qp = t->q_prev + (size_t)b * na;
        for (int d = 0; d < na; ++d) {
            r[d] = (real)t->weights[d] * (q[d] - (real)qp[d]);
            if (J) J[d * na + d] = (real)t->weights[d];
        }
    } else if (t->kind == ORK_COLLISION) {
        const ork_scene *sc = t->scene;
        for (int k = 0; k < sc->nsph; ++k) {
            const int link = sc->sph_link[k];
            const real *Tk = Tl + 7 * link;
            real lc[3] = {(real)sc->sph_c[3 * k], (real)sc->sph_c[3 * k + 1], (real)sc->sph_c[3 * k + 2]}, c[3], n[3];
            quat_apply(Tk + 3, lc, c);
            for (int i = 0; i < 3; ++i) c[i] += Tk[i];
            const real sd = scene_sdf_all(sc->nobs_s, sc->obs_s, sc->nobs_b, sc->obs_b, sc->ngeom, sc->geoms, c, n) -
                            (real)sc->sph_r[k];
            real pen, dpen;
            coll_penalty(sd, (real)sc->margin, sc->penalty, &pen, &dpen);
            const real w = (real)sc->sph_w[k];
            r[k] = w * pen;
            if (J && (sc->penalty == 2 || sd <= (real)sc->margin)) {
                real cn[3];
                cross3(c, n, cn);
                for (int a = 0; a < na; ++a) {
                    real s6[6];
                    spatial_column(rb, S, link, a, s6);
                    J[k * na + a] = w * dpen * ((n[0] * s6[0] + n[1] * s6[1] + n[2] * s6[2]) +
                                                (cn[0] * s6[3] + cn[1] * s6[4] + cn[2] * s6[5]));
                }
            }
        }
    }
}

/* ------------------------------------------------------------------------- */
/* LM: opt/lm_optimizer.py generic (CPU) path                                */
/* ------------------------------------------------------------------------- */
typedef struct {
    float lm_lambda, lambda_factor, lambda_min, lambda_max, rho_min, gain_ratio_epsilon;
} ork_lm_params;

/*
 * Dense damped Cholesky solve, the algorithm of _cholesky_kernel dense branch
 * (lm_optimizer.py:56-79): A = H + lam I; L L^T = A; delta = -A^{-1} g;
 * pred = 0.5 * sum(delta * (lam*delta - g)).  Returns 0 on a non-positive pivot
 * (delta/pred then hold NaN, which the accept test rejects, lm_optimizer.py:352-362).
 */
static int damped_cholesky(int D, const real *H, const real *g, real lam, real *L, real *delta, real *pred) {
    for (int i = 0; i < D * D; ++i) L[i] = H[i];
    for (int i = 0; i < D; ++i) L[i * D + i] = L[i * D + i] + lam;
    int ok = 1;
    for (int j = 0; j < D; ++j) {
        real s = L[j * D + j];
        for (int k = 0; k < j; ++k) s -= L[j * D + k] * L[j * D + k];
        if (!(s > (real)0)) ok = 0;
        real d = RSQRT(s);
        L[j * D + j] = d;
        for (int i = j + 1; i < D; ++i) {
            real v = L[i * D + j];
            for (int k = 0; k < j; ++k) v -= L[i * D + k] * L[j * D + k];
            L[i * D + j] = v / d;
        }
    }
    real *y = (real *)malloc(sizeof(real) * D);
    for (int i = 0; i < D; ++i) {
        real v = -g[i];
        for (int k = 0; k < i; ++k) v -= L[i * D + k] * y[k];
        y[i] = v / L[i * D + i];
    }
    for (int i = D - 1; i >= 0; --i) {
        real v = y[i];
        for (int k = i + 1; k < D; ++k) v -= L[k * D + i] * delta[k];
        delta[i] = v / L[i * D + i];
    }
    free(y);
    real p = 0;
    for (int i = 0; i < D; ++i) p += delta[i] * (lam * delta[i] - g[i]);
    *pred = (real)0.5 * p;
    return ok;
}

/* lm_accept_reject_cost_kernel, lm_optimizer.py:337-362 */
EXPORT void ork_accept(int n, real *cost, const real *cost_prop, const real *pred, const ork_lm_params *p, real *lam,
                       int32_t *acc) {
    for (int b = 0; b < n; ++b) {
        int accepted = pred[b] > (real)0;
        if (accepted) {
            real rho = (cost[b] - cost_prop[b]) / (pred[b] + (real)p->gain_ratio_epsilon);
            accepted = rho >= (real)p->rho_min;
        }
        if (accepted) {
            acc[b] = 1;
            real l = lam[b] / (real)p->lambda_factor;
            lam[b] = l > (real)p->lambda_min ? l : (real)p->lambda_min;
            cost[b] = cost_prop[b];
        } else {
            acc[b] = 0;
            real l = lam[b] * (real)p->lambda_factor;
            if (l < (real)p->lambda_min) l = (real)p->lambda_min;
            if (l > (real)p->lambda_max) l = (real)p->lambda_max;
            lam[b] = l;
        }
    }
}

/* damped solve only (exported for the one-step delta test, tests/opt/test_optimizer.py:366-429) */
EXPORT int ork_damped_solve(int D, const real *H, const real *g, real lam, real *delta, real *pred) {
    real *L = (real *)malloc(sizeof(real) * D * D);
    int ok = damped_cholesky(D, H, g, lam, L, delta, pred);
    free(L);
    return ok;
}

/*
 * Residual + Jacobian + normal equations of all terms at q for problem b:
 * the per-term ResidualTask.accumulate_normal_equations (terms/task.py:176-232) with
 * _accum_kernel (lm_optimizer.py:228-257): H += J^T J, g += J^T r, cost += 0.5 r^T r.
 * Also returns the stacked residual r_all (rows in term order) when non-NULL.
 */
typedef struct {
    real *Xl, *Tj, *Tl, *S, *r, *J, *H, *g, *L, *delta, *q_prop;
    int R;
} ork_ws;

static void ws_alloc(ork_ws *w, const ork_robot *rb, int nterms, const ork_term *terms) {
    int R = 0;
    for (int t = 0; t < nterms; ++t) R += term_rows(rb, &terms[t]);
    w->R = R;
    int nj = rb->nj, na = rb->na;
    w->Xl = (real *)calloc(7 * (nj + 1), sizeof(real));
    w->Tj = (real *)calloc(7 * (nj + 1), sizeof(real));
    w->Tl = (real *)calloc(7 * (rb->nl + 1), sizeof(real));
    w->S = (real *)calloc(6 * (nj + 1), sizeof(real));
    w->r = (real *)calloc(R + 1, sizeof(real));
    w->J = (real *)calloc((size_t)(R + 1) * na, sizeof(real));
    w->H = (real *)calloc(na * na, sizeof(real));
    w->g = (real *)calloc(na, sizeof(real));
    w->L = (real *)calloc(na * na, sizeof(real));
    w->delta = (real *)calloc(na, sizeof(real));
    w->q_prop = (real *)calloc(na, sizeof(real));
}

static void ws_free(ork_ws *w) {
    free(w->Xl); free(w->Tj); free(w->Tl); free(w->S); free(w->r); free(w->J);
    free(w->H); free(w->g); free(w->L); free(w->delta); free(w->q_prop);
}

static real normal_equations(const ork_robot *rb, int nterms, const ork_term *terms, int b, const real *q,
                             const real *base, ork_ws *w) {
    int na = rb->na;
    memset(w->S, 0, sizeof(real) * 6 * rb->nj);
    fk(rb, q, base, w->Xl, w->Tj, w->Tl, w->S);
    memset(w->H, 0, sizeof(real) * na * na);
    memset(w->g, 0, sizeof(real) * na);
    real cost = 0;
    int row = 0;
    for (int t = 0; t < nterms; ++t) {
        int Rt = term_rows(rb, &terms[t]);
        real *r = w->r + row;
        real *J = w->J + (size_t)row * na;
        memset(J, 0, sizeof(real) * Rt * na);
        term_eval(rb, &terms[t], b, q, w->Tl, w->S, r, J);
        for (int i = 0; i < na; ++i)
            for (int j = 0; j < na; ++j) {
                real s = 0;
                for (int k = 0; k < Rt; ++k) s += J[k * na + i] * J[k * na + j];
                w->H[i * na + j] = w->H[i * na + j] + s;
            }
        for (int i = 0; i < na; ++i) {
            real s = 0;
            for (int k = 0; k < Rt; ++k) s += J[k * na + i] * r[k];
            w->g[i] = w->g[i] + s;
        }
        real rr = 0;
        for (int k = 0; k < Rt; ++k) rr += r[k] * r[k];
        cost = cost + (real)0.5 * rr;
        row += Rt;
    }
    return cost;
}

/* cost-only pass: per-term residual then _strided_cost_kernel add (optimizer.py:80-105) */
static real cost_only(const ork_robot *rb, int nterms, const ork_term *terms, int b, const real *q, const real *base,
                      ork_ws *w) {
    fk(rb, q, base, w->Xl, w->Tj, w->Tl, NULL);
    real cost = 0;
    for (int t = 0; t < nterms; ++t) {
        int Rt = term_rows(rb, &terms[t]);
        term_eval(rb, &terms[t], b, q, w->Tl, NULL, w->r, NULL);
        real s = 0;
        for (int k = 0; k < Rt; ++k) s += w->r[k] * w->r[k];
        cost = cost + (real)0.5 * s;
    }
    return cost;
}

/* stacked residual and Jacobian at q (exported for finite-difference and one-step tests) */
EXPORT int ork_residual_jacobian(const ork_robot *rb, int nterms, const ork_term *terms, int b, const real *q,
                                 const real *base, real *r_out, real *J_out, real *H_out, real *g_out,
                                 real *cost_out) {
    ork_ws w;
    ws_alloc(&w, rb, nterms, terms);
    const real *bs = base ? base : IDENT7;
    real c = normal_equations(rb, nterms, terms, b, q, bs, &w);
    if (r_out) memcpy(r_out, w.r, sizeof(real) * w.R);
    if (J_out) memcpy(J_out, w.J, sizeof(real) * w.R * rb->na);
    if (H_out) memcpy(H_out, w.H, sizeof(real) * rb->na * rb->na);
    if (g_out) memcpy(g_out, w.g, sizeof(real) * rb->na);
    if (cost_out) *cost_out = c;
    int R = w.R;
    ws_free(&w);
    return R;
}

/*
 * Optimizer.solve (opt/optimizer.py:216-240) with LMOptimizer's generic step
 * (_propose_body generic branch lm_optimizer.py:912-962, _accept_body :964-982),
 * fixed iteration count (early_stopping_interval = 0).  Problems are independent,
 * so the batch runs problem by problem.
 *
 *   q        [B][na] in: seeds, out: solution
 *   cost_trace [B][iters+1] (nullable): cost at the seed, then after every step
 *   accept_trace [B][iters] (nullable), lambda_trace [B][iters] (nullable)
 */
EXPORT void ork_lm_solve(const ork_robot *rb, int nterms, const ork_term *terms, const ork_lm_params *p, int B,
                         int iters, real *q, const real *base, real *cost_out, real *cost_trace,
                         int32_t *accept_trace, real *lambda_trace) {
    int na = rb->na;
    ork_ws w;
    ws_alloc(&w, rb, nterms, terms);
    for (int b = 0; b < B; ++b) {
        real *qb = q + (size_t)b * na;
        const real *bs = base ? base + 7 * b : IDENT7;
        real lam = (real)p->lm_lambda;
        real cost = 0;
        if (cost_trace) cost_trace[(size_t)b * (iters + 1)] = cost_only(rb, nterms, terms, b, qb, bs, &w);
        if (iters == 0) cost = cost_only(rb, nterms, terms, b, qb, bs, &w);
        for (int it = 0; it < iters; ++it) {
            cost = normal_equations(rb, nterms, terms, b, qb, bs, &w);
            real pred;
            damped_cholesky(na, w.H, w.g, lam, w.L, w.delta, &pred);
            /* integrate_joint_positions_kernel, robot_kernels.py:562-590 (mask 1, no clamp, no decay) */
            for (int d = 0; d < na; ++d) w.q_prop[d] = qb[d] + (real)1 * (w.delta[d] - (real)0 * qb[d]);
            real cp = cost_only(rb, nterms, terms, b, w.q_prop, bs, &w);
            int32_t acc;
            ork_accept(1, &cost, &cp, &pred, p, &lam, &acc);
            if (acc) memcpy(qb, w.q_prop, sizeof(real) * na);
            if (cost_trace) cost_trace[(size_t)b * (iters + 1) + it + 1] = cost;
            if (accept_trace) accept_trace[(size_t)b * iters + it] = acc;
            if (lambda_trace) lambda_trace[(size_t)b * iters + it] = lam;
        }
        if (cost_out) cost_out[b] = cost;
    }
    ws_free(&w);
}

/*
 * One LM step (the body of ork_lm_solve's loop) from given iterates: problem b starts at
 * q[b] with damping lam[b].  Outputs the cost at q, the proposal's cost, the predicted
 * reduction, the accept flag, the next iterate and damping.  Used to check each GPU step
 * from the GPU's own iterate ("shadowing"), which is insensitive to the rounding-level forks
 * a free-running comparison of two LM trajectories suffers.
 */
EXPORT void ork_lm_step(const ork_robot *rb, int nterms, const ork_term *terms, const ork_lm_params *p, int B,
                        const real *q, const real *lam, real *cost_cur, real *cost_prop, real *pred_out,
                        int32_t *acc_out, real *q_new, real *lam_new) {
    int na = rb->na;
    ork_ws w;
    ws_alloc(&w, rb, nterms, terms);
    for (int b = 0; b < B; ++b) {
        const real *qb = q + (size_t)b * na;
        real l = lam[b];
        real cost = normal_equations(rb, nterms, terms, b, qb, IDENT7, &w);
        cost_cur[b] = cost;
        real pred;
        damped_cholesky(na, w.H, w.g, l, w.L, w.delta, &pred);
        for (int d = 0; d < na; ++d) w.q_prop[d] = qb[d] + (real)1 * (w.delta[d] - (real)0 * qb[d]);
        real cp = cost_only(rb, nterms, terms, b, w.q_prop, IDENT7, &w);
        cost_prop[b] = cp;
        pred_out[b] = pred;
        int32_t acc;
        ork_accept(1, &cost, &cp, &pred, p, &l, &acc);
        acc_out[b] = acc;
        lam_new[b] = l;
        for (int d = 0; d < na; ++d) q_new[(size_t)b * na + d] = acc ? w.q_prop[d] : qb[d];
    }
    ws_free(&w);
}

/*
 * _ik_metrics_kernel, helpers/ik/ik.py:59-82: per target position error |t(T_link^-1 T_target)|,
 * angle 2 asin(min(|q_vec|, 1)); success iff every target is strictly under both thresholds.
 */
EXPORT void ork_ik_metrics(const ork_robot *rb, int B, const real *q, const real *base, int ntargets,
                           const int32_t *links, const float *targets, float pos_thr, float ang_thr, real *pos_err,
                           real *ang_err, int32_t *success) {
    real *Tl = (real *)malloc(sizeof(real) * 7 * (rb->nl + 1));
    real *Xl = (real *)malloc(sizeof(real) * 7 * (rb->nj + 1));
    real *Tj = (real *)malloc(sizeof(real) * 7 * (rb->nj + 1));
    for (int b = 0; b < B; ++b) {
        const real *bs = base ? base + 7 * b : IDENT7;
        fk(rb, q + (size_t)b * rb->na, bs, Xl, Tj, Tl, NULL);
        int reached = 1;
        for (int i = 0; i < ntargets; ++i) {
            real inv[7], e[7], tg[7];
            for (int k = 0; k < 7; ++k) tg[k] = (real)targets[((size_t)b * ntargets + i) * 7 + k];
            se3_inverse(Tl + 7 * links[i], inv);
            se3_compose(inv, tg, e);
            real pe = RSQRT(e[0] * e[0] + e[1] * e[1] + e[2] * e[2]);
            real v = RSQRT(e[4] * e[4] + e[5] * e[5] + e[6] * e[6]);
            real ae = (real)2 * RASIN(v < (real)1 ? v : (real)1);
            pos_err[(size_t)b * ntargets + i] = pe;
            ang_err[(size_t)b * ntargets + i] = ae;
            if (pe >= (real)pos_thr || ae >= (real)ang_thr) reached = 0;
        }
        success[b] = reached;
    }
    free(Tl);
    free(Xl);
    free(Tj);
}


/* ------------------------------------------------------------------------- */
/* trajectory LM: T frames of one serial robot, joint-major per frame         */
/* ------------------------------------------------------------------------- */
/*
 * Terms, in this row order (the reference's cost sums term costs in list order):
 *   goal      TrajectoryTask(FrameTask) on the last frame (terms/trajectory_task.py:65-295,
 *             frame_task.py:459-502): 6 log rows
 *   velocity  TrajectorySmoothnessTask order 1 (terms/sparse/trajectory_smoothness_task.py:506-557):
 *             rows (row, dof), row = 0..T-2, r = (w / dt) (q[row+1] - q[row])
 *   accel.    TrajectorySmoothnessTask order 2: rows row = 0..T-2 with window start row-1; the first
 *             window reaches before frame 0 and has weight history_valid * w = 0 (no history), the
 *             others r = (w / dt^2) (q[s] - 2 q[s+1] + q[s+2])
 *   collision TrajectoryCollisionTask, primitive scene, penalty "plain", no sweep
 *             (terms/sparse/trajectory_collision_task.py:176-293, sphere_collision_kernels.py:19-42):
 *             rows (frame, sphere): sd = scene SDF at the sphere centre - radius,
 *             r = w (margin - sd) if sd < margin else 0; J = -w n^T (v + w x c) with n the SDF
 *             gradient (closest-point direction, trajectory_collision_task.py:62-72)
 * Scene SDF (geom/sdf_kernels.py:223-295): minimum over the obstacle spheres, then the
 * axis-aligned boxes, the first strict minimum kept.
 */
typedef struct {
    int T;
    int goal_link;
    const float *goal_targets; /* [B][7] */
    const float *goal_w;       /* [6] */
    float wv, wa, dt;          /* velocity / acceleration weights (0: term absent) */
    int nsph;
    const int32_t *sph_link;   /* [nsph] */
    const float *sph_c;        /* [nsph][3] link-frame centres */
    const float *sph_r;        /* [nsph] */
    float coll_w, margin;
    int nobs_s;
    const float *obs_s;        /* [n][4] centre, radius */
    int nobs_b;
    const float *obs_b;        /* [n][6] centre, half extents */
    int lim_orders;            /* bit o: TrajectoryLimitTask rows of order o, forward stencil */
    const float *lim_w, *lim_lo, *lim_hi; /* [3][na] */
    int cg_iters;              /* > 0: conjugate-gradient steps (traj_cg) instead of the Cholesky */
    int ngeom;                 /* scene volumes / meshes after the boxes */
    const ork_geom *geoms;
} ork_traj;

static int traj_rows(const ork_robot *rb, const ork_traj *tj, int *Rg, int *Rv, int *Ra, int *Rc, int *Rl) {
    *Rg = 6;
    *Rv = tj->wv != 0.f ? (tj->T - 1) * rb->na : 0;
    *Ra = tj->wa != 0.f ? (tj->T - 1) * rb->na : 0;
    *Rc = tj->nsph > 0 ? tj->T * tj->nsph : 0;
    *Rl = 0;
    for (int o = 0; o < 3; ++o)
        if ((tj->lim_orders >> o) & 1) *Rl += (tj->T - o) * rb->na;
    return *Rg + *Rv + *Ra + *Rc + *Rl;
}

/* scene SDF of a trajectory problem set */
static real scene_sdf(const ork_traj *tj, const real *p, real *n) {
    return scene_sdf_all(tj->nobs_s, tj->obs_s, tj->nobs_b, tj->obs_b, tj->ngeom, tj->geoms, p, n);
}

/* sparse row: up to 7 columns (one frame) or 3 (smoothness) */
typedef struct { int n; int col[8]; real val[8]; } traj_row;

/*
 * All rows of problem b at the trajectory q[T*na]: r[R], the rows' sparse Jacobian entries
 * (optionally dense J[R][T*na]) and, accumulated row by row like the sparse path
 * (terms/task.py:373-400, _sparse_normal_equations_kernel :469-488), H and g.
 * cost_terms[7] = 1/2 sum of squares of the goal, velocity, acceleration, collision rows and of the
 * bound rows of orders 0, 1, 2.
 *
 * Bound rows: TrajectoryLimitTask with the forward stencil (terms/sparse/trajectory_limit_task.py:
 * 81-118 stencils, _limit_kernel :436-498): for order o, window f = 0..T-1-o and dof d,
 * v = (sum_k c_k q[f+k][d]) / dt^o with c = (1), (-1, 1), (1, -2, 1) summed from k = 0,
 * r = w max(max(lo - v, v - hi), 0), Jacobian w sign c_k / dt^o on an active bound (sign +1 above,
 * -1 below), nothing otherwise.
 */
static int traj_eval(const ork_robot *rb, const ork_traj *tj, int b, const real *q, real *r, real *J, real *H,
                     real *g, real *cost_terms, traj_row *rows_out);

EXPORT int ork_traj_residual_jacobian(const ork_robot *rb, const ork_traj *tj, int b, const real *q, real *r,
                                      real *J, real *H, real *g, real *cost_terms) {
    return traj_eval(rb, tj, b, q, r, J, H, g, cost_terms, NULL);
}

/* as ork_traj_residual_jacobian; rows_out [R] (nullable) receives every row's sparse entries */
static int traj_eval(const ork_robot *rb, const ork_traj *tj, int b, const real *q, real *r, real *J, real *H,
                     real *g, real *cost_terms, traj_row *rows_out) {
    const int na = rb->na, T = tj->T, D = T * na;
    int Rg, Rv, Ra, Rc, Rl;
    const int R = traj_rows(rb, tj, &Rg, &Rv, &Ra, &Rc, &Rl);
    real *Xl = (real *)calloc(7 * (rb->nj + 1), sizeof(real));
    real *Tj = (real *)calloc(7 * (rb->nj + 1), sizeof(real));
    real *Tl = (real *)calloc(7 * (rb->nl + 1), sizeof(real));
    real *S = (real *)calloc(6 * (rb->nj + 1), sizeof(real));
    real *Jf = (real *)calloc(6 * na, sizeof(real));
    if (J) memset(J, 0, sizeof(real) * (size_t)R * D);
    if (H) memset(H, 0, sizeof(real) * (size_t)D * D);
    if (g) memset(g, 0, sizeof(real) * D);
    real ct[7] = {0, 0, 0, 0, 0, 0, 0};
    int row = 0;
    traj_row rw;
#define EMIT(term, rv)                                                                        \
    do {                                                                                      \
        r[row] = (rv);                                                                        \
        ct[term] += r[row] * r[row];                                                          \
        if (rows_out) rows_out[row] = rw;                                                     \
        for (int u = 0; u < rw.n; ++u) {                                                      \
            if (J) J[(size_t)row * D + rw.col[u]] = rw.val[u];                                \
            if (g) g[rw.col[u]] = g[rw.col[u]] + rw.val[u] * r[row];                          \
            if (H)                                                                            \
                for (int v = 0; v < rw.n; ++v)                                                \
                    H[(size_t)rw.col[u] * D + rw.col[v]] =                                    \
                        H[(size_t)rw.col[u] * D + rw.col[v]] + rw.val[u] * rw.val[v];          \
        }                                                                                     \
        ++row;                                                                                \
    } while (0)
    /* goal: the FrameTask log rows of the last frame */
    {
        const real *qT = q + (size_t)(T - 1) * na;
        memset(S, 0, sizeof(real) * 6 * rb->nj);
        fk(rb, qT, IDENT7, Xl, Tj, Tl, S);
        ork_term ft = {ORK_FRAME, 1, &tj->goal_link, tj->goal_targets, tj->goal_w, NULL};
        real rg[6];
        memset(Jf, 0, sizeof(real) * 6 * na);
        term_eval(rb, &ft, b, qT, Tl, S, rg, Jf);
        for (int i = 0; i < 6; ++i) {
            rw.n = na;
            for (int a = 0; a < na; ++a) { rw.col[a] = (T - 1) * na + a; rw.val[a] = Jf[i * na + a]; }
            EMIT(0, rg[i]);
        }
    }
    /* velocity rows */
    if (Rv) {
        const real w = (real)tj->wv / (real)tj->dt;
        for (int t = 0; t + 1 < T; ++t)
            for (int d = 0; d < na; ++d) {
                rw.n = 2;
                rw.col[0] = t * na + d; rw.val[0] = -w;
                rw.col[1] = (t + 1) * na + d; rw.val[1] = w;
                EMIT(1, w * (q[(t + 1) * na + d] - q[t * na + d]));
            }
    }
    /* acceleration rows (window start s = row - 1; the first window has zero weight) */
    if (Ra) {
        const real w = (real)tj->wa / ((real)tj->dt * (real)tj->dt);
        for (int rr = 0; rr + 1 < T; ++rr)
            for (int d = 0; d < na; ++d) {
                const int s0 = rr - 1;
                if (s0 < 0) {
                    rw.n = 0;
                    EMIT(2, (real)0);
                    continue;
                }
                rw.n = 3;
                rw.col[0] = s0 * na + d; rw.val[0] = w;
                rw.col[1] = (s0 + 1) * na + d; rw.val[1] = (real)-2 * w;
                rw.col[2] = (s0 + 2) * na + d; rw.val[2] = w;
                EMIT(2, w * (q[s0 * na + d] + (real)-2 * q[(s0 + 1) * na + d] + q[(s0 + 2) * na + d]));
            }
    }
    /* collision rows */
    if (Rc) {
        for (int t = 0; t < T; ++t) {
            const real *qt = q + (size_t)t * na;
            memset(S, 0, sizeof(real) * 6 * rb->nj);
            fk(rb, qt, IDENT7, Xl, Tj, Tl, S);
            for (int k = 0; k < tj->nsph; ++k) {
                const int link = tj->sph_link[k];
                real lc[3] = {(real)tj->sph_c[3 * k], (real)tj->sph_c[3 * k + 1], (real)tj->sph_c[3 * k + 2]};
                real c[3];
                const real *Tk = Tl + 7 * link;
                quat_apply(Tk + 3, lc, c);
                for (int i = 0; i < 3; ++i) c[i] += Tk[i];
                real n[3];
                const real sd = scene_sdf(tj, c, n) - (real)tj->sph_r[k];
                const real w = (real)tj->coll_w;
                real res = 0;
                rw.n = 0;
                if (sd < (real)tj->margin) {
                    res = w * ((real)tj->margin - sd);
                    real cn[3];
                    cross3(c, n, cn);
                    rw.n = na;
                    for (int a = 0; a < na; ++a) {
                        real sc[6];
                        spatial_column(rb, S, link, a, sc);
                        rw.col[a] = t * na + a;
                        rw.val[a] = -w * ((n[0] * sc[0] + n[1] * sc[1] + n[2] * sc[2]) +
                                          (cn[0] * sc[3] + cn[1] * sc[4] + cn[2] * sc[5]));
                    }
                }
                EMIT(3, res);
            }
        }
    }
    /* bound rows */
    for (int o = 0; o < 3 && Rl; ++o) {
        if (!((tj->lim_orders >> o) & 1)) continue;
        static const real cf[3][3] = {{1, 0, 0}, {-1, 1, 0}, {1, -2, 1}};
        real sc = 1;
        for (int k = 0; k < o; ++k) sc = sc * (real)tj->dt;
        sc = (real)1 / sc;
        for (int f = 0; f + o < T; ++f)
            for (int d = 0; d < na; ++d) {
                real v = 0;
                for (int k = 0; k <= o; ++k) v = v + cf[o][k] * q[(size_t)(f + k) * na + d];
                v = v * sc;
                const real w = (real)tj->lim_w[o * na + d];
                const real below = (real)tj->lim_lo[o * na + d] - v, above = v - (real)tj->lim_hi[o * na + d];
                const real m = below > above ? below : above;
                real res = 0;
                rw.n = 0;
                if (below > 0 || above > 0) {
                    res = w * (m > 0 ? m : 0);
                    const real sgn = above > 0 ? (real)1 : (real)-1;
                    rw.n = o + 1;
                    for (int k = 0; k <= o; ++k) {
                        rw.col[k] = (f + k) * na + d;
                        rw.val[k] = w * sgn * cf[o][k] * sc;
                    }
                }
                EMIT(4 + o, res);
            }
    }
#undef EMIT
    for (int i = 0; i < 7; ++i) ct[i] = (real)0.5 * ct[i];
    if (cost_terms) memcpy(cost_terms, ct, sizeof(ct));
    free(Xl); free(Tj); free(Tl); free(S); free(Jf);
    return R;
}

static real traj_cost_sum(const real *ct) {
    return (((((ct[0] + ct[1]) + ct[2]) + ct[3]) + ct[4]) + ct[5]) + ct[6];
}

static real traj_cost(const ork_robot *rb, const ork_traj *tj, int b, const real *q, real *r) {
    real ct[7];
    ork_traj_residual_jacobian(rb, tj, b, q, r, NULL, NULL, NULL, ct);
    return traj_cost_sum(ct);
}

/*
 * Matrix-free conjugate gradients on (J^T J + lam I) delta = -g over the sparse rows, the
 * reference's CG step (_compute_sparse_delta, opt/lm_optimizer.py:781-832, with the kernels of
 * :370-499): delta = 0, r = p = -g; per iteration Ap = J^T (J p) + lam p (_apply_linear_operator,
 * add_lambda_parallel_kernel), alpha = r.r / (p.Ap + 1e-8), delta += alpha p, r -= alpha Ap,
 * beta = r'.r' / (r.r + 1e-8), p = r' + beta p; `iters` iterations (the reference's budget without
 * its host-side early exit).  Then pred = -g.delta - delta.(J^T J delta) / 2
 * (compute_predicted_reduction_fixed_kernel).  work: 4 D + R reals.
 */
static void traj_cg(int D, int R, const traj_row *rows, const real *g, real lam, int iters, real *x, real *pred,
                    real *work) {
    real *res = work, *p = res + D, *Ap = p + D, *jtj = Ap + D, *jx = jtj + D;
    /* out = J^T (J v) (+ lam v when damped) */
    #define CG_APPLY(v, out, damped)                                                        \
        do {                                                                                \
            for (int i_ = 0; i_ < R; ++i_) {                                                \
                real a_ = 0;                                                                \
                for (int u_ = 0; u_ < rows[i_].n; ++u_) a_ += rows[i_].val[u_] * (v)[rows[i_].col[u_]]; \
                jx[i_] = a_;                                                                \
            }                                                                               \
            for (int d_ = 0; d_ < D; ++d_) jtj[d_] = 0;                                     \
            for (int i_ = 0; i_ < R; ++i_)                                                  \
                for (int u_ = 0; u_ < rows[i_].n; ++u_)                                     \
                    jtj[rows[i_].col[u_]] = jtj[rows[i_].col[u_]] + rows[i_].val[u_] * jx[i_]; \
            for (int d_ = 0; d_ < D; ++d_) (out)[d_] = (damped) ? jtj[d_] + lam * (v)[d_] : jtj[d_]; \
        } while (0)
    real rr = 0;
    for (int d = 0; d < D; ++d) {
        x[d] = 0;
        res[d] = -g[d];
        p[d] = res[d];
        rr += res[d] * res[d];
    }
    for (int k = 0; k < iters; ++k) {
        CG_APPLY(p, Ap, 1);
        real pap = 0;
        for (int d = 0; d < D; ++d) pap += p[d] * Ap[d];
        const real alpha = rr / (pap + (real)1e-8);
        real nr = 0;
        for (int d = 0; d < D; ++d) {
            x[d] = x[d] + alpha * p[d];
            res[d] = res[d] - alpha * Ap[d];
            nr += res[d] * res[d];
        }
        const real beta = nr / (rr + (real)1e-8);
        rr = nr;
        for (int d = 0; d < D; ++d) p[d] = res[d] + beta * p[d];
    }
    CG_APPLY(x, Ap, 0);
    #undef CG_APPLY
    real gd = 0, qd = 0;
    for (int d = 0; d < D; ++d) {
        gd += x[d] * g[d];
        qd += x[d] * Ap[d];
    }
    *pred = -gd - (real)0.5 * qd;
}

/* one LM step of trajectory problem b from (qb, lam); as the body of ork_lm_step */
typedef struct { real *r, *H, *L, *g, *dl, *qp, *cgw; traj_row *rows; } traj_ws;

static void traj_step_one(const ork_robot *rb, const ork_traj *tj, const ork_lm_params *p, int b, const real *qb,
                          real lam, traj_ws *w, real *cost_cur, real *cost_prop, real *pred_out, int32_t *acc_out,
                          real *q_new, real *lam_new) {
    const int D = tj->T * rb->na;
    real ct[7];
    const int cg = tj->cg_iters > 0;
    const int R = traj_eval(rb, tj, b, qb, w->r, NULL, cg ? NULL : w->H, w->g, ct, cg ? w->rows : NULL);
    real cost = traj_cost_sum(ct);
    *cost_cur = cost;
    real l = lam, pred;
    if (cg)
        traj_cg(D, R, w->rows, w->g, l, tj->cg_iters, w->dl, &pred, w->cgw);
    else
        damped_cholesky(D, w->H, w->g, l, w->L, w->dl, &pred);
    for (int d = 0; d < D; ++d) w->qp[d] = qb[d] + w->dl[d];
    real cp = traj_cost(rb, tj, b, w->qp, w->r);
    *cost_prop = cp;
    *pred_out = pred;
    int32_t acc;
    ork_accept(1, &cost, &cp, &pred, p, &l, &acc);
    *acc_out = acc;
    *lam_new = l;
    for (int d = 0; d < D; ++d) q_new[d] = acc ? w->qp[d] : qb[d];
}

static void traj_ws_alloc(traj_ws *w, const ork_robot *rb, const ork_traj *tj) {
    const int D = tj->T * rb->na;
    int Rg, Rv, Ra, Rc, Rl;
    const int R = traj_rows(rb, tj, &Rg, &Rv, &Ra, &Rc, &Rl);
    w->r = (real *)malloc(sizeof(real) * R);
    w->H = (real *)malloc(sizeof(real) * (size_t)D * D);
    w->L = (real *)malloc(sizeof(real) * (size_t)D * D);
    w->g = (real *)malloc(sizeof(real) * D);
    w->dl = (real *)malloc(sizeof(real) * D);
    w->qp = (real *)malloc(sizeof(real) * D);
    w->cgw = (real *)malloc(sizeof(real) * (4 * (size_t)D + R));
    w->rows = (traj_row *)malloc(sizeof(traj_row) * R);
}

static void traj_ws_free(traj_ws *w) {
    free(w->r); free(w->H); free(w->L); free(w->g); free(w->dl); free(w->qp); free(w->cgw); free(w->rows);
}

/* the linear step alone of problem b at q with damping lam: delta [D] and pred (the CG or the
 * Cholesky step, as tj->cg_iters selects) */
EXPORT void ork_traj_delta(const ork_robot *rb, const ork_traj *tj, int b, const real *q, real lam, real *delta,
                           real *pred) {
    const int D = tj->T * rb->na;
    traj_ws w;
    traj_ws_alloc(&w, rb, tj);
    real ct[7];
    const int cg = tj->cg_iters > 0;
    const int R = traj_eval(rb, tj, b, q, w.r, NULL, cg ? NULL : w.H, w.g, ct, cg ? w.rows : NULL);
    if (cg)
        traj_cg(D, R, w.rows, w.g, lam, tj->cg_iters, delta, pred, w.cgw);
    else
        damped_cholesky(D, w.H, w.g, lam, w.L, delta, pred);
    traj_ws_free(&w);
}

/* one LM step of each trajectory problem from (q[b], lam[b]); as ork_lm_step */
EXPORT void ork_traj_lm_step(const ork_robot *rb, const ork_traj *tj, const ork_lm_params *p, int B, const real *q,
                             const real *lam, real *cost_cur, real *cost_prop, real *pred_out, int32_t *acc_out,
                             real *q_new, real *lam_new) {
    const int D = tj->T * rb->na;
    traj_ws w;
    traj_ws_alloc(&w, rb, tj);
    for (int b = 0; b < B; ++b)
        traj_step_one(rb, tj, p, b, q + (size_t)b * D, lam[b], &w, cost_cur + b, cost_prop + b, pred_out + b,
                      acc_out + b, q_new + (size_t)b * D, lam_new + b);
    traj_ws_free(&w);
}

/* fixed-iteration trajectory LM (as ork_lm_solve); q [B][T*na] in/out */
EXPORT void ork_traj_lm_solve(const ork_robot *rb, const ork_traj *tj, const ork_lm_params *p, int B, int iters,
                              real *q, real *cost_out, real *cost_trace, int32_t *accept_trace) {
    const int D = tj->T * rb->na;
    traj_ws w;
    traj_ws_alloc(&w, rb, tj);
    real *qn = (real *)malloc(sizeof(real) * D);
    for (int b = 0; b < B; ++b) {
        real *qb = q + (size_t)b * D;
        real lam = (real)p->lm_lambda, cost = traj_cost(rb, tj, b, qb, w.r), cp, pred, ln, c0;
        int32_t acc;
        if (cost_trace) cost_trace[(size_t)b * (iters + 1)] = cost;
        for (int it = 0; it < iters; ++it) {
            traj_step_one(rb, tj, p, b, qb, lam, &w, &c0, &cp, &pred, &acc, qn, &ln);
            cost = acc ? cp : c0;
            memcpy(qb, qn, sizeof(real) * D);
            lam = ln;
            if (cost_trace) cost_trace[(size_t)b * (iters + 1) + it + 1] = cost;
            if (accept_trace) accept_trace[(size_t)b * iters + it] = acc;
        }
        if (cost_out) cost_out[b] = cost;
    }
    free(qn);
    traj_ws_free(&w);
}

/* goal pose errors at the last frame and the smallest sphere-to-scene distance over all frames */
EXPORT void ork_traj_metrics(const ork_robot *rb, const ork_traj *tj, int B, const real *q, real *pos_err,
                             real *ang_err, real *min_sd) {
    const int na = rb->na, T = tj->T;
    real *Xl = (real *)calloc(7 * (rb->nj + 1), sizeof(real));
    real *Tjt = (real *)calloc(7 * (rb->nj + 1), sizeof(real));
    real *Tl = (real *)calloc(7 * (rb->nl + 1), sizeof(real));
    for (int b = 0; b < B; ++b) {
        const real *qb = q + (size_t)b * T * na;
        real msd = (real)1e6;
        for (int t = 0; t < T; ++t) {
            fk(rb, qb + (size_t)t * na, IDENT7, Xl, Tjt, Tl, NULL);
            for (int k = 0; k < tj->nsph; ++k) {
                const real *Tk = Tl + 7 * tj->sph_link[k];
                real lc[3] = {(real)tj->sph_c[3 * k], (real)tj->sph_c[3 * k + 1], (real)tj->sph_c[3 * k + 2]}, c[3], n[3];
                quat_apply(Tk + 3, lc, c);
                for (int i = 0; i < 3; ++i) c[i] += Tk[i];
                real sd = scene_sdf(tj, c, n) - (real)tj->sph_r[k];
                if (sd < msd) msd = sd;
            }
            if (t == T - 1) {
                real inv[7], e[7], tg[7];
                for (int k = 0; k < 7; ++k) tg[k] = (real)tj->goal_targets[(size_t)b * 7 + k];
                se3_inverse(Tl + 7 * tj->goal_link, inv);
                se3_compose(inv, tg, e);
                pos_err[b] = RSQRT(e[0] * e[0] + e[1] * e[1] + e[2] * e[2]);
                real v = RSQRT(e[4] * e[4] + e[5] * e[5] + e[6] * e[6]);
                ang_err[b] = (real)2 * RASIN(v < (real)1 ? v : (real)1);
            }
        }
        min_sd[b] = msd;
    }
    free(Xl); free(Tjt); free(Tl);
}

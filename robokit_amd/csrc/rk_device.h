// rk_device.h -- SE(3) math and the device robot model for the batched LM kernels (gfx950).
//
// Pose = 7 floats [x,y,z,qw,qx,qy,qz]; twist = [v; w]; joint types FIXED 0 / PRISMATIC 1 /
// REVOLUTE 2.  The formulas and branch thresholds are the reference's:
//   quaternion ops      xform/warp/rotation_conversions.py:174-279
//   se3 compose/inverse lie/se3_kernels.py:69-112
//   se3 log, V^-1       lie/se3_kernels.py:230-309
//   se3 jlog            lie/se3_kernels.py:358-428, so3_jac_left_inv lie/so3_kernels.py:13-40
// written here as register-resident straight-line code (no 6x6 temporaries) for CDNA4 VALU.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define RK_NJ_MAX 64
#define RK_NL_MAX 64
#define RK_NA_MAX 32

namespace rk {

struct Pose {
  float t[3];
  float q[4];  // w x y z
};

struct Mat3 {
  float m[9];  // row-major
};

// ---------------------------------------------------------------------------------------------
// device robot model: one struct per uploaded robot (host fills it from RobotSpec arrays).
// ---------------------------------------------------------------------------------------------
struct Robot {
  int nj, nl, na, nlevels;
  int chain;    // serial arm: parent(j) = j-1, joints 0..na-1 actuated 1:1, the rest fixed
  int nrounds;  // pointer-jumping rounds that cover the deepest chain: ceil(log2(depth + 1))
  int joint_type[RK_NJ_MAX];
  int parent[RK_NJ_MAX];
  int act[RK_NJ_MAX];    // actuated index or -1
  int mimic[RK_NJ_MAX];  // mimicked actuated index or -1
  float mul[RK_NJ_MAX];
  float off[RK_NJ_MAX];
  float axis[RK_NJ_MAX][3];
  float pT[RK_NJ_MAX][7];      // parent joint transforms
  float pqa[RK_NJ_MAX][4];     // pT's rotation times the pure axis quaternion: q_P (0, axis)
  int link_joint[RK_NL_MAX];   // link -> parent joint (-1 = base)
  float lo[RK_NA_MAX], hi[RK_NA_MAX];  // float32-narrowed actuated limits
  int order[RK_NJ_MAX];        // joints sorted by depth (identity: joints are renumbered in depth order)
  int level_start[RK_NJ_MAX + 1];
  unsigned long long anc[RK_NL_MAX];  // bit j: joint j is an ancestor of link (incl. its parent joint)
  int col_start[RK_NA_MAX + 1];       // CSR over actuated dofs: joints j (ascending) with map[j][a] != 0
  int col_joint[RK_NJ_MAX];
  float col_w[RK_NJ_MAX];
};

// ---------------------------------------------------------------------------------------------
// quaternion / SE(3)
// ---------------------------------------------------------------------------------------------
__host__ __device__ __forceinline__ void cross3(const float* a, const float* b, float* o) {
  float x = a[1] * b[2] - a[2] * b[1];
  float y = a[2] * b[0] - a[0] * b[2];
  float z = a[0] * b[1] - a[1] * b[0];
  o[0] = x; o[1] = y; o[2] = z;
}

// quaternion_multiply_func + standardize (rotation_conversions.py:259-279)
__host__ __device__ __forceinline__ void quat_mul(const float* a, const float* b, float* o) {
  float w = a[0] * b[0] - a[1] * b[1] - a[2] * b[2] - a[3] * b[3];
  float x = a[0] * b[1] + a[1] * b[0] + a[2] * b[3] - a[3] * b[2];
  float y = a[0] * b[2] - a[1] * b[3] + a[2] * b[0] + a[3] * b[1];
  float z = a[0] * b[3] + a[1] * b[2] - a[2] * b[1] + a[3] * b[0];
  float s = w < 0.f ? -1.f : 1.f;
  o[0] = s * w; o[1] = s * x; o[2] = s * y; o[3] = s * z;
}

// v + 2 (w (u x v) + u x (u x v))   (rotation_conversions.py:193-200)
__host__ __device__ __forceinline__ void quat_apply(const float* q, const float* v, float* o) {
  float u[3] = {q[1], q[2], q[3]};
  float uv[3], uuv[3];
  cross3(u, v, uv);
  cross3(u, uv, uuv);
  o[0] = v[0] + 2.f * (q[0] * uv[0] + uuv[0]);
  o[1] = v[1] + 2.f * (q[0] * uv[1] + uuv[1]);
  o[2] = v[2] + 2.f * (q[0] * uv[2] + uuv[2]);
}

__host__ __device__ __forceinline__ Mat3 quat_to_mat(const float* q) {
  float w = q[0], x = q[1], y = q[2], z = q[3];
  float s = 2.f / (w * w + x * x + y * y + z * z);
  Mat3 R;
  R.m[0] = 1.f - s * (y * y + z * z);
  R.m[1] = s * (x * y - z * w);
  R.m[2] = s * (x * z + y * w);
  R.m[3] = s * (x * y + z * w);
  R.m[4] = 1.f - s * (x * x + z * z);
  R.m[5] = s * (y * z - x * w);
  R.m[6] = s * (x * z - y * w);
  R.m[7] = s * (y * z + x * w);
  R.m[8] = 1.f - s * (x * x + y * y);
  return R;
}

__host__ __device__ __forceinline__ void mat3_vec(const Mat3& A, const float* v, float* o) {
  float a = A.m[0] * v[0] + A.m[1] * v[1] + A.m[2] * v[2];
  float b = A.m[3] * v[0] + A.m[4] * v[1] + A.m[5] * v[2];
  float c = A.m[6] * v[0] + A.m[7] * v[1] + A.m[8] * v[2];
  o[0] = a; o[1] = b; o[2] = c;
}

__host__ __device__ __forceinline__ Pose compose(const Pose& a, const Pose& b) {
  Pose o;
  quat_mul(a.q, b.q, o.q);
  float t[3];
  quat_apply(a.q, b.t, t);
  o.t[0] = a.t[0] + t[0]; o.t[1] = a.t[1] + t[1]; o.t[2] = a.t[2] + t[2];
  return o;
}

__host__ __device__ __forceinline__ Pose inverse(const Pose& a) {
  Pose o;
  o.q[0] = a.q[0]; o.q[1] = -a.q[1]; o.q[2] = -a.q[2]; o.q[3] = -a.q[3];
  float t[3];
  quat_apply(o.q, a.t, t);
  o.t[0] = -t[0]; o.t[1] = -t[1]; o.t[2] = -t[2];
  return o;
}

// quaternion_to_axis_angle_func (rotation_conversions.py:226-246)
__device__ __forceinline__ void quat_to_axis_angle(const float* q, float* aa) {
  float norm = sqrtf(q[1] * q[1] + q[2] * q[2] + q[3] * q[3]);
  float half = atan2f(norm, q[0]);
  float angle = 2.f * half;
  float s;
  if (fabsf(angle) < 1e-6f) s = 0.5f - (angle * angle) / 48.f;
  else s = sinf(half) / angle;
  aa[0] = q[1] / s; aa[1] = q[2] / s; aa[2] = q[3] / s;
}

// hat(w)^2 = w w^T - |w|^2 I ; these helpers form I + a K + b K^2 applied to v without 3x3 temporaries.
__device__ __forceinline__ void ikk_apply(float a, float b, const float* w, const float* v, float* o) {
  // (I + a K + b K^2) v = v + a (w x v) + b (w x (w x v))
  float wv[3], wwv[3];
  cross3(w, v, wv);
  cross3(w, wv, wwv);
  o[0] = v[0] + a * wv[0] + b * wwv[0];
  o[1] = v[1] + a * wv[1] + b * wwv[1];
  o[2] = v[2] + a * wv[2] + b * wwv[2];
}

// se3_log_map_func with se3_V_matrix_inverse_func (se3_kernels.py:230-309), eps = 1e-4
__device__ __forceinline__ void se3_log(const Pose& T, float* out) {
  const float eps = 1e-4f;
  float w[3];
  quat_to_axis_angle(T.q, w);
  float angle = sqrtf(w[0] * w[0] + w[1] * w[1] + w[2] * w[2]);
  float c2;
  if (angle < eps) {
    c2 = 1.f / 12.f;
  } else {
    float sa, ca;
    sincosf(angle, &sa, &ca);
    float omc = 1.f - ca;
    if (fabsf(omc) < eps) c2 = 1.f / 12.f;
    else c2 = (1.f / (angle * angle)) * (1.f - (angle * sa) / (2.f * omc));
  }
  float v[3];
  ikk_apply(-0.5f, c2, w, T.t, v);
  out[0] = v[0]; out[1] = v[1]; out[2] = v[2];
  out[3] = w[0]; out[4] = w[1]; out[5] = w[2];
}

// se3_jlog_func (se3_kernels.py:358-428), eps = 1e-4.  Returns the 3x3 blocks of
// [[A, B],[0, A]] with A = so3_jac_left_inv(w)^T.
__device__ __forceinline__ void se3_jlog_blocks(const Pose& T, Mat3& A, Mat3& B) {
  const float eps = 1e-4f;
  const float* p = T.t;
  float w[3];
  quat_to_axis_angle(T.q, w);
  float t2 = w[0] * w[0] + w[1] * w[1] + w[2] * w[2];
  float th = sqrtf(t2);
  bool taylor = t2 < eps;
  // so3_jac_left_inv (so3_kernels.py:13-40): I - 0.5 K + coeff K^2 ; transpose = I + 0.5 K + coeff K^2
  float t2s = taylor ? 1.f : t2;
  float coeff;
  if (taylor) {
    coeff = t2s / 12.f;
  } else {
    float ts = sqrtf(t2s);
    float sh, ch;
    sincosf(ts / 2.f, &sh, &ch);
    coeff = (1.f - ts * ch / (2.f * sh)) / t2s;
  }
  // K^2 = w w^T - t2 I
  float K2[9] = {w[0] * w[0] - t2, w[0] * w[1], w[0] * w[2],
                 w[1] * w[0], w[1] * w[1] - t2, w[1] * w[2],
                 w[2] * w[0], w[2] * w[1], w[2] * w[2] - t2};
  float K[9] = {0.f, -w[2], w[1], w[2], 0.f, -w[0], -w[1], w[0], 0.f};
#pragma unroll
  for (int i = 0; i < 9; ++i) A.m[i] = ((i % 4) == 0 ? 1.f : 0.f) + 0.5f * K[i] + coeff * K2[i];
  if (taylor) {
    float sk[9] = {0.f, -p[2], p[1], p[2], 0.f, -p[0], -p[1], p[0], 0.f};
#pragma unroll
    for (int i = 0; i < 9; ++i) B.m[i] = 0.5f * sk[i];
    return;
  }
  float th_inv = 1.f / th;
  float th2_inv = th_inv * th_inv;
  float st, ct;
  sincosf(th, &st, &ct);
  float inv_2_2ct = 1.f / (2.f * (1.f - ct));
  float beta = th2_inv - st * th_inv * inv_2_2ct;
  float bdot = -2.f * th2_inv * th2_inv + (1.f + st * th_inv) * th2_inv * inv_2_2ct;
  float wTp = w[0] * p[0] + w[1] * p[1] + w[2] * p[2];
  float v3[3];
#pragma unroll
  for (int i = 0; i < 3; ++i) v3[i] = bdot * wTp * w[i] - (t2 * bdot + 2.f * beta) * p[i];
  float Cm[9];
  float sk[9] = {0.f, -p[2], p[1], p[2], 0.f, -p[0], -p[1], p[0], 0.f};
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j)
      Cm[i * 3 + j] = v3[i] * w[j] + beta * (w[i] * p[j]) + (i == j ? wTp * beta : 0.f) + 0.5f * sk[i * 3 + j];
  // B = C * jlog_so3^T = C * A
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j)
      B.m[i * 3 + j] = Cm[i * 3 + 0] * A.m[0 * 3 + j] + Cm[i * 3 + 1] * A.m[1 * 3 + j] + Cm[i * 3 + 2] * A.m[2 * 3 + j];
}

// sin and cos of a joint's half angle.  Cody-Waite reduction by pi/2 in three parts (exact
// products for |k| < 2^12) and the single-precision minimax polynomials of the Cephes library
// (S. L. Moshier, "Methods and Programs for Mathematical Functions", 1989: sinf.c / cosf.c) on
// [-pi/4, pi/4], then the quadrant's swap and signs; about 20 instructions against ~40 for the
// library's, whose large-argument path (Payne-Hanek) serves |x| > 4096.
__device__ __forceinline__ void rk_sincos(float x, float* sp, float* cp) {
  if (fabsf(x) > 4096.f) {
    sincosf(x, sp, cp);
    return;
  }
  const float k = __builtin_rintf(x * 0.636619772367581343f);  // x / (pi / 2), nearest
  float r = fmaf(-k, 1.5703125f, x);
  r = fmaf(-k, 4.837512969970703125e-4f, r);
  r = fmaf(-k, 7.54978995489188216e-8f, r);
  const float z = r * r;
  const float sr = fmaf(fmaf(fmaf(-1.9515295891e-4f, z, 8.3321608736e-3f), z, -1.6666654611e-1f) * z, r, r);
  const float cr =
      fmaf(fmaf(fmaf(2.443315711809948e-5f, z, -1.388731625493765e-3f), z, 4.166664568298827e-2f), z * z,
           fmaf(-0.5f, z, 1.f));
  const int q = (int)k;
  const float s1 = (q & 1) ? cr : sr, c1 = (q & 1) ? sr : cr;
  *sp = (q & 2) ? -s1 : s1;
  *cp = ((q + 1) & 2) ? -c1 : c1;
}

// joint local transform P_j * M(q_j)   (_compute_joint_local_transform_func, robot_kernels.py:264-297)
template <class RB>
__device__ __forceinline__ Pose joint_local(RB rb, int j, const float* q) {
  float pos;
  int mi = rb->mimic[j];
  if (mi == -1) pos = rb->act[j] == -1 ? 0.f : q[rb->act[j]];
  else pos = q[mi] * rb->mul[j] + rb->off[j];
  Pose P;
#pragma unroll
  for (int k = 0; k < 3; ++k) P.t[k] = rb->pT[j][k];
#pragma unroll
  for (int k = 0; k < 4; ++k) P.q[k] = rb->pT[j][3 + k];
  Pose X;
  X.t[0] = X.t[1] = X.t[2] = 0.f;
  X.q[0] = 1.f; X.q[1] = X.q[2] = X.q[3] = 0.f;
  int jt = rb->joint_type[j];
  if (jt == 2) {
    float s, c;
    rk_sincos(pos * 0.5f, &s, &c);
    X.q[0] = c; X.q[1] = rb->axis[j][0] * s; X.q[2] = rb->axis[j][1] * s; X.q[3] = rb->axis[j][2] * s;
  } else if (jt == 1) {
    X.t[0] = rb->axis[j][0] * pos; X.t[1] = rb->axis[j][1] * pos; X.t[2] = rb->axis[j][2] * pos;
  }
  return compose(P, X);
}

__device__ __forceinline__ void load_pose(const float* p, Pose& T) {
  T.t[0] = p[0]; T.t[1] = p[1]; T.t[2] = p[2];
  T.q[0] = p[3]; T.q[1] = p[4]; T.q[2] = p[5]; T.q[3] = p[6];
}
__device__ __forceinline__ void store_pose(float* p, const Pose& T) {
  p[0] = T.t[0]; p[1] = T.t[1]; p[2] = T.t[2];
  p[3] = T.q[0]; p[4] = T.q[1]; p[5] = T.q[2]; p[6] = T.q[3];
}

// world motion-subspace column S_j = Ad(T_wj) twist_j  (robot_kernels.py:650-664)
template <class RB>
__device__ __forceinline__ void subspace_column(RB rb, int j, const Pose& T, float* S6) {
  int jt = rb->joint_type[j];
  float a[3] = {rb->axis[j][0], rb->axis[j][1], rb->axis[j][2]};
  float Ra[3];
  quat_apply(T.q, a, Ra);
  if (jt == 2) {
    float txRa[3];
    cross3(T.t, Ra, txRa);
    S6[0] = txRa[0]; S6[1] = txRa[1]; S6[2] = txRa[2];
    S6[3] = Ra[0]; S6[4] = Ra[1]; S6[5] = Ra[2];
  } else if (jt == 1) {
    S6[0] = Ra[0]; S6[1] = Ra[1]; S6[2] = Ra[2];
    S6[3] = S6[4] = S6[5] = 0.f;
  } else {
    S6[0] = S6[1] = S6[2] = S6[3] = S6[4] = S6[5] = 0.f;
  }
}

}  // namespace rk

// rk_scene.h -- primitive scene distances shared by the IK collision rows and the trajectory
// kernel (included by rk_lm.hip inside namespace rk, before the kernels).  A scene holder is a
// pointer to any table with ns obstacle spheres obs_s[][4] (centre, radius) and nb axis-aligned
// boxes obs_b[][6] (centre, half extents), padded with far-away primitives to whole groups.

// scene signed distance at p: minimum over obstacle spheres, then boxes
// (geom/sdf_kernels.py:256-295, trajectory_collision_task.py:62-72); 1e6 when no primitive is
// closer.  The tables are padded with far-away primitives to whole groups (8 spheres, 4 boxes);
// a group's data is read through the scalar cache before any of it is used, and its distances
// are reduced by a min tree (exact, so the order does not change the value).
#define RK_SDF_NONE (1 << 30)
#define RK_SDF_FAR 1e18f  // centre coordinate of a padding primitive (its squared distance stays finite)
__device__ __forceinline__ float sdf_sphere(const float* p, const float (&o)[4]) {
  const float d0 = p[0] - o[0], d1 = p[1] - o[1], d2 = p[2] - o[2];
  return __builtin_amdgcn_sqrtf(d0 * d0 + d1 * d1 + d2 * d2) - o[3];
}
__device__ __forceinline__ float sdf_box(const float* p, const float (&o)[6]) {
  float qv[3];
#pragma unroll
  for (int k = 0; k < 3; ++k) qv[k] = fabsf(p[k] - o[k]) - o[3 + k];
  const float qmax = fmaxf(fmaxf(qv[0], qv[1]), qv[2]);
  const float p0 = fmaxf(qv[0], 0.f), p1 = fmaxf(qv[1], 0.f), p2 = fmaxf(qv[2], 0.f);
  return __builtin_amdgcn_sqrtf(p0 * p0 + p1 * p1 + p2 * p2) + fminf(qmax, 0.f);
}
// also returns in grp the first group holding the minimum (sphere groups 0..7, box groups 8..11;
// -1: none closer than 1e6), for scene_winner
template <class SC>
__device__ __forceinline__ float scene_sdf(SC tt, const float* p, int& grp) {
  float best = 1e6f;
  grp = -1;
  const int ns = tt->ns, nb = tt->nb;
  for (int i0 = 0; i0 < ns; i0 += 8) {
    float o[8][4], d[8];
#pragma unroll
    for (int u = 0; u < 8; ++u)
#pragma unroll
      for (int c = 0; c < 4; ++c) o[u][c] = tt->obs_s[i0 + u][c];
#pragma unroll
    for (int u = 0; u < 8; ++u) d[u] = sdf_sphere(p, o[u]);
    const float gm = fminf(fminf(fminf(d[0], d[1]), fminf(d[2], d[3])), fminf(fminf(d[4], d[5]), fminf(d[6], d[7])));
    grp = gm < best ? i0 / 8 : grp;
    best = fminf(best, gm);
  }
  for (int i0 = 0; i0 < nb; i0 += 4) {
    float o[4][6], d[4];
#pragma unroll
    for (int u = 0; u < 4; ++u)
#pragma unroll
      for (int c = 0; c < 6; ++c) o[u][c] = tt->obs_b[i0 + u][c];
#pragma unroll
    for (int u = 0; u < 4; ++u) d[u] = sdf_box(p, o[u]);
    const float gm = fminf(fminf(d[0], d[1]), fminf(d[2], d[3]));
    grp = gm < best ? 8 + i0 / 4 : grp;
    best = fminf(best, gm);
  }
  return best;
}

// the primitive that decides the scene distance at p: the first strict minimum in the
// reference's order (spheres, then boxes), found inside grp, the first group holding the minimum
// (the groups' minima are compared strictly in order, so the first group wins, and inside it the
// strict scan keeps the first).  win >= 0: sphere win; win < 0: box -1 - win; RK_SDF_NONE: no
// primitive closer than 1e6.  Only needed for a gradient, so only where a sphere is within the
// collision margin; grp differs between lanes, so the group is read with per-lane addresses.
template <class SC>
__device__ __forceinline__ int scene_winner(SC tt, const float* p, int grp) {
  if (grp < 0) return RK_SDF_NONE;
  float best = 1e6f;
  int win = RK_SDF_NONE;
  if (grp < 8) {
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      float o[4];
#pragma unroll
      for (int c = 0; c < 4; ++c) o[c] = tt->obs_s[8 * grp + u][c];
      const float d = sdf_sphere(p, o);
      win = d < best ? 8 * grp + u : win;
      best = d < best ? d : best;
    }
  } else {
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      float o[6];
#pragma unroll
      for (int c = 0; c < 6; ++c) o[c] = tt->obs_b[4 * (grp - 8) + u][c];
      const float d = sdf_box(p, o);
      win = d < best ? -1 - (4 * (grp - 8) + u) : win;
      best = d < best ? d : best;
    }
  }
  return win;
}

template <class SC>
__device__ __forceinline__ void scene_normal(SC tt, const float* p, int win, float* n) {
  n[0] = 1.f;
  n[1] = n[2] = 0.f;
  if (win == RK_SDF_NONE) return;  // the reference's initial direction
  if (win >= 0) {
    const float d0 = p[0] - tt->obs_s[win][0], d1 = p[1] - tt->obs_s[win][1], d2 = p[2] - tt->obs_s[win][2];
    const float len = sqrtf(d0 * d0 + d1 * d1 + d2 * d2);
    if (len > 1e-10f) {
      n[0] = d0 / len; n[1] = d1 / len; n[2] = d2 / len;
    }
    return;
  }
  const int i = -1 - win;
  float ql[3], qv[3], h[3];
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    ql[k] = p[k] - tt->obs_b[i][k];
    h[k] = tt->obs_b[i][3 + k];
    qv[k] = fabsf(ql[k]) - h[k];
  }
  float qmax = qv[0] > qv[1] ? qv[0] : qv[1];
  if (qv[2] > qmax) qmax = qv[2];
  float nn[3] = {0.f, 0.f, 1.f};
  if (qmax < 0.f) {
    if (qv[0] >= qv[1] && qv[0] >= qv[2]) {
      nn[0] = ql[0] < 0.f ? -1.f : 1.f;
      nn[2] = 0.f;
    } else if (qv[1] >= qv[2]) {
      nn[1] = ql[1] < 0.f ? -1.f : 1.f;
      nn[2] = 0.f;
    } else {
      nn[2] = ql[2] < 0.f ? -1.f : 1.f;
    }
  } else {
    float df[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      const float c = ql[k] < -h[k] ? -h[k] : (ql[k] > h[k] ? h[k] : ql[k]);
      df[k] = ql[k] - c;
    }
    const float dl = sqrtf(df[0] * df[0] + df[1] * df[1] + df[2] * df[2]);
    if (dl > 1e-10f) {
      nn[0] = df[0] / dl; nn[1] = df[1] / dl; nn[2] = df[2] / dl;
    }
  }
  n[0] = nn[0]; n[1] = nn[1]; n[2] = nn[2];
}


// rk_scene.h -- scene distances shared by the IK collision rows and the trajectory kernel
// (included by rk_lm.hip inside namespace rk, before the kernels).  A scene holder is a pointer
// to any table with ns obstacle spheres obs_s[][4] (centre, radius) and nb axis-aligned boxes
// obs_b[][6] (centre, half extents), padded with far-away primitives to whole groups, then ng
// volumes / meshes geom[] (SceneGeom).

// scene signed distance at p: minimum over obstacle spheres, then boxes
// (geom/sdf_kernels.py:256-295, trajectory_collision_task.py:62-72); 1e6 when no primitive is
// closer.  The tables are padded with far-away primitives to whole groups (8 spheres, 4 boxes);
// a group's data is read through the scalar cache before any of it is used, and its distances
// are reduced by a min tree (exact, so the order does not change the value).
__device__ __forceinline__ float dot3(const float* a, const float* b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }

#define RK_SDF_NONE (1 << 30)
#define RK_SDF_GEOM_GRP 16   // scene_sdf group code of geometry k: 16 + k
#define RK_SDF_GEOM_WIN (-1000)  // scene_winner code of geometry k: -1000 - k
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
// ---- volumes and meshes (SceneGeom; geom/sdf_kernels.py:151-222, 346-380)

// signed distance of a volume at p; grad (nullable) receives the direction of its gradient:
// inside [lo, hi] the trilinear interpolant of the 8 grid values around p and its gradient in
// index space (a uniform grid: the world direction), outside the box distance + pad and the
// direction from the box
__device__ __forceinline__ float volume_sdf(const SceneGeom& g, const float* p, float* grad) {
  const bool inside = p[0] >= g.lo[0] && p[0] <= g.hi[0] && p[1] >= g.lo[1] && p[1] <= g.hi[1] && p[2] >= g.lo[2] &&
                      p[2] <= g.hi[2];
  if (inside) {
    float f[3];
    int i[3];
    const int n[3] = {g.nx, g.ny, g.nz};
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      const float u = (p[k] - g.org[k]) * g.inv_h;
      i[k] = min(max((int)floorf(u), 0), n[k] - 2);
      f[k] = u - (float)i[k];
    }
    const float* v = g.data + ((size_t)i[2] * g.ny + i[1]) * g.nx + i[0];
    const size_t sy = g.nx, sz = (size_t)g.nx * g.ny;
    const float v000 = v[0], v100 = v[1], v010 = v[sy], v110 = v[sy + 1];
    const float v001 = v[sz], v101 = v[sz + 1], v011 = v[sz + sy], v111 = v[sz + sy + 1];
    const float c00 = v000 + f[0] * (v100 - v000), c10 = v010 + f[0] * (v110 - v010);
    const float c01 = v001 + f[0] * (v101 - v001), c11 = v011 + f[0] * (v111 - v011);
    const float c0 = c00 + f[1] * (c10 - c00), c1 = c01 + f[1] * (c11 - c01);
    if (grad) {
      const float e00 = v100 - v000, e10 = v110 - v010, e01 = v101 - v001, e11 = v111 - v011;
      const float e0 = e00 + f[1] * (e10 - e00), e1 = e01 + f[1] * (e11 - e01);
      const float d0 = c10 - c00, d1 = c11 - c01;
      grad[0] = e0 + f[2] * (e1 - e0);
      grad[1] = d0 + f[2] * (d1 - d0);
      grad[2] = c1 - c0;
    }
    return c0 + f[2] * (c1 - c0);
  }
  float o[3];
#pragma unroll
  for (int k = 0; k < 3; ++k) o[k] = p[k] - fminf(fmaxf(p[k], g.lo[k]), g.hi[k]);
  const float ad = sqrtf(o[0] * o[0] + o[1] * o[1] + o[2] * o[2]);
  if (grad) {
    grad[0] = o[0] / ad;
    grad[1] = o[1] / ad;
    grad[2] = o[2] / ad;
  }
  return ad + g.pad;
}

// signed distance of a mesh at p: the nearest point of each triangle (the plane projection when
// its barycentric coordinates are all >= 0, else the nearest of the three clamped edge points --
// no per-lane branches), the sign of (p - nearest) . the pseudonormal of the nearest feature
// (face; edge interior; edge end = vertex) of the nearest triangle, read once after the scan.
// The first strict minimum over the triangles wins; face receives its index.  A lane whose point
// is outside the mesh's box by at least `cut` cannot be closer than cut (the mesh lies in the
// box): when every lane is such, the triangles are skipped, and such a lane's result is 3e38
// (never a minimum).
__device__ __forceinline__ float mesh_sdf(const SceneGeom& g, const float* p, float cut, int& face,
                                          float* dir = nullptr) {
  float o[3];
#pragma unroll
  for (int k = 0; k < 3; ++k) o[k] = p[k] - fminf(fmaxf(p[k], g.lo[k]), g.hi[k]);
  const float ad = sqrtf(o[0] * o[0] + o[1] * o[1] + o[2] * o[2]);
  const bool culled = ad > 0.f && ad >= cut;
  face = -1;
  if (__all(culled)) return 3e38f;
  typedef const __attribute__((address_space(4))) float TriF;
  TriF* tri = (TriF*)g.data;
  float best2 = 3e38f, br0 = 0.f, br1 = 0.f, br2 = 1.f;
  int bpn = 9;
  for (int t = 0; t < g.nx; ++t) {
    TriF* T = tri + t * RK_TRI_FLOATS;
    const float a0 = T[0], a1 = T[1], a2 = T[2], b0 = T[3], b1 = T[4], b2 = T[5], c0 = T[6], c1 = T[7], c2 = T[8];
    const float ab0 = b0 - a0, ab1 = b1 - a1, ab2 = b2 - a2, ac0 = c0 - a0, ac1 = c1 - a1, ac2 = c2 - a2;
    const float ap0 = p[0] - a0, ap1 = p[1] - a1, ap2 = p[2] - a2;
    const float bc0 = c0 - b0, bc1 = c1 - b1, bc2 = c2 - b2, bp0 = p[0] - b0, bp1 = p[1] - b1, bp2 = p[2] - b2;
    const float ca0 = a0 - c0, ca1 = a1 - c1, ca2 = a2 - c2, cp0 = p[0] - c0, cp1 = p[1] - c1, cp2 = p[2] - c2;
    const float d00 = ab0 * ab0 + ab1 * ab1 + ab2 * ab2, d01 = ab0 * ac0 + ab1 * ac1 + ab2 * ac2;
    const float d11 = ac0 * ac0 + ac1 * ac1 + ac2 * ac2, d20 = ap0 * ab0 + ap1 * ab1 + ap2 * ab2;
    const float d21 = ap0 * ac0 + ap1 * ac1 + ap2 * ac2;
    const float den = d00 * d11 - d01 * d01;
    const float v = (d11 * d20 - d01 * d21) / den, w = (d00 * d21 - d01 * d20) / den;
    const bool inside = v >= 0.f && w >= 0.f && v + w <= 1.f;  // NaN (a degenerate triangle): false
    const float t0 = fminf(fmaxf(d20 / d00, 0.f), 1.f);
    const float t1 = fminf(fmaxf((bp0 * bc0 + bp1 * bc1 + bp2 * bc2) / (bc0 * bc0 + bc1 * bc1 + bc2 * bc2), 0.f), 1.f);
    const float t2 = fminf(fmaxf((cp0 * ca0 + cp1 * ca1 + cp2 * ca2) / (ca0 * ca0 + ca1 * ca1 + ca2 * ca2), 0.f), 1.f);
    // p minus each candidate point: edges ab, bc, ca, then the face
    const float x0 = ap0 - t0 * ab0, y0 = ap1 - t0 * ab1, z0 = ap2 - t0 * ab2;
    const float x1 = bp0 - t1 * bc0, y1 = bp1 - t1 * bc1, z1 = bp2 - t1 * bc2;
    const float x2 = cp0 - t2 * ca0, y2 = cp1 - t2 * ca1, z2 = cp2 - t2 * ca2;
    const float e0 = x0 * x0 + y0 * y0 + z0 * z0, e1 = x1 * x1 + y1 * y1 + z1 * z1, e2 = x2 * x2 + y2 * y2 + z2 * z2;
    // nearest edge point (first strict minimum); feature offsets of the pseudonormals in the
    // triangle's table: edges ab 12, bc 15, ca 18; vertices a 21, b 24, c 27; face 9
    float e = e0, tt = t0, rx = x0, ry = y0, rz = z0;
    int pe = 12, ps = 21, pf = 24;
    if (e1 < e) { e = e1; tt = t1; rx = x1; ry = y1; rz = z1; pe = 15; ps = 24; pf = 27; }
    if (e2 < e) { e = e2; tt = t2; rx = x2; ry = y2; rz = z2; pe = 18; ps = 27; pf = 21; }
    int pn = tt <= 0.f ? ps : (tt >= 1.f ? pf : pe);
    if (inside) {
      rx = ap0 - v * ab0 - w * ac0;
      ry = ap1 - v * ab1 - w * ac1;
      rz = ap2 - v * ab2 - w * ac2;
      e = rx * rx + ry * ry + rz * rz;
      pn = 9;
    }
    if (e < best2) {
      best2 = e;
      br0 = rx; br1 = ry; br2 = rz;
      bpn = pn;
      face = t;
    }
  }
  if (culled || face < 0) {
    face = -1;
    return 3e38f;
  }
  const float* pnv = g.data + (size_t)face * RK_TRI_FLOATS + bpn;  // this lane's feature
  const float sd = br0 * pnv[0] + br1 * pnv[1] + br2 * pnv[2];
  const float d = sqrtf(best2);
  if (dir) {  // the outward direction from the nearest point; the face normal on the surface
    const float s = sd < 0.f ? -1.f : 1.f;
    if (d > 1e-12f) {
      dir[0] = s * br0 / d; dir[1] = s * br1 / d; dir[2] = s * br2 / d;
    } else {
      const float* fn = g.data + (size_t)face * RK_TRI_FLOATS + 9;
      dir[0] = fn[0]; dir[1] = fn[1]; dir[2] = fn[2];
    }
  }
  return sd < 0.f ? -d : d;
}

// Volumes and meshes are evaluated out of line: the register-resident LM kernels inline the scene
// query once per joint of the chain, and inlined geometry code there pushed their private arrays
// to scratch.  Calls happen only when a scene has geometries; kernels instantiated with GEO false
// (the trajectory kernel without bound rows or geometries) compile no geometry code at all.
typedef const __attribute__((address_space(4))) SceneGeom SceneGeomK;

__device__ __forceinline__ SceneGeom load_geom(SceneGeomK* gk) {
  SceneGeom g;
  g.kind = gk->kind;
  g.data = gk->data;
  g.nx = gk->nx;
  g.ny = gk->ny;
  g.nz = gk->nz;
  g.inv_h = gk->inv_h;
  g.pad = gk->pad;
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    g.org[c] = gk->org[c];
    g.lo[c] = gk->lo[c];
    g.hi[c] = gk->hi[c];
  }
  return g;
}

// distance of geometry gk at (x, y, z) (cut: the scene's best so far, for the mesh's box skip)
__device__ __noinline__ float geom_sdf_at(SceneGeomK* gk, float x, float y, float z, float cut) {
  const SceneGeom g = load_geom(gk);
  const float p[3] = {x, y, z};
  if (g.kind == 0) return volume_sdf(g, p, nullptr);
  int face;
  return mesh_sdf(g, p, cut, face);
}

// normal of geometry gk at (x, y, z): a volume's normalised gradient, (0, 0, 1) where it vanishes
// (sdf_kernels.py:540-545); for a mesh, the outward direction from the nearest surface point, which
// is the nearest triangle's face normal (the reference's, sdf_kernels.py:206-213) wherever that
// point lies inside a face, and on edges and vertices -- where the reference takes the face normal
// of whichever adjacent triangle its search meets first -- the distance gradient
__device__ __noinline__ float3 geom_normal_at(SceneGeomK* gk, float x, float y, float z) {
  const SceneGeom g = load_geom(gk);
  const float p[3] = {x, y, z};
  float3 n = make_float3(1.f, 0.f, 0.f);
  if (g.kind == 0) {
    float gr[3];
    volume_sdf(g, p, gr);
    const float len = sqrtf(gr[0] * gr[0] + gr[1] * gr[1] + gr[2] * gr[2]);
    n = len > 1e-8f ? make_float3(gr[0] / len, gr[1] / len, gr[2] / len) : make_float3(0.f, 0.f, 1.f);
  } else {
    int face;
    float dir[3];
    mesh_sdf(g, p, 3e38f, face, dir);
    if (face >= 0) n = make_float3(dir[0], dir[1], dir[2]);
  }
  return n;
}

template <class SC>
__device__ __forceinline__ SceneGeomK* geom_ptr(SC tt, int k) {
  return (SceneGeomK*)&tt->geom[k];
}

// also returns in grp the first group holding the minimum (sphere groups 0..7, box groups 8..11,
// geometry k: 16 + k; -1: none closer than 1e6), for scene_winner
template <bool GEO = true, class SC>
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
  if constexpr (GEO) {
    for (int k = 0; k < tt->ng; ++k) {
      const float d = geom_sdf_at(geom_ptr(tt, k), p[0], p[1], p[2], best);
      grp = d < best ? RK_SDF_GEOM_GRP + k : grp;
      best = fminf(best, d);
    }
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
  if (grp >= RK_SDF_GEOM_GRP) return RK_SDF_GEOM_WIN - (grp - RK_SDF_GEOM_GRP);
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

template <bool GEO = true, class SC>
__device__ __forceinline__ void scene_normal(SC tt, const float* p, int win, float* n) {
  n[0] = 1.f;
  n[1] = n[2] = 0.f;
  if (win == RK_SDF_NONE) return;  // the reference's initial direction
  if (GEO && win <= RK_SDF_GEOM_WIN) {
    const float3 g = geom_normal_at(geom_ptr(tt, RK_SDF_GEOM_WIN - win), p[0], p[1], p[2]);
    n[0] = g.x;
    n[1] = g.y;
    n[2] = g.z;
    return;
  }
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


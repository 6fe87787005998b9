// rk_traj.h -- collision-aware trajectory LM, two wavefronts per problem (included by rk_lm.hip
// inside namespace rk, after rk_thread.h).
//
// A problem is a trajectory of T <= 64 frames of a serial arm with NA dofs: D = T*NA variables,
// frame-major (index t*NA + d).  Lane t of a wave owns frame t: its FK, its collision spheres
// against the primitive scene, the goal pose on the last frame, its finite-difference rows, and
// the NA x NA diagonal block of J^T J and the frame's part of J^T r, all in registers.  The two
// wavefronts of a problem split the robot spheres; the first adds the second's rows from LDS and
// solves.  J^T J is banded (half-bandwidth W = 2*NA: acceleration rows couple frames t and t+2):
// the band is built in LDS ([D][W+1], rows padded to 16 floats) and factorised by a
// right-looking band Cholesky whose W+1 active rows live in registers of the top W+1 lanes,
// shifted one lane per column; the forward substitution is fused into it and the backward one
// runs the same way.  Rows and term costs in this order: goal, velocity, acceleration,
// collision, then the bound rows of orders 0, 1, 2 (the reference sums term costs in list order).
// The bound rows (TrajectoryLimitTask, forward stencil) and scene volumes / meshes are in a
// separate kernel instantiation (LIM), so that trajectories without them keep their register
// allocation; with no bound rows and no geometries it computes the same bits as the other.

#define RK_TRAJ_MAXT 64
#define RK_TRAJ_MAXS 64  // robot collision spheres

struct TrajTerms {
  int robot, T;
  int goal_link;
  float goal_w[6];
  float wv, wa;  // finite-difference row weights already divided by dt, dt^2 (0: term absent)
  float coll_w, margin;
  int nsph;      // robot spheres, sorted by the joint that moves their link (-1: base link first)
  int sph_joint[RK_TRAJ_MAXS];
  float sph_c[RK_TRAJ_MAXS][3];
  float sph_r[RK_TRAJ_MAXS];
  int ns, nb;
  float obs_s[RK_SCENE_MAXS][4];
  float obs_b[RK_SCENE_MAXB][6];
  int ng;  // the scene's volumes / meshes, after the boxes
  SceneGeom geom[RK_SCENE_MAXG];
  // TrajectoryLimitTask rows, forward stencil (terms/sparse/trajectory_limit_task.py:436-498):
  // for order o (bit o of lim_orders) and window f = 0..T-1-o, per dof, v = lim_scale[o] *
  // sum_k c_k q[f + k] with c = (1), (-1, 1), (1, -2, 1), r = w max(lo - v, v - hi, 0)
  int lim_orders;
  float lim_scale[3];  // 1 / dt^o
  float lim_w[3][8], lim_lo[3][8], lim_hi[3][8];
  // > 0: each LM step solves by this many conjugate-gradient iterations instead of the band
  // Cholesky (LMOptimizerConfig(linear_solver="cg"), opt/lm_optimizer.py:781-832)
  int cg_iters;
};
typedef const __attribute__((address_space(4))) TrajTerms TrajK;

// per-frame evaluation: FK streamed joint by joint with the spheres of each link processed as
// its pose appears; JAC also accumulates the frame's block of J^T J (lower, hb[a][b], b <= a)
// and J^T r (gf) of the goal and collision rows.  Returns the half sums of squares of the goal
// rows (lane T-1 only) in cg and of the collision rows in cc.  The two wavefronts of a problem
// split the robot spheres: PART 0 takes the even-numbered ones and the goal rows, PART 1 the
// odd-numbered ones.
template <int NA, int NJ, bool JAC, int PART, bool GEO>
__device__ __forceinline__ void traj_frame(RobotK* rb, TrajK* tt, const float (&q)[NA], bool last, const Pose& goal,
                                           float (&hb)[NA][NA], float (&gf)[NA], float& cg, float& cc) {
  const int nj = rb->nj;
  float S6[NA][6];
  Pose T = identity_pose(), Tg = identity_pose();
  const int gj = rb->link_joint[tt->goal_link];
  const float w = tt->coll_w, margin = tt->margin;
  int k = 0;
  auto spheres = [&](int joint, const Pose& P, int ndof) {
    for (; k < tt->nsph && tt->sph_joint[k] == joint; ++k) {
      if ((k & 1) != PART) continue;
      const float lc[3] = {tt->sph_c[k][0], tt->sph_c[k][1], tt->sph_c[k][2]};
      float c[3];
      quat_apply(P.q, lc, c);
      c[0] += P.t[0];
      c[1] += P.t[1];
      c[2] += P.t[2];
      int grp;
      const float sd = scene_sdf<GEO>(tt, c, grp) - tt->sph_r[k];
      if (sd < margin) {
        const float res = w * (margin - sd);
        cc += res * res;
        if (JAC) {
          float n[3];
          scene_normal<GEO>(tt, c, scene_winner(tt, c, grp), n);
          const float cn0 = c[1] * n[2] - c[2] * n[1], cn1 = c[2] * n[0] - c[0] * n[2], cn2 = c[0] * n[1] - c[1] * n[0];
          float jr[NA];
#pragma unroll
          for (int a = 0; a < NA; ++a)
            jr[a] = a < ndof ? -w * ((n[0] * S6[a][0] + n[1] * S6[a][1] + n[2] * S6[a][2]) +
                                     (cn0 * S6[a][3] + cn1 * S6[a][4] + cn2 * S6[a][5]))
                             : 0.f;
#pragma unroll
          for (int a = 0; a < NA; ++a) {
#pragma unroll
            for (int b2 = 0; b2 <= a; ++b2) hb[a][b2] += jr[a] * jr[b2];
            gf[a] += jr[a] * res;
          }
        }
      }
    }
  };
  spheres(-1, T, 0);
#pragma unroll
  for (int j = 0; j < NJ; ++j) {
    if (j < nj) {
      Pose L;
#pragma unroll
      for (int c = 0; c < 3; ++c) L.t[c] = rb->pT[j][c];
#pragma unroll
      for (int c = 0; c < 4; ++c) L.q[c] = rb->pT[j][3 + c];
      if (j < NA) {
        Pose X = identity_pose();
        const float ax0 = rb->axis[j][0], ax1 = rb->axis[j][1], ax2 = rb->axis[j][2];
        if (rb->joint_type[j] == 2) {
          float s, c;
          sincosf(q[j] * 0.5f, &s, &c);
          X.q[0] = c; X.q[1] = ax0 * s; X.q[2] = ax1 * s; X.q[3] = ax2 * s;
        } else {
          X.t[0] = ax0 * q[j]; X.t[1] = ax1 * q[j]; X.t[2] = ax2 * q[j];
        }
        L = compose(L, X);
      }
      T = j == 0 ? L : compose(T, L);
      if (JAC && j < NA) subspace_column(rb, j, T, S6[j]);
      if (j == gj) Tg = T;
      spheres(j, T, j < NA ? j + 1 : NA);
    }
  }
  if (PART != 0) return;
  // goal: the FrameTask log rows of the last frame (every lane evaluates; lane T-1 keeps them)
  float gw[6];
#pragma unroll
  for (int i = 0; i < 6; ++i) gw[i] = tt->goal_w[i];
  float r6[6];
  pose_target<false>(Tg, goal, gw, r6);
  float sg = 0.f;
#pragma unroll
  for (int i = 0; i < 6; ++i) sg += r6[i] * r6[i];
  cg = last ? sg : 0.f;
  if (JAC && last) {
    float tg[36];
    pose_blocks<false>(Tg, goal, gw, tg);
    float J6[6][NA];
#pragma unroll
    for (int a = 0; a < NA; ++a) {
      float c6[6] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
      if (a <= gj) pose_jac_column(tg, S6[a], c6);
#pragma unroll
      for (int i = 0; i < 6; ++i) J6[i][a] = c6[i];
    }
#pragma unroll
    for (int a = 0; a < NA; ++a) {
#pragma unroll
      for (int b2 = 0; b2 <= a; ++b2) {
        float sp = 0.f;
#pragma unroll
        for (int i = 0; i < 6; ++i) sp += J6[i][a] * J6[i][b2];
        hb[a][b2] += sp;
      }
      float sp = 0.f;
#pragma unroll
      for (int i = 0; i < 6; ++i) sp += J6[i][a] * r6[i];
      gf[a] += sp;
    }
  }
}

template <int NA>
struct TrajLDS {
  static constexpr int W = 2 * NA;          // half-bandwidth
  static constexpr int BS = (W + 4) & ~3;   // band row stride (>= W + 1, 16-byte rows)
  // band[r][k] = A[r][r-k]; after the solve L[r][r-k], 1/L[r][r] at k = 0.  Rows D..D+W+1 stay
  // zero (they enter the register window past the end); sink takes the stores of idle lanes.
  float band[RK_TRAJ_MAXT * NA + W + 2][BS];
  float rhs[RK_TRAJ_MAXT * NA + W + 2];
  float sink[64];
  // from the second wavefront: its frame block of J^T J (lower), J^T r and collision cost at the
  // iterate, and its collision cost at the proposal; to it: the iterate after each step
  static constexpr int NPART = NA * (NA + 1) / 2 + NA + 1;
  float part[NPART][64];
  float pcc[64];
  float qx[NA][64];
  int moved;  // the last step was accepted (the iterate in qx is new)
  // the second wavefront's Schur update of the separator rows (upper triangle) and right-hand side
  float sx[W][W];
  float rx[W];
};

// one-lane wavefront shifts (DPP wave_shl1 / wave_shr1): lane l takes lane l+1 (l-1); the lane
// with no source (63 / 0) takes `fill`
__device__ __forceinline__ float wv_shl1(float src, float fill) {
  return __int_as_float(__builtin_amdgcn_update_dpp(__float_as_int(fill), __float_as_int(src), 0x130, 0xF, 0xF, false));
}
__device__ __forceinline__ float wv_shr1(float src, float fill) {
  return __int_as_float(__builtin_amdgcn_update_dpp(__float_as_int(fill), __float_as_int(src), 0x138, 0xF, 0xF, false));
}

// The band's rank-1 update X[c] += (L of lane WL+c) * nl, c = 1..W, with the window on lanes
// WL..63 = lanes WL-48..15 of the wavefront's last 16-lane row: each broadcast is a DPP
// row_newbcast operand of one v_fmac_f32 (lane WL+c of the row to the whole row), so an update
// is W instructions with no scalar round trip (a v_readlane and an FMA each, otherwise).  The
// other rows broadcast their own lanes, outside the window.  DPP reads of a VGPR written by the
// VALU need two wait states, and the compiler does not see the writes inside an asm block: the
// leading s_nop covers l, written just before, the trailing one the shifts of X that follow.
// fma(L_c, -l, x) rounds as fma(-l, L_c, x).
#define RK_BC(i, n) "v_fmac_f32_dpp %" #i ", %14, %15 row_newbcast:" #n " row_mask:0xf bank_mask:0xf\n\t"
#define RK_BC12(i, n) "v_fmac_f32_dpp %" #i ", %12, %13 row_newbcast:" #n " row_mask:0xf bank_mask:0xf\n\t"
template <int W>
__device__ __forceinline__ void band_rank1(float (&X)[W + 1], float l, float nl) {
  static_assert(W == 14 || W == 12, "window within one 16-lane row: 6- or 7-dof frames");
  if constexpr (W == 14) {  // WL = 49: lanes 50..63 are row lanes 2..15
    asm volatile("s_nop 1\n\t" RK_BC(0, 2) RK_BC(1, 3) RK_BC(2, 4) RK_BC(3, 5) RK_BC(4, 6) RK_BC(5, 7)
                 RK_BC(6, 8) RK_BC(7, 9) RK_BC(8, 10) RK_BC(9, 11) RK_BC(10, 12) RK_BC(11, 13) RK_BC(12, 14)
                 RK_BC(13, 15) "s_nop 1"
                 : "+v"(X[1]), "+v"(X[2]), "+v"(X[3]), "+v"(X[4]), "+v"(X[5]), "+v"(X[6]), "+v"(X[7]),
                   "+v"(X[8]), "+v"(X[9]), "+v"(X[10]), "+v"(X[11]), "+v"(X[12]), "+v"(X[13]), "+v"(X[14])
                 : "v"(l), "v"(nl));
  } else {  // WL = 51: lanes 52..63 are row lanes 4..15
    asm volatile("s_nop 1\n\t" RK_BC12(0, 4) RK_BC12(1, 5) RK_BC12(2, 6) RK_BC12(3, 7) RK_BC12(4, 8)
                 RK_BC12(5, 9) RK_BC12(6, 10) RK_BC12(7, 11) RK_BC12(8, 12) RK_BC12(9, 13) RK_BC12(10, 14)
                 RK_BC12(11, 15) "s_nop 1"
                 : "+v"(X[1]), "+v"(X[2]), "+v"(X[3]), "+v"(X[4]), "+v"(X[5]), "+v"(X[6]), "+v"(X[7]),
                   "+v"(X[8]), "+v"(X[9]), "+v"(X[10]), "+v"(X[11]), "+v"(X[12])
                 : "v"(l), "v"(nl));
  }
}
#undef RK_BC
#undef RK_BC12

// ---- the damped band solve (J^T J + lam I) dq = -g, both wavefronts taking part.
//
// Register window (one wavefront): the active rows j..j+W sit on lanes WL..63 (row j+a on lane
// WL+a) in column-relative form C[c] = A[j+a][j+c], so column j is C[0] on every lane.  Per
// column: the pivot from lane WL, multipliers L = C[0] / d, the rank-1 update C[c] -= L * (L of
// lane WL+c), the forward substitution on rc, then a one-lane DPP shift moves the window down and
// row j+W+1 enters at lane 63 as the shift's fill value.  L goes back to the band in LDS, 1/d to
// its diagonal slot, y to rhs.  REV runs the same on the reversed band: index i is row D-1-i, and
// entry (r, c), c <= r, of the reversed matrix sits at band[D-1-c][r-c].  ZERO: the entering
// rows are zero (nothing follows the columns factorised).  Factorises columns [j0, j0 + n) from
// the window in (C, rc) and leaves the next window there; L is stored for rows below lim.
template <int NA, bool REV, bool ZERO, bool BC>
__device__ __forceinline__ void band_factor(TrajLDS<NA>* sh, int D, int j0, int n, int lim, float (&C)[2 * NA + 1],
                                            float& rc) {
  constexpr int W = 2 * NA, WL = 63 - W, BS = TrajLDS<NA>::BS;
  const int lane = threadIdx.x & 63, wa = lane - WL;
  auto at = [&](int r, int c) -> float* { return REV ? &sh->band[D - 1 - c][r - c] : &sh->band[r][r - c]; };
  auto orig = [&](int i) { return REV ? D - 1 - i : i; };
  float piv = wv_readlane(C[0], WL);
  // one column: X is the window, Y the row entering it at this step (j + W + 1), Z receives the
  // row of the next step.  The shifted window is written over Y (the DPP fill operand), so three
  // register sets rotate through the steps without copies.
  auto step = [&](int j, float (&X)[W + 1], float (&Y)[W + 1], float (&Z)[W + 1], float& xr, float& yr,
                  float& zr) {
    if (ZERO) {
#pragma unroll
      for (int c = 0; c <= W; ++c) Z[c] = 0.f;
      zr = 0.f;
    } else {
      // entry (j+W+2, j+2+c): one base address and constant offsets (a diagonal walk when REV);
      // the base goes through a vector register so that the offsets fold into the LDS reads
      const int rn = j + W + 2;
      int bo = (REV ? D - 3 - j - W : rn) * BS;
      asm("" : "+v"(bo));
      const float* const b = &sh->band[0][0] + bo;
#pragma unroll
      for (int c = 0; c <= W; ++c) Z[c] = b[REV ? (W - c) * (BS + 1) : W - c];
      zr = sh->rhs[orig(rn)];
    }
    const float inv = __builtin_amdgcn_rsqf(piv);  // 1 / L[j][j]
    const float Lv = X[0] * inv;
    const float y = wv_readlane(xr, WL) * inv;
    float* const dl = wa >= 1 && j + wa < lim ? at(j + wa, j) : &sh->sink[lane];
    *dl = Lv;
    float* const di = wa == 0 ? at(j, j) : &sh->sink[lane];
    float* const dy = wa == 0 ? &sh->rhs[orig(j)] : &sh->sink[lane];
    *di = inv;
    *dy = y;
    xr = fmaf(-Lv, y, xr);
    if constexpr (BC) {
      band_rank1<W>(X, Lv, -Lv);
    } else {
#pragma unroll
      for (int c = 1; c <= W; ++c) X[c] = fmaf(-Lv, wv_readlane(Lv, WL + c), X[c]);
    }
    piv = wv_readlane(X[1], WL + 1);  // the next pivot (lane WL+1 moves to lane WL)
#pragma unroll
    for (int c = 0; c < W; ++c) Y[c] = wv_shl1(X[c + 1], Y[c]);
    // (Y[W] keeps the entering row's diagonal; lanes below 63 hold the upper triangle there)
    yr = wv_shl1(xr, yr);
  };
  float P1[W + 1], P2[W + 1], r1, r2;
  if (ZERO) {
#pragma unroll
    for (int c = 0; c <= W; ++c) P1[c] = 0.f;
    r1 = 0.f;
  } else {
    const int rn = j0 + W + 1;
    const float* const b = REV ? &sh->band[D - 2 - j0 - W][0] : &sh->band[rn][0];
#pragma unroll
    for (int c = 0; c <= W; ++c) P1[c] = b[REV ? (W - c) * (BS + 1) : W - c];
    r1 = sh->rhs[orig(rn)];
  }
  const int je = j0 + n;
  int j = j0;
  for (; j + 2 < je; j += 3) {
    step(j, C, P1, P2, rc, r1, r2);
    step(j + 1, P1, P2, C, r1, r2, rc);
    step(j + 2, P2, C, P1, r2, rc, r1);
  }
  if (j < je) {
    step(j, C, P1, P2, rc, r1, r2);
    if (j + 1 < je) {
      step(j + 1, P1, P2, C, r1, r2, rc);
#pragma unroll
      for (int c = 0; c <= W; ++c) C[c] = P2[c];
      rc = r2;
    } else {
#pragma unroll
      for (int c = 0; c <= W; ++c) C[c] = P1[c];
      rc = r1;
    }
  }
}

// the first window of a factorisation from column 0: rows 0..W of the (reversed) band
template <int NA, bool REV>
__device__ __forceinline__ void band_window0(TrajLDS<NA>* sh, int D, float (&C)[2 * NA + 1], float& rc) {
  constexpr int W = 2 * NA, WL = 63 - W;
  const int lane = threadIdx.x & 63, wa = lane - WL;
#pragma unroll
  for (int c = 0; c <= W; ++c) {
    const bool ok = wa >= c;
    const int r = ok ? wa : c;
    C[c] = REV ? sh->band[D - 1 - c][r - c] : sh->band[r][r - c];
    C[c] = ok ? C[c] : 0.f;
  }
  const int r = wa >= 0 ? wa : 0;
  rc = sh->rhs[REV ? D - 1 - r : r];
  rc = wa >= 0 ? rc : 0.f;
}

// L^T x = y for pivots jtop down to jstop, the same way: rows j-W..j on lanes 0..W (row j-W+l on
// lane l), the pivot on lane W; the window (rb) moves one lane up and row j-W-1 enters at lane 0.
// Lane l applies L[j][j-(W-l)]; entries left of column 0 are zero, and the lanes outside the
// window (and the pivot lane, which leaves it) read a harmless in-range entry.  The values of the
// next TB pivots are loaded while the current TB are solved.
template <int NA, bool REV>
__device__ __forceinline__ void band_back(TrajLDS<NA>* sh, int D, int jtop, int jstop, float& rb) {
  constexpr int W = 2 * NA, TB = 8, BS = TrajLDS<NA>::BS;
  const int lane = threadIdx.x & 63;
  const int wb = W - lane;
  const int wbc = wb < 0 ? 0 : wb;
  auto orig = [&](int i) { return REV ? D - 1 - i : i; };
  float* const dx = lane == W ? &sh->rhs[0] : &sh->sink[lane];
  const int dxs = lane == W ? 1 : 0;
  auto step = [&](int j, float Lj, float ij, float pj) {
    const float x = wv_readlane(rb, W) * ij;
    dx[dxs * orig(j)] = x;
    rb = fmaf(-Lj, x, rb);
    rb = wv_shr1(rb, pj);
  };
  float Lc[TB], ic[TB], pc[TB];
  const float* const lb = &sh->band[0][0] + (REV ? wbc * (BS + 1) : wbc);  // this lane's L entry of row 0
  auto load = [&](int jb, float (&Lo)[TB], float (&io)[TB], float (&po)[TB]) {
#pragma unroll
    for (int u = 0; u < TB; ++u) {
      const int j = max(jb - u, jstop);
      Lo[u] = lb[orig(j) * BS];
      io[u] = sh->band[orig(j)][0];
      const int e = jb - u - W - 1;  // (the rows before 0 never become pivots)
      po[u] = sh->rhs[REV ? D - 1 - max(e, -W - 2) : max(e, 0)];
    }
  };
  int jb = jtop;
  load(jb, Lc, ic, pc);
  for (; jb - (TB - 1) >= jstop; jb -= TB) {
    float Ln[TB], in[TB], pn[TB];
    load(jb - TB, Ln, in, pn);
#pragma unroll
    for (int u = 0; u < TB; ++u) step(jb - u, Lc[u], ic[u], pc[u]);
#pragma unroll
    for (int u = 0; u < TB; ++u) {
      Lc[u] = Ln[u];
      ic[u] = in[u];
      pc[u] = pn[u];
    }
  }
#pragma unroll
  for (int u = 0; u < TB - 1; ++u)
    if (jb - u >= jstop) step(jb - u, Lc[u], ic[u], pc[u]);
}

// the first backward window: rows jtop-W..jtop
template <int NA, bool REV>
__device__ __forceinline__ float band_back_window(TrajLDS<NA>* sh, int D, int jtop) {
  constexpr int W = 2 * NA;
  const int lane = threadIdx.x & 63;
  const int r = jtop - (W - lane);
  const bool ok = lane <= W && r >= 0;
  const float v = sh->rhs[ok ? (REV ? D - 1 - r : r) : 0];
  return ok ? v : 0.f;
}

// Nested dissection over the two wavefronts.  The frames split into A = [0, m), a separator
// S = {m, m+1} (W = 2*NA rows) and B = [m+2, T); no row of A couples to a row of B (the band
// spans two frames).  The first wavefront (wv 0) eliminates A in order while the second
// eliminates B from the end (the reversed band), each leaving its update of S in its window; the
// first then factorises S0 + S1 - A_SS, solves S backwards and publishes x_S; then each runs the
// backward substitution of its part.  Short trajectories (T < 6) are solved whole by the first.
// Returns with dq in sh->rhs; takes three barriers in both wavefronts.
template <int NA, bool BC>
__device__ __forceinline__ void band_solve(TrajLDS<NA>* sh, int T, int wv) {
  constexpr int W = 2 * NA, WL = 63 - W;
  const int lane = threadIdx.x & 63, wa = lane - WL;
  const int D = T * NA;
  const bool split = T >= 6;
  const int m = (T - 2) / 2;
  const int nA = m * NA, nB = (T - m - 2) * NA;
  float C[W + 1], rc;
  if (wv == 0) {
    band_window0<NA, false>(sh, D, C, rc);
    if (split)
      band_factor<NA, false, false, BC>(sh, D, 0, nA, nA + W, C, rc);
    else
      band_factor<NA, false, false, BC>(sh, D, 0, D, D, C, rc);
    __syncthreads();  // the second wavefront's update of S is in sx, rx
    float rb = 0.f;
    if (split) {
      if (wa >= 0 && wa < W) {
#pragma unroll
        for (int c = 0; c < W; ++c)
          if (c <= wa) C[c] += sh->sx[c][wa] - sh->band[nA + wa][wa - c];
        rc += sh->rx[wa] - sh->rhs[nA + wa];
      }
      if (wa == W) {
#pragma unroll
        for (int c = 0; c <= W; ++c) C[c] = 0.f;
        rc = 0.f;
      }
      band_factor<NA, false, true, BC>(sh, D, nA, W, nA + W, C, rc);
      rb = band_back_window<NA, false>(sh, D, nA + W - 1);
      band_back<NA, false>(sh, D, nA + W - 1, nA, rb);
    }
    __syncthreads();  // x_S published
    if (split)
      band_back<NA, false>(sh, D, nA - 1, 0, rb);
    else {
      rb = band_back_window<NA, false>(sh, D, D - 1);
      band_back<NA, false>(sh, D, D - 1, 0, rb);
    }
  } else {
    if (split) {
      band_window0<NA, true>(sh, D, C, rc);
      band_factor<NA, true, false, BC>(sh, D, 0, nB, D, C, rc);
      // the window now holds the reversed separator rows: lane WL+a is row W-1-a of S
      if (wa >= 0 && wa < W) {
#pragma unroll
        for (int c = 0; c < W; ++c)
          if (c <= wa) sh->sx[W - 1 - wa][W - 1 - c] = C[c];
        sh->rx[W - 1 - wa] = rc;
      }
    }
    __syncthreads();  // update of S handed over
    __syncthreads();  // x_S published
    if (split) {
      // the separator pivots (reversed nB+W-1 .. nB) with x known: their L^T rows update B only
      float rb = band_back_window<NA, true>(sh, D, nB + W - 1);
      const int wb = W - lane, wbc = wb < 0 ? 0 : wb;
      float xs[W], Ls[W], ps[W];
#pragma unroll
      for (int u = 0; u < W; ++u) {
        const int j = nB + W - 1 - u;
        xs[u] = sh->rhs[D - 1 - j];
        Ls[u] = sh->band[D - 1 - j + wbc][wbc];
        ps[u] = sh->rhs[D - 1 - (j - W - 1)];
      }
#pragma unroll
      for (int u = 0; u < W; ++u) {
        rb = fmaf(-Ls[u], xs[u], rb);
        rb = wv_shr1(rb, ps[u]);
      }
      band_back<NA, true>(sh, D, nB - 1, 0, rb);
    }
  }
  __syncthreads();  // dq complete
}

// finite-difference rows of lane t (velocity row t, acceleration window starting at t) and their
// J^T r contribution to frame t; q of frames t-2..t+2 come from the neighbouring lanes
template <int NA, bool JAC>
__device__ __forceinline__ void traj_smooth(TrajK* tt, int t, int T, const float (&q)[NA], float (&gf)[NA], float& cv,
                                            float& ca) {
  const float wv = tt->wv, wa = tt->wa;
  const int l1 = t + 1 < 64 ? t + 1 : 63, l2 = t + 2 < 64 ? t + 2 : 63;
  const bool has_v = t + 1 < T, has_a = t + 2 < T;
  float v[NA], ac[NA];
#pragma unroll
  for (int d = 0; d < NA; ++d) {
    const float q1 = __shfl(q[d], l1), q2 = __shfl(q[d], l2);
    v[d] = has_v ? wv * (q1 - q[d]) : 0.f;
    ac[d] = has_a ? wa * (q[d] + -2.f * q1 + q2) : 0.f;
    cv += v[d] * v[d];
    ca += ac[d] * ac[d];
  }
  if (JAC) {
    const int m1 = t >= 1 ? t - 1 : 0, m2 = t >= 2 ? t - 2 : 0;
#pragma unroll
    for (int d = 0; d < NA; ++d) {
      // every lane takes part in each shuffle (a lane reading an inactive lane gets nothing)
      const float vs1 = __shfl(v[d], m1), as1 = __shfl(ac[d], m1), as2 = __shfl(ac[d], m2);
      const float vm1 = t >= 1 ? vs1 : 0.f;
      const float am1 = t >= 1 ? as1 : 0.f;
      const float am2 = t >= 2 ? as2 : 0.f;
      gf[d] += (-wv * v[d] + wv * vm1) + (wa * ac[d] + -2.f * wa * am1 + wa * am2);
    }
  }
}

// bound rows of lane t: its own window f = t of each present order (q of frames t+1, t+2 from the
// neighbouring lanes), the sums of squares cl[o]; with JAC the frame's J^T r and band entries of
// every window that reaches frame t (windows t, t-1, t-2, read from the lanes that own them): the
// diagonal hl0, the coupling to frame t-1 (hl1) and to frame t-2 (hl2), per dof
template <int NA, bool JAC>
__device__ __forceinline__ void traj_limits(TrajK* tt, int t, int T, const float (&q)[NA], float (&gf)[NA],
                                            float (&hl0)[NA], float (&hl1)[NA], float (&hl2)[NA], float (&cl)[3]) {
  const int l1 = t + 1 < 64 ? t + 1 : 63, l2 = t + 2 < 64 ? t + 2 : 63;
  const int m1 = t >= 1 ? t - 1 : 0, m2 = t >= 2 ? t - 2 : 0;
  const int orders = tt->lim_orders;
#pragma unroll
  for (int d = 0; d < NA; ++d) {
    const float q1 = __shfl(q[d], l1), q2 = __shfl(q[d], l2);
#pragma unroll
    for (int o = 0; o < 3; ++o) {
      if (!((orders >> o) & 1)) continue;  // wave-uniform
      const float sc = tt->lim_scale[o], w = tt->lim_w[o][d];
      const bool has = t + o < T;
      float v = o == 0 ? q[d] : o == 1 ? -q[d] + q1 : (q[d] + -2.f * q1) + q2;
      v = v * sc;
      const float below = tt->lim_lo[o][d] - v, above = v - tt->lim_hi[o][d];
      const bool act = has && (below > 0.f || above > 0.f);
      const float rr = act ? w * fmaxf(fmaxf(below, above), 0.f) : 0.f;
      cl[o] += rr * rr;
      if (JAC) {
        const float ws = act ? (above > 0.f ? w : -w) * sc : 0.f;  // w * sign * scale
        // the window's Jacobian entries a_k = ws * c_k
        const float a0 = o == 0 ? ws : o == 1 ? -ws : ws, a1 = o == 1 ? ws : -2.f * ws, a2 = ws;
        gf[d] += a0 * rr;
        hl0[d] += a0 * a0;
        if (o >= 1) {  // window t-1: frame t is its entry 1
          const float pa0 = __shfl(a0, m1), pa1 = __shfl(a1, m1), pr = __shfl(rr, m1);
          if (t >= 1) {
            gf[d] += pa1 * pr;
            hl0[d] += pa1 * pa1;
            hl1[d] += pa1 * pa0;
          }
        }
        if (o == 2) {  // window t-2: frame t is its entry 2
          const float qa0 = __shfl(a0, m2), qa1 = __shfl(a1, m2), qa2 = __shfl(a2, m2), qr = __shfl(rr, m2);
          if (t >= 2) {
            gf[d] += qa2 * qr;
            hl0[d] += qa2 * qa2;
            hl1[d] += qa2 * qa1;
            hl2[d] += qa2 * qa0;
          }
        }
      }
    }
  }
}

// Conjugate gradients on (J^T J + lam I) dq = -g with a fixed budget of `iters` iterations, the
// reference's CG step (_compute_sparse_delta, opt/lm_optimizer.py:781-832; its kernels :370-499;
// eps 1e-8 in both ratios) run without its early exit, as the reference does under graph capture
// or with early_stopping_interval = 0.  The matrix is block pentadiagonal over frames: lane t
// holds its frame's diagonal block (hb lower triangle + dg on the diagonal) and the diagonal
// coupling blocks o1 (to frame t-1) and o2 (to frame t-2); the couplings of frames t+1 and t+2
// to t are those lanes' o1 / o2, and every neighbour vector comes by a one-lane DPP shift, so the
// product never leaves registers.  Lanes past the last frame hold zeros.  Returns dq in x and
// the reference's predicted reduction -g.dq - dq^T (J^T J) dq / 2.
template <int NA>
__device__ __forceinline__ float traj_cg(const float (&hb)[NA][NA], const float (&dg)[NA], const float (&o1)[NA],
                                         const float (&o2)[NA], const float (&gf)[NA], float lam, bool own, int iters,
                                         float (&x)[NA]) {
  float o1n[NA], o2n[NA];
#pragma unroll
  for (int i = 0; i < NA; ++i) {
    o1n[i] = wv_shl1(o1[i], 0.f);
    o2n[i] = wv_shl1(wv_shl1(o2[i], 0.f), 0.f);
  }
  auto apply = [&](const float (&v)[NA], float l, float (&y)[NA]) {
#pragma unroll
    for (int i = 0; i < NA; ++i) {
      const float m1 = wv_shr1(v[i], 0.f), p1 = wv_shl1(v[i], 0.f);
      const float m2 = wv_shr1(m1, 0.f), p2 = wv_shl1(p1, 0.f);
      float acc = (dg[i] + l) * v[i];
#pragma unroll
      for (int j = 0; j < NA; ++j)
        if (j != i) acc = fmaf(j < i ? hb[i][j] : hb[j][i], v[j], acc);
      acc = fmaf(o1[i], m1, acc);
      acc = fmaf(o2[i], m2, acc);
      acc = fmaf(o1n[i], p1, acc);
      acc = fmaf(o2n[i], p2, acc);
      y[i] = own ? acc : 0.f;
    }
  };
  float r[NA], p[NA], Ap[NA];
  float rr = 0.f;
#pragma unroll
  for (int i = 0; i < NA; ++i) {
    x[i] = 0.f;
    r[i] = own ? -gf[i] : 0.f;
    p[i] = r[i];
    rr = fmaf(r[i], r[i], rr);
  }
  rr = wv_sum(rr);
  for (int k = 0; k < iters; ++k) {
    apply(p, lam, Ap);
    float pap = 0.f;
#pragma unroll
    for (int i = 0; i < NA; ++i) pap = fmaf(p[i], Ap[i], pap);
    const float alpha = rr / (wv_sum(pap) + 1e-8f);
    float nr = 0.f;
#pragma unroll
    for (int i = 0; i < NA; ++i) {
      x[i] = fmaf(alpha, p[i], x[i]);
      r[i] = fmaf(-alpha, Ap[i], r[i]);
      nr = fmaf(r[i], r[i], nr);
    }
    nr = wv_sum(nr);
    const float beta = nr / (rr + 1e-8f);
    rr = nr;
#pragma unroll
    for (int i = 0; i < NA; ++i) p[i] = fmaf(beta, p[i], r[i]);
  }
  apply(x, 0.f, Ap);  // J^T J dq, undamped
  float gd = 0.f, qd = 0.f;
#pragma unroll
  for (int i = 0; i < NA; ++i) {
    gd = fmaf(x[i], own ? gf[i] : 0.f, gd);
    qd = fmaf(x[i], Ap[i], qd);
  }
  return -wv_sum(gd) - 0.5f * wv_sum(qd);
}

template <int NA, int NJ, bool LIM, bool BC>
__device__ __forceinline__ void traj_solve(const Problems& Pg, TrajK* tt, int s, const LMParams& lp, int it0, int iters,
                                           TrajLDS<NA>* sh) {
  constexpr int W = TrajLDS<NA>::W;
  RobotK* rb = as_const(Pg.robot);
  const int lane = threadIdx.x & 63;
  const int T = tt->T, D = T * NA, S = Pg.stride, cgi = tt->cg_iters;
  const int t = lane < T ? lane : T - 1;
  const bool own = lane < T;  // lanes >= T mirror frame T-1 and contribute nothing
  const bool last = lane == T - 1;
  float q[NA];
#pragma unroll
  for (int d = 0; d < NA; ++d) {
    const int k = t * NA + d;
    q[d] = (it0 == 0 && Pg.q_seed) ? Pg.q_seed[k * S + s] : Pg.q[k * S + s];
  }
  Pose goal;
  {
    float v[7];
#pragma unroll
    for (int c = 0; c < 7; ++c) v[c] = Pg.pose_tgt[c * S + s];
    load_pose(v, goal);
  }
  float lam = it0 == 0 ? lp.lm_lambda : Pg.lam[s];
  int nacc = it0 == 0 ? 0 : Pg.nacc[s];
  // smoothness band of J^T J is constant: entries (t,d)-(t,d), (t+1,d)-(t,d), (t+2,d)-(t,d)
  const float wv2 = tt->wv * tt->wv, wa2 = tt->wa * tt->wa;
  auto in_a = [&](int st) { return st >= 0 && st <= T - 3; };
  auto in_v = [&](int st) { return st >= 0 && st <= T - 2; };
  const float hs0 = wv2 * (float)(in_v(t) + in_v(t - 1)) + wa2 * (1.f * in_a(t) + 4.f * in_a(t - 1) + 1.f * in_a(t - 2));
  const float hs1m = wv2 * (in_v(t - 1) ? -1.f : 0.f) + wa2 * (-2.f * in_a(t - 1) - 2.f * in_a(t - 2));  // (t, t-1)
  const float hs2m = wa2 * (float)in_a(t - 2);                                                            // (t, t-2)
  // bound-row terms after the four others, order by order
  auto limit_cost = [&](float c, const float (&cl)[3]) {
    if constexpr (LIM) {
#pragma unroll
      for (int o = 0; o < 3; ++o)
        if ((tt->lim_orders >> o) & 1) c = c + 0.5f * wv_sum(own ? cl[o] : 0.f);
    }
    return c;
  };
  auto eval_cost = [&](const float (&qq)[NA]) {
    float hb[NA][NA], gf[NA];
    float cg = 0.f, cc = 0.f, cv = 0.f, ca = 0.f;
    traj_frame<NA, NJ, false, 0, LIM>(rb, tt, qq, last, goal, hb, gf, cg, cc);
    traj_smooth<NA, false>(tt, t, T, qq, gf, cv, ca);
    float cl[3] = {0.f, 0.f, 0.f};
    if constexpr (LIM) {
      float h0[NA], h1[NA], h2[NA];
      traj_limits<NA, false>(tt, t, T, qq, gf, h0, h1, h2, cl);
    }
    __syncthreads();  // the second wavefront's collision cost (odd spheres)
    cc += sh->pcc[lane];
    if (!own) cg = cc = cv = ca = 0.f;
    const float c0 = 0.5f * wv_sum(cg), c1 = 0.5f * wv_sum(cv), c2 = 0.5f * wv_sum(ca), c3 = 0.5f * wv_sum(cc);
    return limit_cost(((c0 + c1) + c2) + c3, cl);
  };
  // rows D..D+W of the band and the right-hand side are zero: they enter the Cholesky window past the end
  for (int e = lane; e < (W + 2) * TrajLDS<NA>::BS; e += 64) sh->band[D + e / TrajLDS<NA>::BS][e % TrajLDS<NA>::BS] = 0.f;
  if (lane <= W + 1) sh->rhs[D + lane] = 0.f;
  float cost = 0.f;
  if (iters == 0) cost = eval_cost(q);
  // the frame's block of J^T J, J^T r and the cost at the iterate: a rejected step leaves the
  // iterate, hence these, unchanged, so they are rebuilt only after an accepted step (both
  // wavefronts skip the rows together: the decision is published with the iterate)
  float hb[NA][NA], gf[NA];
  float hl0[NA], hl1[NA], hl2[NA];  // bound rows' band entries (LIM)
#pragma unroll
  for (int d = 0; d < NA; ++d) hl0[d] = hl1[d] = hl2[d] = 0.f;
  bool moved = true;
  for (int it = 0; it < iters; ++it) {
    const int gi = it0 + it;
    if (Pg.q_trace && gi < Pg.trace_iters) {
      if (own)
#pragma unroll
        for (int d = 0; d < NA; ++d) Pg.q_trace[((size_t)gi * D + t * NA + d) * S + s] = q[d];
      if (lane == 0) Pg.lam_trace[gi * S + s] = lam;
    }
    // ---- rows, block of J^T J and J^T r of this frame
    if (moved) {
#pragma unroll
      for (int a = 0; a < NA; ++a) {
        gf[a] = 0.f;
#pragma unroll
        for (int b2 = 0; b2 < NA; ++b2) hb[a][b2] = 0.f;
      }
      float cg = 0.f, cc = 0.f, cv = 0.f, ca = 0.f;
      traj_frame<NA, NJ, true, 0, LIM>(rb, tt, q, last, goal, hb, gf, cg, cc);
      traj_smooth<NA, true>(tt, t, T, q, gf, cv, ca);
      float cl[3] = {0.f, 0.f, 0.f};
      if constexpr (LIM) {
#pragma unroll
        for (int d = 0; d < NA; ++d) hl0[d] = hl1[d] = hl2[d] = 0.f;
        traj_limits<NA, true>(tt, t, T, q, gf, hl0, hl1, hl2, cl);
      }
      __syncthreads();  // the second wavefront's rows (odd spheres)
      {
        int e = 0;
#pragma unroll
        for (int a = 0; a < NA; ++a)
#pragma unroll
          for (int b2 = 0; b2 <= a; ++b2) hb[a][b2] += sh->part[e++][lane];
#pragma unroll
        for (int a = 0; a < NA; ++a) gf[a] += sh->part[e++][lane];
        cc += sh->part[e][lane];
      }
      if (!own) cg = cc = cv = ca = 0.f;
      cost = limit_cost(((0.5f * wv_sum(cg) + 0.5f * wv_sum(cv)) + 0.5f * wv_sum(ca)) + 0.5f * wv_sum(cc), cl);
    }
    float dq[NA], qp[NA];
    float pred;
    if (cgi > 0) {
      // ---- conjugate gradients in registers; the step goes to the second wavefront in rhs
      float dg[NA], o1[NA], o2[NA];
#pragma unroll
      for (int i = 0; i < NA; ++i) {
        dg[i] = LIM ? hb[i][i] + hs0 + hl0[i] : hb[i][i] + hs0;
        o1[i] = own ? (LIM ? hs1m + hl1[i] : hs1m) : 0.f;
        o2[i] = own ? (LIM ? hs2m + hl2[i] : hs2m) : 0.f;
      }
      __syncthreads();  // (pairs with the second wavefront's "band built")
      pred = traj_cg<NA>(hb, dg, o1, o2, gf, lam, own, cgi, dq);
      if (own)
#pragma unroll
        for (int d = 0; d < NA; ++d) sh->rhs[t * NA + d] = dq[d];
      __syncthreads();  // (the three barriers of the band solve)
      __syncthreads();
      __syncthreads();
#pragma unroll
      for (int d = 0; d < NA; ++d) qp[d] = q[d] + dq[d];
    } else {
      // ---- band of J^T J + lam I (rows of frame t), right-hand side J^T r
      if (own) {
#pragma unroll
        for (int i = 0; i < NA; ++i) {
          const int row = t * NA + i;
#pragma unroll
          for (int kk = 0; kk <= W; ++kk) {
            float v = 0.f;
            if (kk <= i) v = hb[i][i - kk];
            if (kk == 0) v = LIM ? v + hs0 + hl0[i] + lam : v + hs0 + lam;
            if (kk == NA) v = LIM ? hs1m + hl1[i] : hs1m;
            if (kk == 2 * NA) v = LIM ? hs2m + hl2[i] : hs2m;
            sh->band[row][kk] = v;
          }
          sh->rhs[row] = -gf[i];
        }
      }
      __syncthreads();  // band built
      band_solve<NA, BC>(sh, T, 0);
      float pr = 0.f;
#pragma unroll
      for (int d = 0; d < NA; ++d) {
        dq[d] = sh->rhs[t * NA + d];
        qp[d] = q[d] + dq[d];
        pr += dq[d] * (lam * dq[d] - gf[d]);
      }
      pred = 0.5f * wv_sum(own ? pr : 0.f);
    }
    const float cp = eval_cost(qp);
    if (Pg.trace && gi == 0 && lane == 0) Pg.trace[s] = cost;
    const bool acc = lm_accept(lp, cost, cp, pred, lam);
    if (acc) {
      ++nacc;
#pragma unroll
      for (int d = 0; d < NA; ++d) q[d] = qp[d];
    }
    if (Pg.trace && gi < Pg.trace_iters && lane == 0) {
      Pg.trace[(gi + 1) * S + s] = cost;
      Pg.acc_trace[gi * S + s] = acc;
    }
    moved = acc;
    if (acc)
#pragma unroll
      for (int d = 0; d < NA; ++d) sh->qx[d][lane] = q[d];
    if (lane == 0) sh->moved = acc ? 1 : 0;
    __syncthreads();
  }
  if (own)
#pragma unroll
    for (int d = 0; d < NA; ++d) Pg.q[(t * NA + d) * S + s] = q[d];
  if (lane == 0) {
    Pg.cost[s] = cost;
    Pg.lam[s] = lam;
    Pg.nacc[s] = nacc;
  }
}

// the second wavefront of a trajectory problem: the collision rows of the odd-numbered robot
// spheres at the iterate (after an accepted step) and at each proposal, handed over in LDS, and
// its half of the band solve.  Its barriers pair one for one with traj_solve's: rows ready (after
// an accepted step only), band built, the three of band_solve, proposal cost ready, iterate
// published.
template <int NA, int NJ, bool LIM, bool BC>
__device__ __forceinline__ void traj_worker(const Problems& Pg, TrajK* tt, int s, int it0, int iters,
                                            TrajLDS<NA>* sh) {
  RobotK* rb = as_const(Pg.robot);
  const int lane = threadIdx.x & 63;
  const int T = tt->T, S = Pg.stride;
  const int t = lane < T ? lane : T - 1;
  float q[NA];
#pragma unroll
  for (int d = 0; d < NA; ++d) {
    const int k = t * NA + d;
    q[d] = (it0 == 0 && Pg.q_seed) ? Pg.q_seed[k * S + s] : Pg.q[k * S + s];
  }
  const Pose goal = identity_pose();  // (no goal rows in this part)
  auto proposal_cost = [&](const float (&qq)[NA]) {
    float hb[NA][NA], gf[NA];
    float cg = 0.f, cc = 0.f;
    traj_frame<NA, NJ, false, 1, LIM>(rb, tt, qq, false, goal, hb, gf, cg, cc);
    sh->pcc[lane] = cc;
    __syncthreads();
  };
  if (iters == 0) {
    proposal_cost(q);
    return;
  }
  bool moved = true;
  for (int it = 0; it < iters; ++it) {
    if (moved) {
      float hb[NA][NA], gf[NA];
#pragma unroll
      for (int a = 0; a < NA; ++a) {
        gf[a] = 0.f;
#pragma unroll
        for (int b2 = 0; b2 < NA; ++b2) hb[a][b2] = 0.f;
      }
      float cg = 0.f, cc = 0.f;
      traj_frame<NA, NJ, true, 1, LIM>(rb, tt, q, false, goal, hb, gf, cg, cc);
      int e = 0;
#pragma unroll
      for (int a = 0; a < NA; ++a)
#pragma unroll
        for (int b2 = 0; b2 <= a; ++b2) sh->part[e++][lane] = hb[a][b2];
#pragma unroll
      for (int a = 0; a < NA; ++a) sh->part[e++][lane] = gf[a];
      sh->part[e][lane] = cc;
      __syncthreads();  // rows ready
    }
    __syncthreads();  // band built
    if (tt->cg_iters > 0) {
      __syncthreads();  // (the first wavefront runs the conjugate gradients alone)
      __syncthreads();
      __syncthreads();
    } else {
      band_solve<NA, BC>(sh, T, 1);
    }
    float qp[NA];
#pragma unroll
    for (int d = 0; d < NA; ++d) qp[d] = q[d] + sh->rhs[t * NA + d];
    proposal_cost(qp);  // (+ barrier: proposal cost ready)
    __syncthreads();  // iterate published
    moved = sh->moved != 0;
    if (moved)
#pragma unroll
      for (int d = 0; d < NA; ++d) q[d] = sh->qx[d][lane];
  }
}

// LIM: the kernel instantiation with bound rows and scene volumes / meshes (the other one, for
// primitive scenes without bound rows, compiles neither)
// BC: the band's rank-1 updates by DPP row broadcasts (band_rank1); without it, by v_readlane
// broadcasts, whose scalar round trips leave VALU issue slots to other kernels' wavefronts on the
// same SIMDs.  The request server launches that form beside its IK launches: on the mixed served
// trace its IK p99 was 1.65-1.80 ms, against 2.16-2.19 ms beside the broadcast form, which solves
// a batch of 512 x 64 trajectories in 6.44 ms instead of 6.88.
// The broadcast form is compiled for one wavefront per SIMD (a batch of 512 trajectories is 1024
// wavefronts on 1024 SIMDs), so that register allocation and scheduling aim at the latency of one
// wavefront's chain: compiled for two per SIMD (held to 256 registers), the 7-dof kernel took
// 7.8 ms per 512 x 64 solve.  The other form keeps the default target (its code as measured).
template <int NA, int NJ, bool LIM = false, bool BC = true>
__global__ __launch_bounds__(128) __attribute__((amdgpu_waves_per_eu(1, BC ? 1 : 10))) void k_traj(
    Problems P, const TrajTerms* tt, LMParams lp, int it0, int iters) {
  __shared__ __align__(16) TrajLDS<NA> sh;
  const int s = xcd_order(blockIdx.x, gridDim.x);
  if (s >= P.n) return;
  if (threadIdx.x < 64)
    traj_solve<NA, NJ, LIM, BC>(P, as_const(tt), s, lp, it0, iters, &sh);
  else
    traj_worker<NA, NJ, LIM, BC>(P, as_const(tt), s, it0, iters, &sh);
}

// Trajectory batches of one arm class (dofs, bound rows / geometries) in a group: consecutive
// workgroup ranges serve the batches of a table, so trajectories of uneven horizons (and scenes)
// share one launch.  Workgroup b solves problem b - block0 of the bucket whose range holds it.
struct TrajBucket {
  Problems P;
  const TrajTerms* tt;
  int block0;
  int pad;
};
typedef const __attribute__((address_space(4))) TrajBucket TrajBucketK;

template <int NA, int NJ, bool LIM = false>
__global__ __launch_bounds__(128) __attribute__((amdgpu_waves_per_eu(1, 1))) void k_traj_multi(const TrajBucket* table_,
                                                                                               int nb, LMParams lp,
                                                                                               int it0, int iters) {
  __shared__ __align__(16) TrajLDS<NA> sh;
  TrajBucketK* table = (TrajBucketK*)table_;
  const int bx = xcd_order(blockIdx.x, gridDim.x);
  int k = 0;
  while (k + 1 < nb && bx >= table[k + 1].block0) ++k;  // wave-uniform
  const __attribute__((address_space(4))) Problems& Q = table[k].P;  // read field by field (scalar loads)
  Problems P;
  P.robot = Q.robot; P.terms = Q.terms; P.q = Q.q; P.q_seed = Q.q_seed; P.pose_tgt = Q.pose_tgt;
  P.pos_tgt = Q.pos_tgt; P.rot_tgt = Q.rot_tgt; P.q_prev = Q.q_prev; P.cost = Q.cost; P.lam = Q.lam;
  P.nacc = Q.nacc; P.trace = Q.trace; P.acc_trace = Q.acc_trace; P.q_trace = Q.q_trace; P.lam_trace = Q.lam_trace;
  P.trace_iters = Q.trace_iters; P.n = Q.n; P.stride = Q.stride;
  const int s = bx - table[k].block0;
  if (s >= P.n) return;
  TrajK* tt = as_const(table[k].tt);
  if (threadIdx.x < 64)
    traj_solve<NA, NJ, LIM, true>(P, tt, s, lp, it0, iters, &sh);
  else
    traj_worker<NA, NJ, LIM, true>(P, tt, s, it0, iters, &sh);
}

// goal errors at the last frame and the smallest sphere-to-scene distance (one wave per problem)
template <int NA, int NJ>
__global__ __launch_bounds__(64) void k_traj_metrics(Problems P, const TrajTerms* ttg, float* pos_err, float* ang_err,
                                                     float* min_sd) {
  const int s = blockIdx.x, lane = threadIdx.x & 63;
  if (s >= P.n) return;
  TrajK* tt = as_const(ttg);
  RobotK* rb = as_const(P.robot);
  const int T = tt->T, S = P.stride, t = lane < T ? lane : T - 1;
  float q[NA];
#pragma unroll
  for (int d = 0; d < NA; ++d) q[d] = P.q[(t * NA + d) * S + s];
  float msd = 1e6f;
  Pose Tt = identity_pose(), Tg = identity_pose();
  const int gj = rb->link_joint[tt->goal_link];
  int k = 0;
  auto spheres = [&](int joint, const Pose& Pw) {
    for (; k < tt->nsph && tt->sph_joint[k] == joint; ++k) {
      const float lc[3] = {tt->sph_c[k][0], tt->sph_c[k][1], tt->sph_c[k][2]};
      float c[3];
      quat_apply(Pw.q, lc, c);
      c[0] += Pw.t[0];
      c[1] += Pw.t[1];
      c[2] += Pw.t[2];
      int grp;
      msd = fminf(msd, scene_sdf(tt, c, grp) - tt->sph_r[k]);
    }
  };
  spheres(-1, Tt);
  for (int j = 0; j < rb->nj && j < NJ; ++j) {
    Pose L = joint_local(rb, j, q);
    Tt = j == 0 ? L : compose(Tt, L);
    if (j == gj) Tg = Tt;
    spheres(j, Tt);
  }
  for (int o = 32; o > 0; o >>= 1) msd = fminf(msd, __shfl_xor(msd, o));
  if (lane == 0) min_sd[s] = msd;
  if (lane == T - 1) {
    float v[7];
    for (int c = 0; c < 7; ++c) v[c] = P.pose_tgt[c * S + s];
    Pose goal;
    load_pose(v, goal);
    const Pose E = compose(inverse(Tg), goal);
    pos_err[s] = sqrtf(E.t[0] * E.t[0] + E.t[1] * E.t[1] + E.t[2] * E.t[2]);
    const float vv = sqrtf(E.q[1] * E.q[1] + E.q[2] * E.q[2] + E.q[3] * E.q[3]);
    ang_err[s] = 2.f * asinf(fminf(vv, 1.f));
  }
}

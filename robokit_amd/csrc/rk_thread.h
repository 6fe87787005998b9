// rk_thread.h -- thread-per-problem LM for serial arms (included by rk_lm.hip inside namespace rk).
//
// A serial arm here: after depth renumbering joint j's parent is j-1, joints 0..na-1 are actuated
// one-to-one (dof a = joint a), the remaining joints are fixed.  Then every index in the solve is
// a compile-time constant: the motion-subspace columns, the Jacobian (6 rows x NA columns),
// J^T J, its Cholesky factor and the solves all live in VGPRs of one lane, robot and term data
// are read through the constant address space (scalar loads), and there is no LDS and no
// synchronisation.  FK is streamed joint by joint (no pose array is kept): once at the iterate
// (with the subspace columns) and once at the proposal (to the target link only).
// Column a of the pose rows is structurally zero when joint a is not an ancestor of the target
// link (a > its parent joint); those products are skipped, which adds exact zeros only, so the
// per-entry operation sequence equals the dense one of the LDS kernel.

typedef const __attribute__((address_space(4))) Robot RobotK;
typedef const __attribute__((address_space(4))) Terms TermsK;

template <class T>
__device__ __forceinline__ const __attribute__((address_space(4))) T* as_const(const T* p) {
  return (const __attribute__((address_space(4))) T*)p;
}

__device__ __forceinline__ Pose identity_pose() {
  Pose P;
  P.t[0] = P.t[1] = P.t[2] = 0.f;
  P.q[0] = 1.f;
  P.q[1] = P.q[2] = P.q[3] = 0.f;
  return P;
}

// FK of a serial arm streamed joint by joint: T_j = T_{j-1} o (P_j o M(q_j)) (fixed tail joints
// compose P_j only).  Returns the world pose of joint jt; with WITH_S also the motion-subspace
// column of every actuated joint, computed as its pose is produced.
struct NoJointHook {
  __device__ __forceinline__ void operator()(int, const Pose&) const {}
};

// on(j, T) runs once joint j's world pose T (and, with WITH_S, its subspace column) is known
template <int NA, int NJ, bool WITH_S, class Hook = NoJointHook>
__device__ __forceinline__ Pose chain_fk(RobotK* rb, int nj, const float* q, int jt, float (&S6)[NA][6],
                                         const Hook& on = Hook()) {
  Pose T = identity_pose(), Tt = identity_pose();
#pragma unroll
  for (int j = 0; j < NJ; ++j) {
    if (j < nj) {
      Pose L;
#pragma unroll
      for (int k = 0; k < 3; ++k) L.t[k] = rb->pT[j][k];
#pragma unroll
      for (int k = 0; k < 4; ++k) L.q[k] = rb->pT[j][3 + k];
      if (j < NA) {
        if (rb->joint_type[j] == 2) {
          // P o M(q): the translation is P's, the rotation q_P (c, s axis) = c q_P + s q_P (0, axis),
          // standardised to w >= 0 as quat_mul's products are
          float s, c;
          rk_sincos(q[j] * 0.5f, &s, &c);
          float w = fmaf(s, rb->pqa[j][0], c * L.q[0]);
          const float sg = w < 0.f ? -1.f : 1.f;
          L.q[0] = sg * w;
#pragma unroll
          for (int k = 1; k < 4; ++k) L.q[k] = sg * fmaf(s, rb->pqa[j][k], c * L.q[k]);
        } else {
          Pose X = identity_pose();
          X.t[0] = rb->axis[j][0] * q[j]; X.t[1] = rb->axis[j][1] * q[j]; X.t[2] = rb->axis[j][2] * q[j];
          L = compose(L, X);
        }
      }
      T = j == 0 ? L : compose(T, L);
      if (WITH_S && j < NA) subspace_column(rb, j, T, S6[j]);
      if (j == jt) Tt = T;
      on(j, T);
    }
  }
  return Tt;
}

// residual rows at a point: the target's six rows, limit rows and smoothness rows; with JAC also
// the target's Jacobian data and the diagonal Jacobian entries of limit / smoothness rows.  The
// target rows are either a FrameTask target (log, 6 rows; tg = its Jacobian blocks) or a
// PositionTask target (rows 0..2, r = w (p_tgt - p_link), position_task.py:188-212; tg[0..2] =
// p_link) and / or a RotationTask target (rows 3..5, r = w axis_angle(q_link q_tgt^-1),
// rotation_task.py:323-349) on links of one parent joint -- the term pair of the IK presets.  In
// that form tgt.t holds the position target and tgt.q the target rotation.
template <int NA, bool JAC, bool PR>
__device__ __forceinline__ void chain_rows(RobotK* rb, TermsK* tm, const Pose& Tf, const Pose& tgt, const float* q,
                                           const float* qprev, float (&r_pose)[6], float (&r_lim)[NA],
                                           float (&r_sm)[NA], float (&tg)[36], float (&jl)[NA], float (&js)[NA]) {
  if constexpr (!PR) {
    if (tm->n_pose) {
      float w[6];
#pragma unroll
      for (int i = 0; i < 6; ++i) w[i] = tm->pose_w[0][i];
      pose_target(Tf, tgt, w, r_pose, JAC ? tg : nullptr);
    }
  } else {
#pragma unroll
    for (int i = 0; i < 6; ++i) r_pose[i] = 0.f;
    if (tm->n_pos) {
#pragma unroll
      for (int i = 0; i < 3; ++i) {
        r_pose[i] = tm->pos_w[0][i] * (tgt.t[i] - Tf.t[i]);
        if (JAC) tg[i] = Tf.t[i];
      }
    }
    if (tm->n_rot) {
      float aa[3];
      rotation_target(Tf, tgt.q, aa);
#pragma unroll
      for (int i = 0; i < 3; ++i) r_pose[3 + i] = tm->rot_w[0][i] * aa[i];
    }
  }
  if (tm->limit) {
#pragma unroll
    for (int d = 0; d < NA; ++d) {
      r_lim[d] = limit_row(q[d], rb->lo[d], rb->hi[d], tm->limit_w[d]);
      if (JAC) jl[d] = limit_diag(q[d], rb->lo[d], rb->hi[d], tm->limit_w[d]);
    }
  }
  if (tm->smooth) {
#pragma unroll
    for (int d = 0; d < NA; ++d) {
      r_sm[d] = tm->smooth_w[d] * (q[d] - qprev[d]);
      if (JAC) js[d] = tm->smooth_w[d];
    }
  }
}

// cost = sum over terms (pose, limit, smoothness) of 0.5 * sum of squares of the term's rows
template <int NA, bool PR>
__device__ __forceinline__ float chain_cost(TermsK* tm, const float* r_pose, const float* r_lim, const float* r_sm) {
  float cost = 0.f;
  if constexpr (!PR) {
    if (tm->n_pose) {
      float s = 0.f;
#pragma unroll
      for (int k = 0; k < 6; ++k) s += r_pose[k] * r_pose[k];
      cost = cost + 0.5f * s;
    }
  } else {
    if (tm->n_pos) {
      float s = 0.f;
#pragma unroll
      for (int k = 0; k < 3; ++k) s += r_pose[k] * r_pose[k];
      cost = cost + 0.5f * s;
    }
    if (tm->n_rot) {
      float s = 0.f;
#pragma unroll
      for (int k = 3; k < 6; ++k) s += r_pose[k] * r_pose[k];
      cost = cost + 0.5f * s;
    }
  }
  if (tm->limit) {
    float s = 0.f;
#pragma unroll
    for (int k = 0; k < NA; ++k) s += r_lim[k] * r_lim[k];
    cost = cost + 0.5f * s;
  }
  if (tm->smooth) {
    float s = 0.f;
#pragma unroll
    for (int k = 0; k < NA; ++k) s += r_sm[k] * r_sm[k];
    cost = cost + 0.5f * s;
  }
  return cost;
}

// SceneCollisionTask penalty of a signed distance (sphere_collision_kernels.py:19-42): the residual
// and its derivative.  mode 0 smooth: margin / 2 - sd inside, (sd - margin)^2 / (2 (margin + 1e-6))
// up to the margin; 1 plain: margin - sd below the margin; 2 surface distance: |sd|.
__device__ __forceinline__ void coll_penalty(float sd, float margin, int mode, float& pen, float& dpen) {
  if (mode == 2) {
    pen = fabsf(sd);
    dpen = sd < 0.f ? -1.f : 1.f;
  } else if (mode == 1) {
    pen = sd < margin ? margin - sd : 0.f;
    dpen = sd < margin ? -1.f : 0.f;
  } else if (sd < 0.f) {
    pen = 0.5f * margin - sd;
    dpen = -1.f;
  } else if (sd <= margin) {
    const float diff = sd - margin, inv = 1.f / (margin + 1e-6f);
    pen = 0.5f * inv * diff * diff;
    dpen = inv * diff;
  } else {
    pen = dpen = 0.f;
  }
}

// One collision row of sphere k at its link's world pose T (scene_collision_task.py:33-81, 550-646):
// r = w pen(sd); with JAC its Jacobian columns a <= j (the joints above the link) are
// w pen'(sd) n . (v_a + w_a x c), n the scene gradient at the centre c, zero beyond the margin
// (surface distance: always), accumulated as the term's J^T J (Hc, lower) and J^T r (gc)
template <int NA, bool JAC>
__device__ __forceinline__ void coll_row(TermsK* tm, int k, int j, const Pose& T, const float (&S6)[NA][6],
                                         float (&Hc)[NA][NA], float (&gc)[NA], float& cc) {
  const float lc[3] = {tm->c_c[k][0], tm->c_c[k][1], tm->c_c[k][2]};
  float c[3];
  quat_apply(T.q, lc, c);
  c[0] += T.t[0];
  c[1] += T.t[1];
  c[2] += T.t[2];
  int grp;
  const float sd = scene_sdf(tm, c, grp) - tm->c_r[k];
  float pen, dpen;
  coll_penalty(sd, tm->c_margin, tm->c_penalty, pen, dpen);
  const float w = tm->c_w[k];
  const float r = w * pen;
  cc += r * r;
  if (JAC && (tm->c_penalty == 2 || sd <= tm->c_margin)) {
    float n[3];
    scene_normal(tm, c, scene_winner(tm, c, grp), n);
    const float ws = w * dpen;
    float Jk[NA];
#pragma unroll
    for (int a = 0; a < NA; ++a) {
      float wxc[3];
      cross3(S6[a] + 3, c, wxc);
      Jk[a] = a <= j ? ws * (n[0] * (S6[a][0] + wxc[0]) + n[1] * (S6[a][1] + wxc[1]) + n[2] * (S6[a][2] + wxc[2]))
                     : 0.f;
    }
#pragma unroll
    for (int a = 0; a < NA; ++a) {
#pragma unroll
      for (int b = 0; b <= a; ++b) Hc[a][b] += Jk[a] * Jk[b];
      gc[a] += Jk[a] * r;
    }
  }
}

// The linearisation of a serial-arm problem at qv: FK with the motion subspace, the residual rows,
// the Jacobian (target rows: FrameTask -w jlog Ad(T_f^-1) S; PositionTask -w (v + w x p_link),
// position_task.py:253-262; RotationTask w * the column's angular part, rotation_task.py:352-401;
// zero for joints after the target link's joint), J^T J (lower triangle, per-term partial sums,
// rows in order), J^T r; returns the cost.
template <int NA, int NJ, bool PR>
__device__ __forceinline__ float chain_linearize(RobotK* rb, TermsK* tm, int nj, int jt, const Pose& tgt,
                                                 const float* qv, const float* qprev, float (&H)[NA][NA],
                                                 float (&g)[NA]) {
  const int n_pose = PR ? 0 : tm->n_pose, n_pos = PR ? tm->n_pos : 0, n_rot = PR ? tm->n_rot : 0;
  float S6[NA][6];
  const Pose Tf = chain_fk<NA, NJ, true>(rb, nj, qv, jt, S6);
  float r_pose[6], r_lim[NA], r_sm[NA], jl[NA], js[NA], tg[36];
  chain_rows<NA, true, PR>(rb, tm, Tf, tgt, qv, qprev, r_pose, r_lim, r_sm, tg, jl, js);
  float J[6][NA];
  float M[27];  // the target's column map, once for all the dofs' columns
  if constexpr (!PR) pose_jac_matrix(tg, M);
#pragma unroll
  for (int a = 0; a < NA; ++a) {
    float c6[6] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    if (a <= jt) {
      if constexpr (!PR) {
        if (n_pose) pose_jac_apply(M, S6[a], c6);
      } else {
        if (n_pos) {
          const float w3[3] = {tm->pos_w[0][0], tm->pos_w[0][1], tm->pos_w[0][2]};
          pos_jac_column(S6[a], tg, w3, c6);
        }
        if (n_rot) {
#pragma unroll
          for (int i = 0; i < 3; ++i) c6[3 + i] = tm->rot_w[0][i] * S6[a][3 + i];
        }
      }
    }
#pragma unroll
    for (int i = 0; i < 6; ++i) J[i][a] = c6[i];
  }
#pragma unroll
  for (int i = 0; i < NA; ++i) {
#pragma unroll
    for (int j = 0; j <= i; ++j) {
      float h = 0.f;
      if (!PR && n_pose && i <= jt) {
        float sp = 0.f;
#pragma unroll
        for (int r = 0; r < 6; ++r) sp += J[r][i] * J[r][j];
        h = sp;
      } else if (PR && i <= jt) {  // position term, then rotation term
        if (n_pos) {
          float sp = 0.f;
#pragma unroll
          for (int r = 0; r < 3; ++r) sp += J[r][i] * J[r][j];
          h = sp;
        }
        if (n_rot) {
          float sr = 0.f;
#pragma unroll
          for (int r = 3; r < 6; ++r) sr += J[r][i] * J[r][j];
          h = h + sr;
        }
      }
      if (i == j) {
        if (tm->limit) h = h + jl[i] * jl[i];
        if (tm->smooth) h = h + js[i] * js[i];
      }
      H[i][j] = h;
    }
    float gv = 0.f;
    if (!PR && n_pose && i <= jt) {
      float sp = 0.f;
#pragma unroll
      for (int r = 0; r < 6; ++r) sp += J[r][i] * r_pose[r];
      gv = sp;
    } else if (PR && i <= jt) {
      if (n_pos) {
        float sp = 0.f;
#pragma unroll
        for (int r = 0; r < 3; ++r) sp += J[r][i] * r_pose[r];
        gv = sp;
      }
      if (n_rot) {
        float sr = 0.f;
#pragma unroll
        for (int r = 3; r < 6; ++r) sr += J[r][i] * r_pose[r];
        gv = gv + sr;
      }
    }
    if (tm->limit) gv = gv + jl[i] * r_lim[i];
    if (tm->smooth) gv = gv + js[i] * r_sm[i];
    g[i] = gv;
  }
  return chain_cost<NA, PR>(tm, r_pose, r_lim, r_sm);
}

// damped Cholesky (left-looking) of H + lam I into L and the solve (H + lam I) dq = -g; returns the
// predicted reduction 1/2 dq.(lam dq - g).  Each pivot's reciprocal square root is the hardware's
// (v_rsq_f32, 1 ulp) and the column scales by it: a correctly rounded division is a dependent chain
// of ~10 instructions, and these kernels are bound by dependency latency.
template <int NA>
__device__ __forceinline__ float chain_step(const float (&H)[NA][NA], const float (&g)[NA], float lam, float (&dq)[NA]) {
  float L[NA][NA], dinv[NA];
#pragma unroll
  for (int j = 0; j < NA; ++j) {
    float sd = H[j][j] + lam;
#pragma unroll
    for (int k = 0; k < j; ++k) sd -= L[j][k] * L[j][k];
    dinv[j] = __builtin_amdgcn_rsqf(sd);  // the pivot itself is never read: only 1 / L_jj
#pragma unroll
    for (int i = j + 1; i < NA; ++i) {
      float v = H[i][j];
#pragma unroll
      for (int k = 0; k < j; ++k) v -= L[i][k] * L[j][k];
      L[i][j] = v * dinv[j];
    }
  }
#pragma unroll
  for (int i = 0; i < NA; ++i) {
    float v = -g[i];
#pragma unroll
    for (int k = 0; k < i; ++k) v -= L[i][k] * dq[k];
    dq[i] = v * dinv[i];
  }
#pragma unroll
  for (int i = NA - 1; i >= 0; --i) {
    float v = dq[i];
#pragma unroll
    for (int k = i + 1; k < NA; ++k) v -= L[k][i] * dq[k];
    dq[i] = v * dinv[i];
  }
  float pr = 0.f;
#pragma unroll
  for (int i = 0; i < NA; ++i) pr += dq[i] * (lam * dq[i] - g[i]);
  return 0.5f * pr;
}

// One serial-arm problem per lane without collision rows.  An accepted proposal's FK is the next
// iterate's, so the problem is linearised once per iteration, at the proposal: its cost decides
// the step, and its J^T J, J^T r and cost replace the iterate's when the step is accepted (a
// rejected step keeps them: the iterate did not move).  One FK per iteration instead of two.
template <int NA, int NJ, bool PR>
__device__ __forceinline__ void thread_solve_lin(const Problems& Pg, int s_idx, const LMParams& lp, int it0,
                                                 int iters) {
  RobotK* rb = as_const(Pg.robot);
  TermsK* tm = as_const(Pg.terms);
  const int S = Pg.stride;
  const bool valid = s_idx < Pg.n;
  const int s = valid ? s_idx : Pg.n - 1;
  const int nj = rb->nj;
  const int n_pose = PR ? 0 : tm->n_pose, n_pos = PR ? tm->n_pos : 0, n_rot = PR ? tm->n_rot : 0;
  const int jt = n_pose ? rb->link_joint[tm->pose_link[0]]
                 : n_pos ? rb->link_joint[tm->pos_link[0]]
                 : n_rot ? rb->link_joint[tm->rot_link[0]]
                         : -1;
  float q[NA], qprev[NA];
#pragma unroll
  for (int d = 0; d < NA; ++d) {
    q[d] = (it0 == 0 && Pg.q_seed) ? Pg.q_seed[d * S + s] : Pg.q[d * S + s];
    qprev[d] = tm->smooth ? Pg.q_prev[d * S + s] : 0.f;
  }
  Pose tgt = identity_pose();
  if constexpr (!PR) {
    if (n_pose) {
      float v[7];
#pragma unroll
      for (int c = 0; c < 7; ++c) v[c] = Pg.pose_tgt[c * S + s];
      load_pose(v, tgt);
    }
  } else {
    if (n_pos) {
#pragma unroll
      for (int c = 0; c < 3; ++c) tgt.t[c] = Pg.pos_tgt[c * S + s];
    }
    if (n_rot) {
#pragma unroll
      for (int c = 0; c < 4; ++c) tgt.q[c] = Pg.rot_tgt[c * S + s];
    }
  }
  float lam = it0 == 0 ? lp.lm_lambda : Pg.lam[s];
  int nacc = it0 == 0 ? 0 : Pg.nacc[s];
  float H[NA][NA], g[NA];
  float cost = iters > 0 ? chain_linearize<NA, NJ, PR>(rb, tm, nj, jt, tgt, q, qprev, H, g) : 0.f;
  for (int it = 0; it < iters; ++it) {
    const int gi = it0 + it;
    if (valid && Pg.q_trace && gi < Pg.trace_iters) {
#pragma unroll
      for (int d = 0; d < NA; ++d) Pg.q_trace[(gi * NA + d) * S + s] = q[d];
      Pg.lam_trace[gi * S + s] = lam;
    }
    float dq[NA];
    const float pred = chain_step<NA>(H, g, lam, dq);
    float qp[NA];
#pragma unroll
    for (int d = 0; d < NA; ++d) qp[d] = q[d] + dq[d];
    float Hp[NA][NA], gp[NA];
    const float cp = chain_linearize<NA, NJ, PR>(rb, tm, nj, jt, tgt, qp, qprev, Hp, gp);
    if (valid && Pg.trace && gi == 0) Pg.trace[s] = cost;
    const bool acc = lm_accept(lp, cost, cp, pred, lam);
    if (acc) {
      ++nacc;
#pragma unroll
      for (int d = 0; d < NA; ++d) {
        q[d] = qp[d];
        g[d] = gp[d];
#pragma unroll
        for (int e = 0; e <= d; ++e) H[d][e] = Hp[d][e];
      }
    }
    if (valid && Pg.trace && gi < Pg.trace_iters) {
      Pg.trace[(gi + 1) * S + s] = cost;
      Pg.acc_trace[gi * S + s] = acc;
    }
  }
  if (valid) {
#pragma unroll
    for (int d = 0; d < NA; ++d) Pg.q[d * S + s] = q[d];
    if (iters > 0) Pg.cost[s] = cost;
    Pg.lam[s] = lam;
    Pg.nacc[s] = nacc;
  }
}

// One serial-arm problem per lane: <= 1 pose target (PR false), or <= 1 position and <= 1 rotation
// target on links of one parent joint (PR true); limit and smoothness rows; with CO the
// SceneCollisionTask rows of the robot's spheres, evaluated as FK reaches each sphere's link.  The
// forms are separate kernels so that each keeps its own register allocation.
template <int NA, int NJ, bool PR = false, bool CO = false>
__device__ __forceinline__ void thread_solve(const Problems& Pg, int s_idx, const LMParams& lp, int it0, int iters) {
  if constexpr (!CO) {
    thread_solve_lin<NA, NJ, PR>(Pg, s_idx, lp, it0, iters);
    return;
  }
  RobotK* rb = as_const(Pg.robot);
  TermsK* tm = as_const(Pg.terms);
  const int S = Pg.stride;
  const bool valid = s_idx < Pg.n;
  const int s = valid ? s_idx : Pg.n - 1;
  const int nj = rb->nj;
  const int n_pose = PR ? 0 : tm->n_pose, n_pos = PR ? tm->n_pos : 0, n_rot = PR ? tm->n_rot : 0;
  const int jt = n_pose ? rb->link_joint[tm->pose_link[0]]
                 : n_pos ? rb->link_joint[tm->pos_link[0]]
                 : n_rot ? rb->link_joint[tm->rot_link[0]]
                         : -1;

  float q[NA], qprev[NA];
#pragma unroll
  for (int d = 0; d < NA; ++d) {
    q[d] = (it0 == 0 && Pg.q_seed) ? Pg.q_seed[d * S + s] : Pg.q[d * S + s];
    qprev[d] = tm->smooth ? Pg.q_prev[d * S + s] : 0.f;
  }
  Pose tgt = identity_pose();
  if constexpr (!PR) {
    if (n_pose) {
      float v[7];
#pragma unroll
      for (int c = 0; c < 7; ++c) v[c] = Pg.pose_tgt[c * S + s];
      load_pose(v, tgt);
    }
  } else {
    if (n_pos) {
#pragma unroll
      for (int c = 0; c < 3; ++c) tgt.t[c] = Pg.pos_tgt[c * S + s];
    }
    if (n_rot) {
#pragma unroll
      for (int c = 0; c < 4; ++c) tgt.q[c] = Pg.rot_tgt[c * S + s];
    }
  }
  float lam = it0 == 0 ? lp.lm_lambda : Pg.lam[s];
  int nacc = it0 == 0 ? 0 : Pg.nacc[s];
  float cost = 0.f;

  for (int it = 0; it < iters; ++it) {
    const int gi = it0 + it;
    if (valid && Pg.q_trace && gi < Pg.trace_iters) {
#pragma unroll
      for (int d = 0; d < NA; ++d) Pg.q_trace[(gi * NA + d) * S + s] = q[d];
      Pg.lam_trace[gi * S + s] = lam;
    }
    // ---- FK with motion subspace (and the collision rows as the spheres' links appear), residual
    // rows, Jacobian blocks at q
    float S6[NA][6];
    float Hc[NA][NA], gc[NA], cc = 0.f;
    if constexpr (CO) {
#pragma unroll
      for (int a = 0; a < NA; ++a) {
        gc[a] = 0.f;
#pragma unroll
        for (int b = 0; b < NA; ++b) Hc[a][b] = 0.f;
        for (int r = 0; r < 6; ++r) S6[a][r] = 0.f;
      }
    }
    int kc = 0;
    if constexpr (CO)
      for (; kc < tm->nc && tm->c_joint[kc] < 0; ++kc) coll_row<NA, false>(tm, kc, -1, identity_pose(), S6, Hc, gc, cc);
    auto coll_at = [&](int j, const Pose& T) {
      if constexpr (CO)
        for (; kc < tm->nc && tm->c_joint[kc] == j; ++kc) coll_row<NA, true>(tm, kc, j, T, S6, Hc, gc, cc);
    };
    const Pose Tf = chain_fk<NA, NJ, true>(rb, nj, q, jt, S6, coll_at);
    float r_pose[6], r_lim[NA], r_sm[NA], jl[NA], js[NA], tg[36];
    chain_rows<NA, true, PR>(rb, tm, Tf, tgt, q, qprev, r_pose, r_lim, r_sm, tg, jl, js);
    // ---- target-row Jacobian columns (zero for joints after the target link's joint): FrameTask
    // -w jlog Ad(T_f^-1) S; PositionTask -w (v + w x p_link) (position_task.py:253-262); RotationTask
    // w * (the column's angular part) (rotation_task.py:352-401)
    float J[6][NA];
#pragma unroll
    for (int a = 0; a < NA; ++a) {
      float c6[6] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
      if (a <= jt) {
        if constexpr (!PR) {
          if (n_pose) pose_jac_column(tg, S6[a], c6);
        } else {
          if (n_pos) {
            const float w3[3] = {tm->pos_w[0][0], tm->pos_w[0][1], tm->pos_w[0][2]};
            pos_jac_column(S6[a], tg, w3, c6);
          }
          if (n_rot) {
#pragma unroll
            for (int i = 0; i < 3; ++i) c6[3 + i] = tm->rot_w[0][i] * S6[a][3 + i];
          }
        }
      }
#pragma unroll
      for (int i = 0; i < 6; ++i) J[i][a] = c6[i];
    }
    // ---- normal equations (per-term partial sums, rows in order)
    float H[NA][NA], g[NA];
#pragma unroll
    for (int i = 0; i < NA; ++i) {
#pragma unroll
      for (int j = 0; j <= i; ++j) {
        float h = 0.f;
        if (!PR && n_pose && i <= jt) {
          float sp = 0.f;
#pragma unroll
          for (int r = 0; r < 6; ++r) sp += J[r][i] * J[r][j];
          h = sp;
        } else if (PR && i <= jt) {  // position term, then rotation term
          if (n_pos) {
            float sp = 0.f;
#pragma unroll
            for (int r = 0; r < 3; ++r) sp += J[r][i] * J[r][j];
            h = sp;
          }
          if (n_rot) {
            float sr = 0.f;
#pragma unroll
            for (int r = 3; r < 6; ++r) sr += J[r][i] * J[r][j];
            h = h + sr;
          }
        }
        if (i == j) {
          if (tm->limit) h = h + jl[i] * jl[i];
          if (tm->smooth) h = h + js[i] * js[i];
        }
        H[i][j] = h;
      }
      float gv = 0.f;
      if (!PR && n_pose && i <= jt) {
        float sp = 0.f;
#pragma unroll
        for (int r = 0; r < 6; ++r) sp += J[r][i] * r_pose[r];
        gv = sp;
      } else if (PR && i <= jt) {
        if (n_pos) {
          float sp = 0.f;
#pragma unroll
          for (int r = 0; r < 3; ++r) sp += J[r][i] * r_pose[r];
          gv = sp;
        }
        if (n_rot) {
          float sr = 0.f;
#pragma unroll
          for (int r = 3; r < 6; ++r) sr += J[r][i] * r_pose[r];
          gv = gv + sr;
        }
      }
      if (tm->limit) gv = gv + jl[i] * r_lim[i];
      if (tm->smooth) gv = gv + js[i] * r_sm[i];
      if (CO) gv = gv + gc[i];
      g[i] = gv;
    }
    if constexpr (CO) {  // the collision term after the others
#pragma unroll
      for (int i = 0; i < NA; ++i)
#pragma unroll
        for (int j = 0; j <= i; ++j) H[i][j] = H[i][j] + Hc[i][j];
    }
    cost = chain_cost<NA, PR>(tm, r_pose, r_lim, r_sm);
    if (CO) cost = cost + 0.5f * cc;
    // ---- damped Cholesky (left-looking) and solves.  Each pivot's reciprocal square root is the
    // hardware's (v_rsq_f32, 1 ulp) and the column scales by it: a correctly rounded division is a
    // dependent chain of ~10 instructions, and this kernel is bound by dependency latency.
    float dinv[NA];
#pragma unroll
    for (int j = 0; j < NA; ++j) {
      float sd = H[j][j] + lam;
#pragma unroll
      for (int k = 0; k < j; ++k) sd -= H[j][k] * H[j][k];
      dinv[j] = __builtin_amdgcn_rsqf(sd);  // the pivot itself is never read: only 1 / L_jj
#pragma unroll
      for (int i = j + 1; i < NA; ++i) {
        float v = H[i][j];
#pragma unroll
        for (int k = 0; k < j; ++k) v -= H[i][k] * H[j][k];
        H[i][j] = v * dinv[j];
      }
    }
    float dq[NA];
#pragma unroll
    for (int i = 0; i < NA; ++i) {
      float v = -g[i];
#pragma unroll
      for (int k = 0; k < i; ++k) v -= H[i][k] * dq[k];
      dq[i] = v * dinv[i];
    }
#pragma unroll
    for (int i = NA - 1; i >= 0; --i) {
      float v = dq[i];
#pragma unroll
      for (int k = i + 1; k < NA; ++k) v -= H[k][i] * dq[k];
      dq[i] = v * dinv[i];
    }
    float pr = 0.f;
#pragma unroll
    for (int i = 0; i < NA; ++i) pr += dq[i] * (lam * dq[i] - g[i]);
    const float pred = 0.5f * pr;
    // ---- proposal: FK to the target link only, residual rows, cost
    float qp[NA];
#pragma unroll
    for (int d = 0; d < NA; ++d) qp[d] = q[d] + dq[d];
    float cpc = 0.f;
    int kp = 0;
    if constexpr (CO)
      for (; kp < tm->nc && tm->c_joint[kp] < 0; ++kp) coll_row<NA, false>(tm, kp, -1, identity_pose(), S6, Hc, gc, cpc);
    auto coll_prop = [&](int j, const Pose& T) {
      if constexpr (CO)
        for (; kp < tm->nc && tm->c_joint[kp] == j; ++kp) coll_row<NA, false>(tm, kp, j, T, S6, Hc, gc, cpc);
    };
    const Pose Tfp = chain_fk<NA, NJ, false>(rb, nj, qp, jt, S6, coll_prop);
    chain_rows<NA, false, PR>(rb, tm, Tfp, tgt, qp, qprev, r_pose, r_lim, r_sm, tg, jl, js);
    float cp = chain_cost<NA, PR>(tm, r_pose, r_lim, r_sm);
    if (CO) cp = cp + 0.5f * cpc;
    if (valid && Pg.trace && gi == 0) Pg.trace[s] = cost;
    const bool acc = lm_accept(lp, cost, cp, pred, lam);
    if (acc) {
      ++nacc;
#pragma unroll
      for (int d = 0; d < NA; ++d) q[d] = qp[d];
    }
    if (valid && Pg.trace && gi < Pg.trace_iters) {
      Pg.trace[(gi + 1) * S + s] = cost;
      Pg.acc_trace[gi * S + s] = acc;
    }
  }
  if (valid) {
#pragma unroll
    for (int d = 0; d < NA; ++d) Pg.q[d * S + s] = q[d];
    if (iters > 0) Pg.cost[s] = cost;
    Pg.lam[s] = lam;
    Pg.nacc[s] = nacc;
  }
}

template <int NA, int NJ, bool PR = false, bool CO = false>
__global__ __launch_bounds__(64) void k_thread(Problems P, LMParams lp, int it0, int iters) {
  // wave issue priority 1 (trajectory waves stay at 0): where IK and trajectory launches share a
  // SIMD, as in the request server, the short IK solves issue first
  __builtin_amdgcn_s_setprio(1);
  thread_solve<NA, NJ, PR, CO>(P, xcd_order(blockIdx.x, gridDim.x) * 64 + threadIdx.x, lp, it0, iters);
}

// rk_quad.h -- four lanes per problem LM for serial arms (included by rk_lm.hip inside namespace rk,
// after rk_thread.h; its quad DPP helpers serve rk_wave.h too).
//
// The one-lane-per-problem kernel (rk_thread.h) does the least arithmetic per problem, but one
// problem's iteration is a serial chain of ~3.7k instructions and 6554 problems fill only 103
// wavefronts: it is bound by that chain's latency.  Here a quad of lanes shares a problem (16
// problems per wavefront, four times the wavefronts) and splits the longest serial parts:
//   - forward kinematics: lane k of the quad owns joints 2k and 2k + 1; it forms their local
//     transforms and their product, an inclusive scan over the quad (two rounds, each a quad_perm
//     DPP move of the neighbour's product and one compose) gives every lane the world pose after
//     its joints, and one more compose the pose between them;
//   - the motion-subspace and Jacobian columns of the lane's two dofs;
// then the columns are broadcast over the quad (quad_perm DPP moves) and every lane runs the
// remaining short parts -- the target's log map and blocks, J^T J, the damped Cholesky, the
// solves, the acceptance test -- itself, on identical data, so the quad stays in lockstep without
// any exchange.  The per-entry sums of J^T J, J^T r and the cost are the one-lane kernel's, and as
// there the problem is linearised once per iteration, at the proposal.

// quad_perm DPP control for lanes 0..3 of every quad: lane i reads lane s_i
constexpr int qp_ctrl(int s0, int s1, int s2, int s3) { return s0 | (s1 << 2) | (s2 << 4) | (s3 << 6); }

template <int CTRL>
__device__ __forceinline__ float qperm(float v) {
  return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), CTRL, 0xF, 0xF, false));
}
template <int CTRL>
__device__ __forceinline__ Pose qperm_pose(const Pose& P) {
  Pose o;
#pragma unroll
  for (int k = 0; k < 3; ++k) o.t[k] = qperm<CTRL>(P.t[k]);
#pragma unroll
  for (int k = 0; k < 4; ++k) o.q[k] = qperm<CTRL>(P.q[k]);
  return o;
}
// the value of quad lane K
template <int K>
__device__ __forceinline__ float qbcast(float v) {
  return qperm<qp_ctrl(K, K, K, K)>(v);
}
template <int K>
__device__ __forceinline__ Pose qbcast_pose(const Pose& P) {
  return qperm_pose<qp_ctrl(K, K, K, K)>(P);
}

__device__ __forceinline__ Pose pose_select(bool c, const Pose& a, const Pose& b) {
  Pose o;
#pragma unroll
  for (int k = 0; k < 3; ++k) o.t[k] = c ? a.t[k] : b.t[k];
#pragma unroll
  for (int k = 0; k < 4; ++k) o.q[k] = c ? a.q[k] : b.q[k];
  return o;
}

// one of the quad's two joint slots: the joint's parent transform, axis and type (identity
// transform and type 0 past the actuated joints)
struct QuadJoint {
  Pose P;
  float ax[3];
  int type;
};

// P_j o M(q_j) of a slot
__device__ __forceinline__ Pose quad_local(const QuadJoint& J, float qj) {
  Pose X = identity_pose();
  if (J.type == 2) {
    float s, c;
    rk_sincos(qj * 0.5f, &s, &c);
    X.q[0] = c; X.q[1] = J.ax[0] * s; X.q[2] = J.ax[1] * s; X.q[3] = J.ax[2] * s;
  } else if (J.type == 1) {
    X.t[0] = J.ax[0] * qj; X.t[1] = J.ax[1] * qj; X.t[2] = J.ax[2] * qj;
  }
  return compose(J.P, X);
}

// q of the dof in slot s of quad lane ql (dof 2 ql + s), from the quad's full copy of q
template <int NA>
__device__ __forceinline__ float quad_q(const float (&q)[NA], int ql, int s) {
  float v = 0.f;
#pragma unroll
  for (int k = 0; k < 4; ++k)
    if (2 * k + s < NA) v = ql == k ? q[2 * k + s] : v;
  return v;
}

// world poses of the lane's two joints (T0: joint 2 ql, T1: joint 2 ql + 1) by the quad scan
template <int NA>
__device__ __forceinline__ void quad_fk(const QuadJoint (&J)[2], const float (&q)[NA], int ql, Pose& T0, Pose& T1) {
  const Pose L0 = quad_local(J[0], quad_q<NA>(q, ql, 0));
  const Pose L1 = quad_local(J[1], quad_q<NA>(q, ql, 1));
  Pose B = compose(L0, L1);
  {
    const Pose t = qperm_pose<qp_ctrl(0, 0, 1, 2)>(B);  // lane i <- lane i - 1
    B = pose_select(ql >= 1, compose(t, B), B);
  }
  {
    const Pose t = qperm_pose<qp_ctrl(0, 1, 0, 1)>(B);  // lane i <- lane i - 2
    B = pose_select(ql >= 2, compose(t, B), B);
  }
  const Pose E = pose_select(ql == 0, identity_pose(), qperm_pose<qp_ctrl(0, 0, 1, 2)>(B));
  T0 = compose(E, L0);
  T1 = B;
}

// world pose of joint jt (runtime, uniform) from the quad lanes' joint poses
__device__ __forceinline__ Pose quad_joint_pose(const Pose& T0, const Pose& T1, int jt) {
  const int k = jt >> 1;
  Pose a, b;
  switch (k) {
    case 0: a = qbcast_pose<0>(T0); b = qbcast_pose<0>(T1); break;
    case 1: a = qbcast_pose<1>(T0); b = qbcast_pose<1>(T1); break;
    case 2: a = qbcast_pose<2>(T0); b = qbcast_pose<2>(T1); break;
    default: a = qbcast_pose<3>(T0); b = qbcast_pose<3>(T1); break;
  }
  return (jt & 1) ? b : a;
}

// the world motion-subspace column of a slot's joint at its world pose (robot_kernels.py:650-664)
__device__ __forceinline__ void quad_subspace(const QuadJoint& J, const Pose& T, float* S6) {
  float Ra[3];
  quat_apply(T.q, J.ax, Ra);
  if (J.type == 2) {
    float txRa[3];
    cross3(T.t, Ra, txRa);
    S6[0] = txRa[0]; S6[1] = txRa[1]; S6[2] = txRa[2];
    S6[3] = Ra[0]; S6[4] = Ra[1]; S6[5] = Ra[2];
  } else if (J.type == 1) {
    S6[0] = Ra[0]; S6[1] = Ra[1]; S6[2] = Ra[2];
    S6[3] = S6[4] = S6[5] = 0.f;
  } else {
    S6[0] = S6[1] = S6[2] = S6[3] = S6[4] = S6[5] = 0.f;
  }
}

// One serial-arm problem per quad: <= 1 FrameTask pose target, limit and smoothness rows.
// tile = the wavefront's group of 16 problems.
template <int NA, int NJ>
__device__ __forceinline__ void quad_solve(const Problems& Pg, int tile, const LMParams& lp, int it0, int iters) {
  static_assert(NA >= 2 && NA <= 8, "a quad holds two joints per lane");
  RobotK* rb = as_const(Pg.robot);
  TermsK* tm = as_const(Pg.terms);
  const int S = Pg.stride;
  const int ql = threadIdx.x & 3;
  const int s_idx = tile * (WAVE / 4) + (threadIdx.x >> 2);
  const bool valid = s_idx < Pg.n;
  const bool lead = valid && ql == 0;  // the lane that writes the problem's outputs
  const int s = valid ? s_idx : Pg.n - 1;
  const int nj = rb->nj;
  const int n_pose = tm->n_pose;
  const int jt = n_pose ? rb->link_joint[tm->pose_link[0]] : -1;

  // the lane's two joints, and the fixed joints between the last actuated joint and the target
  QuadJoint J[2];
#pragma unroll
  for (int sl = 0; sl < 2; ++sl) {
    const int j = 2 * ql + sl;
    const int jj = j < NA ? j : 0;
    const Robot* rg = Pg.robot;
    Pose P;
#pragma unroll
    for (int k = 0; k < 3; ++k) P.t[k] = rg->pT[jj][k];
#pragma unroll
    for (int k = 0; k < 4; ++k) P.q[k] = rg->pT[jj][3 + k];
    J[sl].P = j < NA ? P : identity_pose();
#pragma unroll
    for (int k = 0; k < 3; ++k) J[sl].ax[k] = rg->axis[jj][k];
    J[sl].type = j < NA ? rg->joint_type[jj] : 0;
  }
  Pose Ctail = identity_pose();  // P_NA o ... o P_jt when the target sits past the actuated joints
  for (int j = NA; j < nj && j <= jt; ++j) {
    Pose P;
#pragma unroll
    for (int k = 0; k < 3; ++k) P.t[k] = rb->pT[j][k];
#pragma unroll
    for (int k = 0; k < 4; ++k) P.q[k] = rb->pT[j][3 + k];
    Ctail = compose(Ctail, P);
  }
  // world pose of the target's joint from the lanes' joint poses
  auto target_pose = [&](const Pose& T0, const Pose& T1) {
    if (jt < 0) return identity_pose();
    if (jt >= NA - 1) {
      constexpr int k = (NA - 1) / 2;
      const Pose Tl = (NA - 1) & 1 ? qbcast_pose<k>(T1) : qbcast_pose<k>(T0);
      return jt >= NA ? compose(Tl, Ctail) : Tl;
    }
    return quad_joint_pose(T0, T1, jt);
  };

  float q[NA], qprev[NA];
#pragma unroll
  for (int d = 0; d < NA; ++d) {
    q[d] = (it0 == 0 && Pg.q_seed) ? Pg.q_seed[d * S + s] : Pg.q[d * S + s];
    qprev[d] = tm->smooth ? Pg.q_prev[d * S + s] : 0.f;
  }
  Pose tgt = identity_pose();
  if (n_pose) {
    float v[7];
#pragma unroll
    for (int c = 0; c < 7; ++c) v[c] = Pg.pose_tgt[c * S + s];
    load_pose(v, tgt);
  }
  float lam = it0 == 0 ? lp.lm_lambda : Pg.lam[s];
  int nacc = it0 == 0 ? 0 : Pg.nacc[s];

  // the linearisation at qv (every lane of the quad ends with the same bits): FK by the quad scan,
  // the target's rows and blocks, the lane's two Jacobian columns (zero for joints after the
  // target's joint) broadcast over the quad, then J^T J (lower), J^T r and the cost as the
  // one-lane kernel forms them
  auto linearize = [&](const float (&qv)[NA], float (&H)[NA][NA], float (&g)[NA]) {
    Pose T0, T1;
    quad_fk<NA>(J, qv, ql, T0, T1);
    const Pose Tf = target_pose(T0, T1);
    float r_pose[6], r_lim[NA], r_sm[NA], jl[NA], js[NA], tg[36];
    chain_rows<NA, true, false>(rb, tm, Tf, tgt, qv, qprev, r_pose, r_lim, r_sm, tg, jl, js);
    float c6[2][6];
#pragma unroll
    for (int sl = 0; sl < 2; ++sl) {
      float S6[6];
      quad_subspace(J[sl], sl ? T1 : T0, S6);
      float c[6] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
      if (n_pose && 2 * ql + sl <= jt) pose_jac_column(tg, S6, c);
#pragma unroll
      for (int i = 0; i < 6; ++i) c6[sl][i] = c[i];
    }
    float Jm[6][NA];
#pragma unroll
    for (int a = 0; a < NA; ++a)
#pragma unroll
      for (int i = 0; i < 6; ++i) {
        const float v = c6[a & 1][i];
        Jm[i][a] = (a >> 1) == 0 ? qbcast<0>(v) : (a >> 1) == 1 ? qbcast<1>(v) : (a >> 1) == 2 ? qbcast<2>(v) : qbcast<3>(v);
      }
#pragma unroll
    for (int i = 0; i < NA; ++i) {
#pragma unroll
      for (int j = 0; j <= i; ++j) {
        float h = 0.f;
        if (n_pose && i <= jt) {
          float sp = 0.f;
#pragma unroll
          for (int r = 0; r < 6; ++r) sp += Jm[r][i] * Jm[r][j];
          h = sp;
        }
        if (i == j) {
          if (tm->limit) h = h + jl[i] * jl[i];
          if (tm->smooth) h = h + js[i] * js[i];
        }
        H[i][j] = h;
      }
      float gv = 0.f;
      if (n_pose && i <= jt) {
        float sp = 0.f;
#pragma unroll
        for (int r = 0; r < 6; ++r) sp += Jm[r][i] * r_pose[r];
        gv = sp;
      }
      if (tm->limit) gv = gv + jl[i] * r_lim[i];
      if (tm->smooth) gv = gv + js[i] * r_sm[i];
      g[i] = gv;
    }
    return chain_cost<NA, false>(tm, r_pose, r_lim, r_sm);
  };

  // linearised once per iteration, at the proposal (an accepted proposal's FK is the next
  // iterate's; a rejected step keeps the iterate's normal equations)
  float H[NA][NA], g[NA];
  float cost = iters > 0 ? linearize(q, H, g) : 0.f;
  for (int it = 0; it < iters; ++it) {
    const int gi = it0 + it;
    if (lead && Pg.q_trace && gi < Pg.trace_iters) {
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
    const float cp = linearize(qp, Hp, gp);
    if (lead && Pg.trace && gi == 0) Pg.trace[s] = cost;
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
    if (lead && Pg.trace && gi < Pg.trace_iters) {
      Pg.trace[(gi + 1) * S + s] = cost;
      Pg.acc_trace[gi * S + s] = acc;
    }
  }
  if (lead) {
#pragma unroll
    for (int d = 0; d < NA; ++d) Pg.q[d * S + s] = q[d];
    if (iters > 0) Pg.cost[s] = cost;
    Pg.lam[s] = lam;
    Pg.nacc[s] = nacc;
  }
}

template <int NA, int NJ>
__global__ __launch_bounds__(64) void k_quad(Problems P, LMParams lp, int it0, int iters) {
  // wave issue priority 1 (trajectory waves stay at 0): where IK and trajectory launches share a
  // SIMD, as in the request server, the short IK solves issue first
  __builtin_amdgcn_s_setprio(1);
  quad_solve<NA, NJ>(P, xcd_order(blockIdx.x, gridDim.x), lp, it0, iters);
}

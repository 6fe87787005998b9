// rk_wave.h -- register-resident wave kernels for trees (included by rk_lm.hip inside namespace rk).
//
// HW lanes per problem: HW = 64 (one problem per wavefront, trees up to 64 joints) or HW = 32
// (two problems per wavefront, trees up to 32 joints).  For trees up to 32 dofs (e.g. a 29-dof
// humanoid).  Lane roles inside a problem's lanes (lane = threadIdx.x mod HW):
//   lane j  (< nj): joint j -- its local transform, world pose, motion-subspace column;
//   lane a  (< na): dof a   -- q_a, its Jacobian column (all pose / position rows), row a of
//                              J^T J + lam I and of its Cholesky factor;
//   lane t        : pose target t (then position target t - n_pose): residual rows and the
//                   per-target Jacobian blocks, broadcast to the problem's lanes through LDS.
// World poses come from pointer jumping over the parent tree (ceil(log2(depth+1)) rounds of
// __shfl of the ancestor's partial product).  The factorisation is the right-looking Cholesky on
// register rows; each step stages its column (and the right-hand side entry) in LDS, from where
// every lane reads the entries it needs as broadcasts, and the forward substitution rides along.
// The backward solve reads the factor's columns after one transpose through LDS.
// Row blocks of a target are structurally zero in the columns of dofs that do not move the
// target's link; those (target, column) blocks are skipped in J^T J with a wave-uniform mask
// (exact zeros only, so the per-entry sequence equals the dense one).

constexpr int WV_NA = 32;  // padded dof count: rows na..31 of the system are the identity

typedef float wv_v32f __attribute__((ext_vector_type(32)));

// J^T J of the two problems of a wavefront (32 lanes each) on the matrix cores.
// v_mfma_f32_32x32x1_2b_f32 multiplies two independent blocks: lane l supplies entry l % 32 of
// block l / 32's column and row vectors, so with row r of J (lane = dof) as both operands each
// instruction adds J_r^T J_r of both problems; NR of them chained in row order sum the rows.
// Block b's result D_b[i][j] lands in lane j + 32 ((i / 4) % 2), register 16 b + (i % 4) + 4 (i / 8).
// D_b is symmetric, so lane j of either half holds row j's entries at the columns i with
// (i / 4) % 2 == its half; 16 v_permlane32_swap (upper half of register m <-> lower half of
// register 16 + m) bring the other half's entries of its own block over.  Afterwards every lane
// holds its problem's row: column i in register (i % 4) + 4 (i / 8) + 16 ((i / 4) % 2).
template <int NR>
__device__ __forceinline__ void wv_jtj_mfma(const float* Jr, float* A) {
  wv_v32f acc = {};
#pragma unroll
  for (int r = 0; r < NR; ++r) acc = __builtin_amdgcn_mfma_f32_32x32x1f32(Jr[r], Jr[r], acc, 0, 0, 0);
#pragma unroll
  for (int m = 0; m < 16; ++m) {
    const auto sw = __builtin_amdgcn_permlane32_swap(__float_as_uint(acc[m]), __float_as_uint(acc[16 + m]), false,
                                                     false);
    acc[m] = __uint_as_float(sw[0]);
    acc[16 + m] = __uint_as_float(sw[1]);
  }
#pragma unroll
  for (int i = 0; i < WV_NA; ++i) A[i] = acc[(i % 4) + 4 * (i / 8) + 16 * ((i / 4) % 2)];
}

__device__ __forceinline__ float wv_readlane(float v, int lane) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), lane));
}

// value of lane k of the caller's problem (HW lanes per problem)
template <int HW>
__device__ __forceinline__ float wv_bcast(float v, int k) {
  if constexpr (HW == 64) {
    return wv_readlane(v, k);
  } else {
    const float lo = wv_readlane(v, k), hi = wv_readlane(v, k + 32);
    return (threadIdx.x & 32) ? hi : lo;
  }
}

// sum over a problem's HW lanes by a butterfly of partner exchanges that stay in the VALU (DPP
// lane permutes and the 16 / 32-lane row swaps): pairs, quads (quad_perm), 8 lanes
// (row_half_mirror: lane i <-> 7 - i), 16 (row_mirror: i <-> 15 - i), 32 (permlane16_swap),
// 64 (permlane32_swap).  Before each exchange every lane of a group holds the same partial, so
// every lane ends with the bitwise-same total (each add sees the same two operands, commuted).
__device__ __forceinline__ float wv_dpp(float v, int ctrl_case) {
  const int b = __float_as_int(v);
  switch (ctrl_case) {
    case 0: return __int_as_float(__builtin_amdgcn_mov_dpp(b, 0xB1, 0xF, 0xF, false));   // quad_perm [1,0,3,2]
    case 1: return __int_as_float(__builtin_amdgcn_mov_dpp(b, 0x4E, 0xF, 0xF, false));   // quad_perm [2,3,0,1]
    case 2: return __int_as_float(__builtin_amdgcn_mov_dpp(b, 0x141, 0xF, 0xF, false));  // row_half_mirror
    default: return __int_as_float(__builtin_amdgcn_mov_dpp(b, 0x140, 0xF, 0xF, false)); // row_mirror
  }
}
template <int HW = 64>
__device__ __forceinline__ float wv_sum(float v) {
  v += wv_dpp(v, 0);
  v += wv_dpp(v, 1);
  v += wv_dpp(v, 2);
  v += wv_dpp(v, 3);
  {
    const auto sw = __builtin_amdgcn_permlane16_swap(__float_as_uint(v), __float_as_uint(v), false, false);
    v = __uint_as_float(sw[0]) + __uint_as_float(sw[1]);
  }
  if constexpr (HW == 64) {
    const auto sw = __builtin_amdgcn_permlane32_swap(__float_as_uint(v), __float_as_uint(v), false, false);
    v = __uint_as_float(sw[0]) + __uint_as_float(sw[1]);
  }
  return v;
}

template <int HW = 64>
__device__ __forceinline__ Pose wv_shfl_pose(const Pose& P, int src) {
  Pose o;
#pragma unroll
  for (int k = 0; k < 3; ++k) o.t[k] = __shfl(P.t[k], src, HW);
#pragma unroll
  for (int k = 0; k < 4; ++k) o.q[k] = __shfl(P.q[k], src, HW);
  return o;
}

// an empty asm the optimiser cannot see through: keeps a wave-uniform mask in one SGPR pair and
// its bit tests inside the iteration loop (hoisted out, every (target, column) test became a live
// 64-bit condition and the kernel spilled hundreds of SGPRs into VGPR lanes)
__device__ __forceinline__ void wv_opaque(unsigned long long& m) { asm volatile("" : "+s"(m)); }
__device__ __forceinline__ void wv_opaque(unsigned& m) { asm volatile("" : "+s"(m)); }
template <class T>
__device__ __forceinline__ void wv_opaque_v(T& v) { asm volatile("" : "+v"(v)); }

template <int NP, int NQ>
struct WaveLDS {
  static constexpr int RP = 6 * NP + 3 * NQ > 0 ? 6 * NP + 3 * NQ : 1;
  static constexpr int JS = ((RP + 3) & ~3) + 4;  // row stride of J: 16-byte rows, 4-way at worst
  float tg[NP > 0 ? NP : 1][36];        // pose-target Jacobian blocks
  float rows[RP + 2 * WV_NA];           // residual rows, term order
  float plink[NQ > 0 ? NQ : 1][4];      // world position of each position-target link
  float4 blk[2][WV_NA];                 // Cholesky block step: row i's 4 entries of the block columns
  alignas(16) float rhs[2][WV_NA];      //                      and the right-hand side (2 buffers)
  alignas(16) float cst[2][WV_NA];      // backward block step: the running right-hand side (2 buffers)
  alignas(16) float dvs[WV_NA];         //                      and 1 / L_ii
  union {
    alignas(16) float J[WV_NA][JS];     // Jacobian columns (dof-major)
    alignas(16) float L[WV_NA][36];     // Cholesky factor rows, for the transposed backward solve
    // supernodal kernels: rows of J^T J (row stride 36); the rows past WV_NA only absorb the
    // masked reads of a slot's padding rows (a slot of s < SN_NB dofs reads SN_NB rows)
    alignas(16) float M[WV_NA + SN_NB + 1][36];
    // supernodal kernels with pose targets: J^T J of the problem's two 16-lane pair blocks (row
    // major, 16 x 16 each) and each pair's J^T r share; kept across rejected steps
    struct {
      alignas(16) float blk[2 * 16 * 16];
      float gp[WV_NA];
    } pr;
  };
};

// per-problem LDS stride in floats: 16-float aligned and an odd multiple of 16, so the two
// problems of a wavefront (HW = 32) read their broadcasts from disjoint banks
template <int NP, int NQ>
constexpr int wave_lds_stride() {
  const int f = (int)((sizeof(WaveLDS<NP, NQ>) + 63) / 64) * 16;
  return (f / 16) % 2 ? f : f + 16;
}
template <int NP, int NQ, int HW>
constexpr size_t wave_lds_bytes() {
  return (size_t)(64 / HW) * wave_lds_stride<NP, NQ>() * sizeof(float);
}

// per-lane joint data, loaded once per problem
struct WaveJoint {
  Pose P;  // parent joint transform
  float ax[3], mul, off;
  int type, parent, act, mimic;
};

// xor-butterfly exchange inside 32-lane groups (ds_swizzle, bit-mask mode): lane i reads lane i ^ X
template <int X>
__device__ __forceinline__ float sn_swz(float v) {
  return __int_as_float(__builtin_amdgcn_ds_swizzle(__float_as_int(v), 0x1F | (X << 10)));
}

// Supernodal solve of (J^T J + diag) dq = b for the block-arrow form of the elimination order
// (two problems per wavefront; sn_structure builds the form).  With the slots' diagonal blocks
// D_k, their coupling blocks E_k (root rows x slot columns) and the root block C:
//   X_k = D_k^-1 [E_k^T | b_k]                        (quad of lanes 4k..4k+3: column u each)
//   C' = C - sum_k E_k X_k[:, 0:3],  b_r' = b_r - sum_k E_k X_k[:, 3]
//   C' x_r = b_r'                                      (every lane, identically)
//   x_k = X_k[:, 3] - X_k[:, 0:3] x_r                  (summed over the quad)
// which is block elimination in the same order.  Each lane of a slot's quad factors D_k = L L^T
// itself and solves its own right-hand side, so the quad's work runs side by side; the slots'
// shares of the root system are summed by an xor butterfly over the eight quads (the same total
// on every lane), the quad exchanges are DPP moves.  A slot of s < SN_NB dofs is padded with
// identity rows; quads past the slots hold an empty slot and add zeros.  The solution goes to
// X[position]; the pivots' reciprocal square roots are the hardware's (v_rsq_f32, 1 ulp).
// In: L (the slot's block, lower), z (this lane's right-hand side: root row u of E, or b for
// u = 3), C / r (the root block and right-hand side).
__device__ __forceinline__ void sn_core(float (&L)[SN_NB][SN_NB], float (&z)[SN_NB], float (&C)[SN_NR][SN_NR],
                                        float (&r)[SN_NR], int c, int ns, int lc, int u, int r0, int nr, float* X) {
  constexpr int NB = SN_NB, NR = SN_NR;
  static_assert(SN_MAXSLOT == 8 && SN_NR == 3, "eight quads of four lanes: three root columns and b");
  // the root rows of E from the quad's lanes 0..2 (before the solve overwrites z)
  float E[NR][NB];
#pragma unroll
  for (int i = 0; i < NB; ++i) {
    E[0][i] = qbcast<0>(z[i]);
    E[1][i] = qbcast<1>(z[i]);
    E[2][i] = qbcast<2>(z[i]);
  }
  // ---- left-looking Cholesky of D (every lane of the quad), then D x = z
  float di[NB];
#pragma unroll
  for (int k = 0; k < NB; ++k) {
    float d = L[k][k];
#pragma unroll
    for (int m = 0; m < k; ++m) d -= L[k][m] * L[k][m];
    di[k] = __builtin_amdgcn_rsqf(d);
#pragma unroll
    for (int i = k + 1; i < NB; ++i) {
      float v = L[i][k];
#pragma unroll
      for (int m = 0; m < k; ++m) v -= L[i][m] * L[k][m];
      L[i][k] = v * di[k];
    }
    float v = z[k];
#pragma unroll
    for (int m = 0; m < k; ++m) v -= L[k][m] * z[m];
    z[k] = v * di[k];
  }
#pragma unroll
  for (int k = NB - 1; k >= 0; --k) {
    float v = z[k];
#pragma unroll
    for (int i = k + 1; i < NB; ++i) v -= L[i][k] * z[i];
    z[k] = v * di[k];
  }
  // ---- this lane's column of the slot's share [E D^-1 E^T | E D^-1 b], summed over the slots
  float col[NR];
#pragma unroll
  for (int w = 0; w < NR; ++w) {
    float acc = 0.f;
#pragma unroll
    for (int i = 0; i < NB; ++i) acc += E[w][i] * z[i];
    acc += sn_swz<4>(acc);
    acc += sn_swz<8>(acc);
    acc += sn_swz<16>(acc);
    col[w] = acc;
  }
  // root system from the quad's columns: C' = C - S (S[w][v] on lane v), b_r' = b_r - t (lane 3)
#pragma unroll
  for (int w = 0; w < NR; ++w) {
    C[w][0] -= qbcast<0>(col[w]);
    if (w >= 1) C[w][1] -= qbcast<1>(col[w]);
    if (w >= 2) C[w][2] -= qbcast<2>(col[w]);
    r[w] -= qbcast<3>(col[w]);
  }
  float ci[NR];
#pragma unroll
  for (int k = 0; k < NR; ++k) {
    float d = C[k][k];
#pragma unroll
    for (int m = 0; m < k; ++m) d -= C[k][m] * C[k][m];
    ci[k] = __builtin_amdgcn_rsqf(d);
#pragma unroll
    for (int i = k + 1; i < NR; ++i) {
      float v = C[i][k];
#pragma unroll
      for (int m = 0; m < k; ++m) v -= C[i][m] * C[k][m];
      C[i][k] = v * ci[k];
    }
    float v = r[k];
#pragma unroll
    for (int m = 0; m < k; ++m) v -= C[k][m] * r[m];
    r[k] = v * ci[k];
  }
#pragma unroll
  for (int k = NR - 1; k >= 0; --k) {
    float v = r[k];
#pragma unroll
    for (int i = k + 1; i < NR; ++i) v -= C[i][k] * r[i];
    r[k] = v * ci[k];
  }
  // ---- the slot's unknowns x = X[:, 3] - X[:, 0:3] x_r, summed over the quad
  const float xr_u = u == 0 ? r[0] : u == 1 ? r[1] : r[2];
#pragma unroll
  for (int i = 0; i < NB; ++i) {
    float v = u == NR ? z[i] : -(xr_u * z[i]);
    v += qperm<qp_ctrl(1, 0, 3, 2)>(v);
    v += qperm<qp_ctrl(2, 3, 0, 1)>(v);
    if (i < ns && (i & 3) == u) X[c + i] = v;
  }
  if (lc == 0) {
#pragma unroll
    for (int w = 0; w < NR; ++w)
      if (w < nr) X[r0 + w] = r[w];
  }
}

// the supernodal solve from rows of J^T J staged in LDS (row `lane` of each lane: A), the lane's
// diagonal terms + lam (dsum) and right-hand side b; returns dq of the lane's dof
template <int NP, int NQ>
__device__ __forceinline__ float sn_solve(WaveLDS<NP, NQ>* sh, TermsK* tm, int lane, int c, int ns,
                                          const float (&A)[WV_NA], float dsum, float b) {
  constexpr int LS = 36, NB = SN_NB, NR = SN_NR;
  float* const M = &sh->M[0][0];
  float* const dg = sh->cst[0];  // diagonal terms + lam per row
  float* const bv = sh->rhs[0];  // right-hand side per row
  const int r0 = tm->sn_r0, nr = tm->sn_nr;
  const int lc = lane & (WV_NA - 1), u = lane & 3;
#pragma unroll
  for (int j = 0; j < WV_NA; j += 4)
    *reinterpret_cast<float4*>(&M[lc * LS + j]) = make_float4(A[j], A[j + 1], A[j + 2], A[j + 3]);
  dg[lc] = dsum;
  bv[lc] = b;
  __syncthreads();
  float L[NB][NB], z[NB], C[NR][NR], r[NR];
  {
    const float* Dp = M + c * (LS + 1);  // D[i][j] = Dp[i * LS + j]
    const bool ru = u < NR && u < nr;
    const float* Zp = u < NR ? M + (r0 + (u < NR ? u : 0)) * LS + c : bv + c;
    const float* Cp = M + r0 * (LS + 1);  // C[w][v] = Cp[w * LS + v]
    // every load reads an address inside the LDS block (rows past the slot land in M's spare rows)
    // and is selected afterwards: conditional loads would become a branch each
#pragma unroll
    for (int i = 0; i < NB; ++i) {
#pragma unroll
      for (int j = 0; j < i; ++j) {
        const float v = Dp[i * LS + j];
        L[i][j] = i < ns ? v : 0.f;
      }
      const float d = Dp[i * LS + i] + dg[c + i], zi = Zp[i];
      L[i][i] = i < ns ? d : 1.f;
      z[i] = i < ns && (u == NR || ru) ? zi : 0.f;
    }
#pragma unroll
    for (int w = 0; w < NR; ++w) {
#pragma unroll
      for (int v = 0; v < w; ++v) {
        const float cv = Cp[w * LS + v];
        C[w][v] = w < nr ? cv : 0.f;
      }
      const float d = Cp[w * LS + w] + dg[r0 + w], rv = bv[r0 + w];
      C[w][w] = w < nr ? d : 1.f;
      r[w] = w < nr ? rv : 0.f;
    }
  }
  sn_core(L, z, C, r, c, ns, lc, u, r0, nr, sh->dvs);
  __syncthreads();
  return sh->dvs[lc];
}

// The same solve from the pair layout: J^T J in the two 16 x 16 pair blocks (slot k's block at
// its first pair p = sn_poff[k], its root coupling at its target's first root pair; the root
// block summed over the targets that move the root), the pairs' J^T r shares, and per position
// the diagonal terms + lam (dsum) and the limit / smoothness part of the right-hand side (b).
// Returns dq of the lane's dof; g receives the position's full J^T r (for the predicted
// reduction).
template <int NP, int NQ>
__device__ __forceinline__ float sn_solve_pairs(WaveLDS<NP, NQ>* sh, TermsK* tm, int lane, int c, int ns,
                                                float dsum, float b, float& g) {
  constexpr int NB = SN_NB, NR = SN_NR;
  const float* const Bk = sh->pr.blk;
  const float* const Gp = sh->pr.gp;
  float* const dg = sh->cst[0];  // diagonal terms + lam per position
  float* const bv = sh->rhs[0];  // limit / smoothness right-hand side per position
  float* const Gf = sh->cst[1];  // full J^T r per position
  const int r0 = tm->sn_r0, nr = tm->sn_nr;
  const int lc = lane & (WV_NA - 1), u = lane & 3, k = (lc >> 2) & (SN_MAXSLOT - 1);
  dg[lc] = dsum;
  bv[lc] = b;
  __syncthreads();
  const int po = ns > 0 ? tm->sn_poff[k] : -1, pq = ns > 0 ? tm->sn_proot[k] : -1;
  float L[NB][NB], z[NB], C[NR][NR], r[NR], bs[NB];
  {
    const int pb = po < 0 ? 0 : po;
    const float* Dp = Bk + (pb >> 4) * 256 + (pb & 15) * 17;  // D[i][j] = Dp[i * 16 + j]
    const int qb = pq < 0 ? 0 : pq;
    const float* Ep = Bk + (qb >> 4) * 256 + (qb & 15) * 16 + (pb & 15);  // E[w][i] = Ep[w * 16 + i]
    const bool ru = u < NR && u < nr && pq >= 0;
    // every load reads an address inside the WaveLDS block (past the pair blocks it lands in the
    // union's larger members) and is selected afterwards: conditional loads would become a branch
    // each
    const int w = u < NR ? u : 0;
#pragma unroll
    for (int i = 0; i < NB; ++i) {
      const bool live = i < ns && po >= 0;
#pragma unroll
      for (int j = 0; j < i; ++j) {
        const float v = Dp[i * 16 + j];
        L[i][j] = live ? v : 0.f;
      }
      const float dd = Dp[i * 16 + i], dgi = dg[c + i], bvi = bv[c + i], gpi = Gp[pb + i], ei = Ep[w * 16 + i];
      L[i][i] = i < ns ? (live ? dd : 0.f) + dgi : 1.f;
      bs[i] = i < ns ? bvi - (live ? gpi : 0.f) : 0.f;
      z[i] = u == NR ? bs[i] : (i < ns && ru ? ei : 0.f);
    }
#pragma unroll
    for (int w2 = 0; w2 < NR; ++w2) {
#pragma unroll
      for (int v = 0; v <= w2; ++v) C[w2][v] = 0.f;
      r[w2] = 0.f;
    }
    for (int t = 0; t < tm->pr_nrc; ++t) {
      const int rc = tm->pr_rc[t];
      const float* Cp = Bk + (rc >> 4) * 256 + (rc & 15) * 17;
#pragma unroll
      for (int w2 = 0; w2 < NR; ++w2) {
#pragma unroll
        for (int v = 0; v <= w2; ++v) {
          const float cv = Cp[w2 * 16 + v];
          C[w2][v] += w2 < nr ? cv : 0.f;
        }
        const float gv = Gp[rc + w2];
        r[w2] += w2 < nr ? gv : 0.f;
      }
    }
#pragma unroll
    for (int w2 = 0; w2 < NR; ++w2) {
      const float dgw = dg[r0 + w2], bvw = bv[r0 + w2];
      C[w2][w2] = w2 < nr ? C[w2][w2] + dgw : 1.f;
      r[w2] = w2 < nr ? bvw - r[w2] : 0.f;
    }
  }
  // the full right-hand side per position (-J^T r): the quad's lane 3 for the slot, lane 0 the root
  if (u == NR) {
#pragma unroll
    for (int i = 0; i < NB; ++i)
      if (i < ns) Gf[c + i] = -bs[i];
  }
  if (lc == 0) {
#pragma unroll
    for (int w = 0; w < NR; ++w)
      if (w < nr) Gf[r0 + w] = -r[w];
  }
  sn_core(L, z, C, r, c, ns, lc, u, r0, nr, sh->dvs);
  __syncthreads();
  g = Gf[lc];
  return sh->dvs[lc];
}

// pointer-jumping FK: lane j ends with T_world_joint(j) (identity base)
template <int HW>
__device__ __forceinline__ Pose wave_fk(const WaveJoint& jd, int nj, int nrounds, float q_mine) {
  const int lane = threadIdx.x & (HW - 1);
  // q of the dof that drives this joint
  const int src = jd.mimic >= 0 ? jd.mimic : (jd.act >= 0 ? jd.act : 0);
  const float qs = __shfl(q_mine, src, HW);
  float pos = jd.mimic >= 0 ? qs * jd.mul + jd.off : (jd.act >= 0 ? qs : 0.f);
  Pose X = identity_pose();
  if (jd.type == 2) {
    float s, c;
    rk_sincos(pos * 0.5f, &s, &c);
    X.q[0] = c; X.q[1] = jd.ax[0] * s; X.q[2] = jd.ax[1] * s; X.q[3] = jd.ax[2] * s;
  } else if (jd.type == 1) {
    X.t[0] = jd.ax[0] * pos; X.t[1] = jd.ax[1] * pos; X.t[2] = jd.ax[2] * pos;
  }
  Pose A = compose(jd.P, X);
  int ptr = lane < nj ? jd.parent : -1;
  for (int r = 0; r < nrounds; ++r) {
    const int sp = ptr < 0 ? 0 : ptr;
    const Pose Ap = wv_shfl_pose<HW>(A, sp);
    const int pp = __shfl(ptr, sp, HW);
    if (ptr >= 0) {
      A = compose(Ap, A);
      ptr = pp;
    }
  }
  return A;
}

// tile = index of the wavefront's problem group: problems tile * (64 / HW) + threadIdx.x / HW
template <int NP, int NQ, int HW, bool SN = false>
__device__ __forceinline__ void wave_solve(const Problems& Pg, int tile, const LMParams& lp, int it0, int iters,
                                           char* lds_raw) {
  static_assert(HW == 64 || HW == 32, "HW lanes per problem: 64 or 32");
  static_assert(!SN || HW == 32, "the supernodal solve runs two problems per wavefront");
  // supernodal with pose targets: the Jacobian in the (target, dof) pair layout
  constexpr bool PM = SN && NQ == 0 && NP > 0;
  constexpr int RP = 6 * NP + 3 * NQ;
  const int slot = threadIdx.x / HW;
  WaveLDS<NP, NQ>* sh =
      reinterpret_cast<WaveLDS<NP, NQ>*>(reinterpret_cast<float*>(lds_raw) + slot * wave_lds_stride<NP, NQ>());
  const int s_idx = tile * (64 / HW) + slot;
  int lane = threadIdx.x & (HW - 1);
  RobotK* rb = as_const(Pg.robot);
  TermsK* tm = as_const(Pg.terms);
  const int S = Pg.stride;
  const bool valid = s_idx < Pg.n;
  const int s = valid ? s_idx : Pg.n - 1;
  const int nj = rb->nj, na = rb->na, nrounds = rb->nrounds;
  const int n_pose = tm->n_pose, n_pos = tm->n_pos;
  const bool has_lim = tm->limit != 0, has_sm = tm->smooth != 0;
  const int Rf = 6 * n_pose, Rq = 3 * n_pos;
  const int off_l = Rf + Rq, off_s = off_l + (has_lim ? na : 0), R = off_s + (has_sm ? na : 0);

  // ---- per-lane joint data
  WaveJoint jd;
  {
    const int j = lane < nj ? lane : 0;
    const Robot* rg = Pg.robot;
#pragma unroll
    for (int k = 0; k < 3; ++k) jd.P.t[k] = rg->pT[j][k];
#pragma unroll
    for (int k = 0; k < 4; ++k) jd.P.q[k] = rg->pT[j][3 + k];
#pragma unroll
    for (int k = 0; k < 3; ++k) jd.ax[k] = rg->axis[j][k];
    jd.mul = rg->mul[j];
    jd.off = rg->off[j];
    jd.type = lane < nj ? rg->joint_type[j] : 0;
    jd.parent = lane < nj ? rg->parent[j] : -1;
    // the lane holding the dof that drives the joint (dofs sit on lanes in elimination order)
    jd.act = rg->act[j] >= 0 ? Pg.terms->wiperm[rg->act[j]] : -1;
    jd.mimic = rg->mimic[j] >= 0 ? Pg.terms->wiperm[rg->mimic[j]] : -1;
  }
  // ---- per-lane dof data (lane a < na holds dof wperm[a], the elimination order of the term
  // structure): joints driving the dof (CSR, <= 2 entries), limits, weights
  const int dof = lane < na ? Pg.terms->wperm[lane] : 0;
  int cj0 = -1, cj1 = -1;
  float cw0 = 0.f, cw1 = 0.f, lo = 0.f, hi = 0.f, wl = 0.f, ws = 0.f;
  if (lane < na) {
    const Robot* rg = Pg.robot;
    const int c0 = rg->col_start[dof], c1 = rg->col_start[dof + 1];
    if (c1 > c0) { cj0 = rg->col_joint[c0]; cw0 = rg->col_w[c0]; }
    if (c1 > c0 + 1) { cj1 = rg->col_joint[c0 + 1]; cw1 = rg->col_w[c0 + 1]; }
    lo = rg->lo[dof];
    hi = rg->hi[dof];
    wl = Pg.terms->limit_w[dof];
    ws = Pg.terms->smooth_w[dof];
  }
  // which of this dof's joints move each target's link; wave-uniform column masks per target
  unsigned in0 = 0, in1 = 0;  // bit t (pose targets), bit NP + u (position targets)
  unsigned long long cmask[NP + NQ > 0 ? NP + NQ : 1];
#pragma unroll
  for (int t = 0; t < NP + NQ; ++t) {
    const bool pose = t < NP;
    const int tt = pose ? t : t - NP;
    const bool live = pose ? tt < n_pose : tt < n_pos;
    bool b0 = false, b1 = false;
    if (live) {
      const int link = pose ? tm->pose_link[tt] : tm->pos_link[tt];
      const unsigned long long anc = rb->anc[link];
      b0 = cj0 >= 0 && ((anc >> cj0) & 1ull);
      b1 = cj1 >= 0 && ((anc >> cj1) & 1ull);
    }
    in0 |= (unsigned)b0 << t;
    in1 |= (unsigned)b1 << t;
    cmask[t] = __ballot(b0 || b1);
  }
  // pair layout: this lane's (pose target pt, dof) pair -- the dof's joints that move the
  // target's link (pj0 / pj1, -1: none) and their weights -- and its target's index in the block
  int pt = -1, ph = 0, pj0 = -1, pj1 = -1;
  float pw0 = 0.f, pw1 = 0.f;
  if constexpr (PM) {
    pt = Pg.terms->pr_t[lane];
    if (pt >= 0) {
      const Robot* rg = Pg.robot;
      const int pdof = Pg.terms->wperm[Pg.terms->pr_pos[lane]];
      ph = Pg.terms->pr_h[lane];
      const unsigned long long anc = rg->anc[Pg.terms->pose_link[pt]];
      const int c0 = rg->col_start[pdof], c1 = rg->col_start[pdof + 1];
      if (c1 > c0 && ((anc >> rg->col_joint[c0]) & 1ull)) { pj0 = rg->col_joint[c0]; pw0 = rg->col_w[c0]; }
      if (c1 > c0 + 1 && ((anc >> rg->col_joint[c0 + 1]) & 1ull)) {
        pj1 = rg->col_joint[c0 + 1];
        pw1 = rg->col_w[c0 + 1];
      }
    }
  }
  // both problems' LDS (the pair blocks' matrix-core results land on every lane of the wavefront)
  WaveLDS<NP, NQ>* const sh0 = reinterpret_cast<WaveLDS<NP, NQ>*>(lds_raw);
  WaveLDS<NP, NQ>* const sh1 =
      reinterpret_cast<WaveLDS<NP, NQ>*>(reinterpret_cast<float*>(lds_raw) + wave_lds_stride<NP, NQ>());
  (void)sh0;
  (void)sh1;

  // ---- problem data
  float q = 0.f, qprev = 0.f;
  if (lane < na) {
    q = (it0 == 0 && Pg.q_seed) ? Pg.q_seed[dof * S + s] : Pg.q[dof * S + s];
    if (has_sm) qprev = Pg.q_prev[dof * S + s];
  }
  // pose / position target owned by this lane
  Pose tgt = identity_pose();
  float ptgt[3] = {0.f, 0.f, 0.f};
  int tj = -1;  // world joint whose pose the lane's target reads (-1: base)
  if (lane < n_pose) {
    float v[7];
#pragma unroll
    for (int c = 0; c < 7; ++c) v[c] = Pg.pose_tgt[(lane * 7 + c) * S + s];
    load_pose(v, tgt);
    tj = Pg.robot->link_joint[Pg.terms->pose_link[lane]];
  } else if (lane < n_pose + n_pos) {
    const int u = lane - n_pose;
#pragma unroll
    for (int c = 0; c < 3; ++c) ptgt[c] = Pg.pos_tgt[(u * 3 + c) * S + s];
    tj = Pg.robot->link_joint[Pg.terms->pos_link[u]];
  }
  float pw[6] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
  if (lane < n_pose) {
#pragma unroll
    for (int i = 0; i < 6; ++i) pw[i] = Pg.terms->pose_w[lane][i];
  } else if (lane < n_pose + n_pos) {
#pragma unroll
    for (int i = 0; i < 3; ++i) pw[i] = Pg.terms->pos_w[lane - n_pose][i];
  }

  float lam = it0 == 0 ? lp.lm_lambda : Pg.lam[s];
  int nacc = it0 == 0 ? 0 : Pg.nacc[s];
  float cost = 0.f;
  Pose T = wave_fk<HW>(jd, nj, nrounds, q);
  // supernodal solve: the quad of lanes 4k..4k+3 factors slot k (its first position and size)
  int sn_c = 0, sn_n = 0;
  if constexpr (SN) {  // slot k = lane / 4 (a quad of lanes per slot)
    const int k = (lane >> 2) & (SN_MAXSLOT - 1);
    sn_c = Pg.terms->sn_start[k];
    sn_n = k < Pg.terms->sn_nslot ? Pg.terms->sn_size[k] : 0;
  }

  // residual rows of the lane's target at world poses Tw (all lanes hold their joint's pose);
  // JAC also writes the pose target's Jacobian blocks / the position link's point to LDS
  // (and the lane's sums of squares sp / sq; with jac false -- the proposal -- nothing goes to
  // LDS.  Both paths sum the same rows in the same order, so a cost at the proposal and the cost
  // recomputed at the accepted iterate agree bitwise).  In the pair layout (PM) a pose target's
  // log map is evaluated at the proposal only (into Lp): the iterate's (Lq) is the accepted
  // proposal's, kept in registers, so a rebuild at the iterate takes its rows and blocks from Lq
  // (the same bits as evaluating it again at the same pose).  The other layouts, whose register
  // budgets are spent, evaluate it again.
  auto target_frame = [&](const Pose& Tw) {
    const Pose Tl = wv_shfl_pose<HW>(Tw, tj < 0 ? 0 : tj);
    return tj < 0 ? identity_pose() : Tl;
  };
  TargetLog Lq, Lp;
  if constexpr (PM) {
    const Pose Tf = target_frame(T);
    if (lane < n_pose) target_log<true>(Tf, tgt, Lq);
  }
  auto target_rows = [&](const Pose& Tw, bool jac, float& sp, float& sq) {
    const Pose Tf = target_frame(Tw);
    if (lane < n_pose) {
      float r6[6];
      if constexpr (PM) {
        if (!jac) target_log<true>(Tf, tgt, Lp);
        const TargetLog& L = jac ? Lq : Lp;
#pragma unroll
        for (int i = 0; i < 6; ++i) r6[i] = pw[i] * L.e[i];
        if (jac) target_blocks<true>(tgt, Tf, L, pw, sh->tg[lane < NP ? lane : 0]);
      } else {
        pose_target(Tf, tgt, pw, r6, jac ? sh->tg[lane < NP ? lane : 0] : nullptr);
      }
      if (jac) {
#pragma unroll
        for (int i = 0; i < 6; ++i) sh->rows[6 * lane + i] = r6[i];
      }
#pragma unroll
      for (int i = 0; i < 6; ++i) sp += r6[i] * r6[i];
    } else if (lane < n_pose + n_pos) {
      const int u = lane - n_pose;
      float r3[3];
#pragma unroll
      for (int i = 0; i < 3; ++i) r3[i] = pw[i] * (ptgt[i] - Tf.t[i]);
      if (jac) {
#pragma unroll
        for (int i = 0; i < 3; ++i) {
          sh->rows[Rf + 3 * u + i] = r3[i];
          sh->plink[u < NQ ? u : 0][i] = Tf.t[i];
        }
      }
#pragma unroll
      for (int i = 0; i < 3; ++i) sq += r3[i] * r3[i];
    }
  };
  // cost = sum over terms of 0.5 * (sum of the term's squared rows) from the lanes' partial sums
  auto term_cost = [&](float sp, float sq, float sl, float ss) {
    float c = 0.f;
    if (n_pose) c = c + 0.5f * wv_sum<HW>(sp);
    if (n_pos) c = c + 0.5f * wv_sum<HW>(sq);
    if (has_lim) c = c + 0.5f * wv_sum<HW>(sl);
    if (has_sm) c = c + 0.5f * wv_sum<HW>(ss);
    return c;
  };

  // J^T J (row `lane`, without the diagonal terms), J^T r, the squared limit / smoothness
  // Jacobian diagonal and the cost at the current iterate.  A rejected step leaves the iterate,
  // hence all of these, unchanged: they are rebuilt only when some problem of the wavefront
  // moved (the others recompute the same bits).  Two problems per wavefront only.
  float Aj[WV_NA];
  float gv = 0.f, dl = 0.f, ds = 0.f;
  bool moved = true;
  for (int it = 0; it < iters; ++it) {
    // lane-derived predicates recomputed per iteration (hoisted, each is a live SGPR pair)
    wv_opaque_v(lane);
    wv_opaque_v(in0);
    wv_opaque_v(in1);
    const int gi = it0 + it;
    if (valid && Pg.q_trace && gi < Pg.trace_iters) {
      if (lane < na) Pg.q_trace[(gi * na + dof) * S + s] = q;
      if (lane == 0) Pg.lam_trace[gi * S + s] = lam;
    }
    if (HW == 64 || __ballot(moved)) {  // (one problem per wavefront: no registers to spare)
    // ---- motion subspace (lane j), the dof's joint columns (lane a)
    float S6[6];
    {
      const int j = lane < nj ? lane : 0;
      subspace_column(rb, j, T, S6);
      if (lane >= nj) S6[0] = S6[1] = S6[2] = S6[3] = S6[4] = S6[5] = 0.f;
    }
    float Sc0[6], Sc1[6];
    {
      const int j0 = PM ? pj0 : cj0, j1 = PM ? pj1 : cj1;
#pragma unroll
      for (int r = 0; r < 6; ++r) {
        Sc0[r] = __shfl(S6[r], j0 < 0 ? 0 : j0, HW);
        Sc1[r] = __shfl(S6[r], j1 < 0 ? 0 : j1, HW);
      }
    }
    // ---- residual rows and target blocks at q
    float sp = 0.f, sq = 0.f;
    target_rows(T, true, sp, sq);
    float rl = 0.f, jl = 0.f, rs = 0.f, js = 0.f;
    if (lane < na) {
      if (has_lim) {
        rl = limit_row(q, lo, hi, wl);
        jl = limit_diag(q, lo, hi, wl);
        sh->rows[off_l + lane] = rl;
      }
      if (has_sm) {
        rs = ws * (q - qprev);
        js = ws;
        sh->rows[off_s + lane] = rs;
      }
    }
    __syncthreads();
    // the cost at the iterate: after the launch's first rebuild it is already known -- an accepted
    // proposal's cost (lm_accept stores it, summed from the same rows in the same order), or the
    // unchanged cost of a problem that did not move -- so only the first evaluates it
    if (it == 0) cost = term_cost(sp, sq, rl * rl, rs * rs);
    if constexpr (PM) {
      // ---- this pair's Jacobian column (its target's six rows, -w jlog Ad(T_f^-1) S) and its
      // share of J^T r; J^T J of each block's targets on the matrix cores: per target slot h of
      // the block and row i one v_mfma_f32_16x16x1_4b_f32 (lanes of other targets give zeros),
      // so block b's 16 x 16 result is block-diagonal over its targets
      float c6[6] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
      float gp = 0.f;
      if (pt >= 0) {
        float sc[6];
#pragma unroll
        for (int r = 0; r < 6; ++r) sc[r] = (pj0 >= 0 ? Sc0[r] * pw0 : 0.f) + (pj1 >= 0 ? Sc1[r] * pw1 : 0.f);
        pose_jac_column(sh->tg[pt < NP ? pt : 0], sc, c6);
#pragma unroll
        for (int r = 0; r < 6; ++r) gp += c6[r] * sh->rows[6 * pt + r];
      }
      typedef float v16f __attribute__((ext_vector_type(16)));
      v16f acc = {};
      const int kt = tm->pr_kt;
      for (int h = 0; h < kt; ++h) {
#pragma unroll
        for (int r = 0; r < 6; ++r) {
          const float x = ph == h ? c6[r] : 0.f;
          acc = __builtin_amdgcn_mfma_f32_16x16x1f32(x, x, acc, 0, 0, 0);
        }
      }
      // block b (of problem b / 2) row j = lane % 16, columns 4 (lane / 16) .. +3 (symmetric)
      {
        const int wl = threadIdx.x & 63, j = wl & 15, g4 = 4 * (wl >> 4);
        *reinterpret_cast<float4*>(&sh0->pr.blk[j * 16 + g4]) = make_float4(acc[0], acc[1], acc[2], acc[3]);
        *reinterpret_cast<float4*>(&sh0->pr.blk[256 + j * 16 + g4]) = make_float4(acc[4], acc[5], acc[6], acc[7]);
        *reinterpret_cast<float4*>(&sh1->pr.blk[j * 16 + g4]) = make_float4(acc[8], acc[9], acc[10], acc[11]);
        *reinterpret_cast<float4*>(&sh1->pr.blk[256 + j * 16 + g4]) =
            make_float4(acc[12], acc[13], acc[14], acc[15]);
      }
      sh->pr.gp[lane & (WV_NA - 1)] = gp;
      gv = 0.f;
    } else {
    // ---- this dof's Jacobian column (pose rows, then position rows), also to LDS
    float Jc[RP > 0 ? RP : 1];
#pragma unroll
    for (int t = 0; t < NP + NQ; ++t) {
      const bool pose = t < NP;
      const int tt = pose ? t : t - NP;
      float sc[6];
      const bool m0 = (in0 >> t) & 1u, m1 = (in1 >> t) & 1u;
#pragma unroll
      for (int r = 0; r < 6; ++r) sc[r] = (m0 ? Sc0[r] * cw0 : 0.f) + (m1 ? Sc1[r] * cw1 : 0.f);
      if (pose) {
        float c6[6] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
        if (tt < n_pose && (m0 || m1)) pose_jac_column(sh->tg[tt], sc, c6);
#pragma unroll
        for (int i = 0; i < 6; ++i) Jc[6 * tt + i] = c6[i];
      } else {
        float c3[3] = {0.f, 0.f, 0.f};
        if (tt < n_pos && (m0 || m1)) {
          float w3[3] = {tm->pos_w[tt][0], tm->pos_w[tt][1], tm->pos_w[tt][2]};
          pos_jac_column(sc, sh->plink[tt], w3, c3);
        }
#pragma unroll
        for (int i = 0; i < 3; ++i) Jc[6 * NP + 3 * tt + i] = c3[i];
      }
    }
    float* const A = Aj;
    gv = 0.f;
    if constexpr (HW == 32) {
      // ---- row `lane` of J^T J (pose term, then position term) on the matrix cores; J^T r
      if constexpr (NP > 0) {
        wv_jtj_mfma<6 * NP>(Jc, A);
      } else {
#pragma unroll
        for (int j = 0; j < WV_NA; ++j) A[j] = 0.f;
      }
      if constexpr (NQ > 0) {
        float A2[WV_NA];
        wv_jtj_mfma<3 * NQ>(Jc + 6 * NP, A2);
#pragma unroll
        for (int j = 0; j < WV_NA; ++j) A[j] = A[j] + A2[j];
      }
      if (NP > 0 && n_pose) {
#pragma unroll
        for (int t = 0; t < NP; ++t)
          if (t < n_pose) {
#pragma unroll
            for (int r = 0; r < 6; ++r) gv += Jc[6 * t + r] * sh->rows[6 * t + r];
          }
      }
      if (NQ > 0 && n_pos) {
        float g2 = 0.f;
#pragma unroll
        for (int u = 0; u < NQ; ++u)
          if (u < n_pos) {
#pragma unroll
            for (int r = 0; r < 3; ++r) g2 += Jc[6 * NP + 3 * u + r] * sh->rows[Rf + 3 * u + r];
          }
        gv = gv + g2;
      }
    } else {
      if (lane < WV_NA) {
#pragma unroll
        for (int r = 0; r < RP; ++r) sh->J[lane][r] = Jc[r];
      }
      __syncthreads();
      // ---- row `lane` of J^T J and of J^T r (pose term, then position term) on the VALU.
      // Dense over the 32 columns, branch-free: a (target, column) block whose column does not move
      // the target's link holds exact zeros and adds nothing, and without branches the broadcast
      // reads of every column issue back to back (masked blocks cost a wait each and register
      // copies at every merge).
#pragma unroll
      for (int j = 0; j < WV_NA; ++j) A[j] = 0.f;
      if (NP > 0 && n_pose) {
#pragma unroll
        for (int t = 0; t < NP; ++t) {
          if (t < n_pose) {
#pragma unroll
            for (int j = 0; j < WV_NA; ++j)
#pragma unroll
              for (int r = 0; r < 6; ++r) A[j] += Jc[6 * t + r] * sh->J[j][6 * t + r];
#pragma unroll
            for (int r = 0; r < 6; ++r) gv += Jc[6 * t + r] * sh->rows[6 * t + r];
          }
        }
      }
      if (NQ > 0 && n_pos) {
        float A2[WV_NA];
#pragma unroll
        for (int j = 0; j < WV_NA; ++j) A2[j] = 0.f;
        float g2 = 0.f;
#pragma unroll
        for (int u = 0; u < NQ; ++u) {
          if (u < n_pos) {
#pragma unroll
            for (int j = 0; j < WV_NA; ++j)
#pragma unroll
              for (int r = 0; r < 3; ++r) A2[j] += Jc[6 * NP + 3 * u + r] * sh->J[j][6 * NP + 3 * u + r];
#pragma unroll
            for (int r = 0; r < 3; ++r) g2 += Jc[6 * NP + 3 * u + r] * sh->rows[Rf + 3 * u + r];
          }
        }
#pragma unroll
        for (int j = 0; j < WV_NA; ++j) A[j] = A[j] + A2[j];
        gv = gv + g2;
      }
    }
    }  // dof-lane Jacobian
    if (has_lim) gv = gv + jl * rl;
    if (has_sm) gv = gv + js * rs;
    dl = jl * jl;
    ds = js * js;
    }  // moved
    float A[WV_NA];
#pragma unroll
    for (int j = 0; j < WV_NA; ++j) A[j] = PM ? 0.f : Aj[j];
    // diagonal terms (limit, smoothness, then lam) enter row k's pivot when column k is staged.
    // Rows na..31 (padding) and their columns are exact zeros: no Jacobian column, diagonal term
    // or right-hand side reaches them.
    const bool pad = lane >= na;
    // ---- right-looking Cholesky on register rows (lane i holds row i; entries above the
    // diagonal are left as scratch and never read), forward substitution L y = -g alongside,
    // four columns B = [k0, k0 + 4) per step.  Each row stages its entries a_iB (and b_i) in LDS
    // and reads back the diagonal block D = A_BB, b_B and every trailing row's a_jB as
    // broadcasts.  Every lane factors D = L_D L_D^T itself (4 x 4), then
    //   L_iB = L_D^-1 a_iB,  y_B = L_D^-1 b_B,  b_i -= L_iB . y_B,
    //   A_ij -= (D^-1 a_iB) . a_jB = (L_D^-T L_iB) . a_jB        (rows i, j past the block)
    // (the block form of the column-by-column recurrence: one LDS round trip per four columns).
    // Two buffers: step s + 1 writes the other one, and its wait orders step s's reads before
    // step s + 2 overwrites theirs.  Padding rows k >= na stage zeros with pivot lam: they keep
    // y = 0, give dq = 0 and change no other row.  The pivots' reciprocal square roots are the
    // hardware's (v_rsq_f32, 1 ulp).
    float b = -gv, y = 0.f, dinv = 1.f;
    const bool stage = HW == 32 || lane < WV_NA;
    float dsum = lam;  // this row's diagonal terms: limit, smoothness, then lam
    if (has_sm) dsum = ds + dsum;
    if (has_lim) dsum = dl + dsum;
    float dq = 0.f;
    float gfull = gv;  // this position's full J^T r (the pair layout assembles it in the solve)
    if constexpr (PM) {
      dq = sn_solve_pairs<NP, NQ>(sh, as_const(Pg.terms), lane, sn_c, sn_n, dsum, b, gfull);
    } else if constexpr (SN) {
      dq = sn_solve<NP, NQ>(sh, as_const(Pg.terms), lane, sn_c, sn_n, A, dsum, b);
    } else {
#pragma unroll
    for (int s4 = 0; s4 < WV_NA / 4; ++s4) {
      const int k0 = 4 * s4, buf = s4 & 1;
      int ln = lane;
      wv_opaque_v(ln);  // lane tests recomputed per step (kept for all steps they spilled SGPRs)
      const int m = ln - k0;  // position of this row in the block (0..3 inside)
      float4 a4 = make_float4(A[k0], A[k0 + 1], A[k0 + 2], A[k0 + 3]);
      a4.x += m == 0 ? dsum : 0.f;
      a4.y += m == 1 ? dsum : 0.f;
      a4.z += m == 2 ? dsum : 0.f;
      a4.w += m == 3 ? dsum : 0.f;
      if (stage) {
        sh->blk[buf][lane & (WV_NA - 1)] = a4;
        sh->rhs[buf][lane & (WV_NA - 1)] = b;
      }
      __syncthreads();
      const float4 d0 = sh->blk[buf][k0], d1 = sh->blk[buf][k0 + 1], d2 = sh->blk[buf][k0 + 2],
                   d3 = sh->blk[buf][k0 + 3];
      const float4 bB = *reinterpret_cast<const float4*>(&sh->rhs[buf][k0]);
      // L_D (lower) and its reciprocal diagonal
      // (only the reciprocal square roots of the pivots are needed: one v_rsq each)
      const float i00 = __builtin_amdgcn_rsqf(d0.x);
      const float l10 = d1.x * i00, l20 = d2.x * i00, l30 = d3.x * i00;
      const float i11 = __builtin_amdgcn_rsqf(d1.y - l10 * l10);
      const float l21 = (d2.y - l20 * l10) * i11, l31 = (d3.y - l30 * l10) * i11;
      const float i22 = __builtin_amdgcn_rsqf(d2.z - l20 * l20 - l21 * l21);
      const float l32 = (d3.z - l30 * l20 - l31 * l21) * i22;
      const float i33 = __builtin_amdgcn_rsqf(d3.w - l30 * l30 - l31 * l31 - l32 * l32);
      // y_B
      const float y0 = bB.x * i00;
      const float y1 = (bB.y - l10 * y0) * i11;
      const float y2 = (bB.z - l20 * y0 - l21 * y1) * i22;
      const float y3 = (bB.w - l30 * y0 - l31 * y1 - l32 * y2) * i33;
      // this row's L_iB and D^-1 a_iB
      const float li0 = a4.x * i00;
      const float li1 = (a4.y - li0 * l10) * i11;
      const float li2 = (a4.z - li0 * l20 - li1 * l21) * i22;
      const float li3 = (a4.w - li0 * l30 - li1 * l31 - li2 * l32) * i33;
      const float g3 = li3 * i33;
      const float g2 = (li2 - l32 * g3) * i22;
      const float g1 = (li1 - l21 * g2 - l31 * g3) * i11;
      const float g0 = (li0 - l10 * g1 - l20 * g2 - l30 * g3) * i00;
      b -= li0 * y0 + li1 * y1 + li2 * y2 + li3 * y3;
      // a row m of the block solves L_D x = D e_m = L_D (row m of L_D)^T: its L_iB is its row of
      // L_D (to rounding; the entries past the diagonal are scratch).  It keeps y_m and 1 / L_mm.
      A[k0] = li0;
      A[k0 + 1] = li1;
      A[k0 + 2] = li2;
      A[k0 + 3] = li3;
      const bool s0 = m == 0, s1 = m == 1, s2 = m == 2, s3 = m == 3;
      y = s0 ? y0 : y;
      y = s1 ? y1 : y;
      y = s2 ? y2 : y;
      y = s3 ? y3 : y;
      dinv = s0 ? i00 : dinv;
      dinv = s1 ? i11 : dinv;
      dinv = s2 ? i22 : dinv;
      dinv = s3 ? i33 : dinv;
      // trailing rows in chunks of 4.  A chunk whose rows have no structural nonzero in the
      // block's columns (the term structure's fill pattern) would subtract exact zeros: it is
      // skipped (wave-uniform branch).  The updated entries are pinned at each chunk's end: left
      // free, the compiler sank every update to the step consuming its column and kept all
      // multipliers and staged entries live (hundreds of registers).
      unsigned cm = tm->wchol[s4];
      wv_opaque(cm);
#pragma unroll
      for (int jc = k0 + 4; jc < WV_NA; jc += 4) {
        if ((cm >> (jc / 4)) & 1u) {
          float4 c[4];
#pragma unroll
          for (int u = 0; u < 4; ++u) c[u] = sh->blk[buf][jc + u];
#pragma unroll
          for (int u = 0; u < 4; ++u)
            A[jc + u] = fmaf(-g3, c[u].w, fmaf(-g2, c[u].z, fmaf(-g1, c[u].y, fmaf(-g0, c[u].x, A[jc + u]))));
#pragma unroll
          for (int u = 0; u < 4; ++u) wv_opaque_v(A[jc + u]);
        }
      }
    }
    // backward substitution L^T x = y with the factor's rows transposed through LDS (the
    // Jacobian columns sharing that space were consumed before the first step's wait), four
    // unknowns per step from the last block down: the running right-hand side c is staged, every
    // lane solves L_D^T x_B = c_B (rows of L_D and 1 / L_mm read back as broadcasts), and
    // c_j -= sum_m L_{k0+m, j} x_m reads the lane's own column of the block's rows.
    const int lc = lane & (WV_NA - 1);
    if (stage) {
#pragma unroll
      for (int j = 0; j < WV_NA; ++j) sh->L[lc][j] = A[j];
      sh->dvs[lc] = dinv;
    }
    __syncthreads();
    float c = y;
#pragma unroll
    for (int s4 = WV_NA / 4 - 1; s4 >= 0; --s4) {
      const int k0 = 4 * s4, buf = s4 & 1;
      // this lane's entries of the block's rows (independent of c: issued ahead of the wait)
      const float e0 = sh->L[k0][lc], e1 = sh->L[k0 + 1][lc], e2 = sh->L[k0 + 2][lc], e3 = sh->L[k0 + 3][lc];
      if (stage) sh->cst[buf][lc] = c;
      __syncthreads();
      const float4 cB = *reinterpret_cast<const float4*>(&sh->cst[buf][k0]);
      const float4 dB = *reinterpret_cast<const float4*>(&sh->dvs[k0]);
      const float4 r1 = *reinterpret_cast<const float4*>(&sh->L[k0 + 1][k0]);
      const float4 r2 = *reinterpret_cast<const float4*>(&sh->L[k0 + 2][k0]);
      const float4 r3 = *reinterpret_cast<const float4*>(&sh->L[k0 + 3][k0]);
      const float x3 = cB.w * dB.w;
      const float x2 = (cB.z - r3.z * x3) * dB.z;
      const float x1 = (cB.y - r2.y * x2 - r3.y * x3) * dB.y;
      const float x0 = (cB.x - r1.x * x1 - r2.x * x2 - r3.x * x3) * dB.x;
      int ln = lane;
      wv_opaque_v(ln);
      const int m = ln - k0;
      dq = m == 0 ? x0 : dq;
      dq = m == 1 ? x1 : dq;
      dq = m == 2 ? x2 : dq;
      dq = m == 3 ? x3 : dq;
      c -= e0 * x0 + e1 * x1 + e2 * x2 + e3 * x3;
    }
    }  // dense
    if (pad) dq = 0.f;
    // predicted reduction 1/2 d.(lam d - g), summed over the dofs
    const float pr = wv_sum<HW>(pad ? 0.f : dq * (lam * dq - gfull));
    const float pred = 0.5f * pr;
    // ---- proposal: FK, rows and cost in registers (no LDS)
    const float qp = q + dq;
    const Pose Tp = wave_fk<HW>(jd, nj, nrounds, qp);
    float psp = 0.f, psq = 0.f, psl = 0.f, pss = 0.f;
    target_rows(Tp, false, psp, psq);
    if (lane < na) {
      if (has_lim) {
        const float r = limit_row(qp, lo, hi, wl);
        psl = r * r;
      }
      if (has_sm) {
        const float r = ws * (qp - qprev);
        pss = r * r;
      }
    }
    const float cp = term_cost(psp, psq, psl, pss);
    if (valid && Pg.trace && gi == 0 && lane == 0) Pg.trace[s] = cost;
    const bool acc = lm_accept(lp, cost, cp, pred, lam);
    moved = acc;
    if (acc) {
      ++nacc;
      q = qp;
      T = Tp;
      if constexpr (PM) Lq = Lp;
    }
    if (valid && Pg.trace && gi < Pg.trace_iters && lane == 0) {
      Pg.trace[(gi + 1) * S + s] = cost;
      Pg.acc_trace[gi * S + s] = acc;
    }
    __syncthreads();
  }
  (void)R;
  if (valid) {
    if (lane < na) Pg.q[dof * S + s] = q;
    if (lane == 0) {
      if (iters > 0) Pg.cost[s] = cost;
      Pg.lam[s] = lam;
      Pg.nacc[s] = nacc;
    }
  }
  __syncthreads();
}

template <int NP, int NQ, int HW, bool SN = false>
__global__ __launch_bounds__(64) void k_wave(Problems P, LMParams lp, int it0, int iters) {
  // wave issue priority 1 (trajectory waves stay at 0): where IK and trajectory launches share a
  // SIMD, as in the request server, the short IK solves issue first
  __builtin_amdgcn_s_setprio(1);
  __shared__ __align__(16) char lds[wave_lds_bytes<NP, NQ, HW>()];
  wave_solve<NP, NQ, HW, SN>(P, xcd_order(blockIdx.x, gridDim.x), lp, it0, iters, lds);
}

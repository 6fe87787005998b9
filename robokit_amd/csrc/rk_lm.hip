// rk_lm.hip -- batched Levenberg-Marquardt for MI355X (gfx950): kernels + C ABI (include/rkhip.h).
//
// Two execution forms of the same per-problem math:
//   * fused:  one kernel runs the whole fixed-iteration solve of a problem inside one wavefront;
//             G lanes (4..64) own a problem, its FK poses, motion subspace, Jacobian, normal
//             matrix and Cholesky factor live in LDS; nothing but inputs/outputs touches HBM.
//             The persistent variant pulls tiles of problems of several batches (different
//             robots / sizes) from a device work queue.
//   * plain:  the reference's GPU decomposition (lm_optimizer.py:834-982 fast path): per-term
//             residual / Jacobian kernels writing J to HBM, a solve kernel, integrate, FK at the
//             proposal, per-term residuals, accept -- a host loop of ~13 launches per iteration.
//
// Determinism: every reduction (each H entry, each g entry, the cost, each Cholesky entry,
// the triangular solves) is summed by ONE lane in a fixed row order, so a problem's result does
// not depend on G, on its slot in the wave, or on what else is in the batch.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstdio>
#include <cstring>
#include <deque>
#include <map>
#include <mutex>
#include <set>
#include <string>
#include <thread>
#include <array>
#include <cmath>
#include <vector>

#include "../../include/rkhip.h"
#include "rk_device.h"

#define RK_MAXT 4   // pose targets per problem
#define RK_MAXQ 12  // position targets per problem
#define RK_MAX_ROBOTS 256
#define RK_SCENE_MAXS 64  // obstacle spheres of a scene
#define RK_SCENE_MAXB 16  // obstacle boxes (axis-aligned) of a scene
#define RK_COLL_MAXS 64   // robot collision spheres of an IK term structure
#define RK_SCENE_MAXG 4   // volumes / meshes of a scene
#define RK_GEOM_MAX 64    // geometries resident in a context
// supernodal half-wave kernels: dofs per slot, root dofs, slots per problem
constexpr int SN_NB = 7, SN_NR = 3, SN_MAXSLOT = 8;

namespace rk {

constexpr int WAVE = 64;
// gfx950: 8 XCDs, each with its own L2; a launch's workgroups go to them round-robin (workgroup i
// to XCD i mod 8)
constexpr int RK_XCD = 8;
// ints per variant in a group's queue counters: the tile queue and the exit count
constexpr int RK_QSTRIDE = 2;

// Workgroup b of n renumbered so that the workgroups of one XCD (the same b mod RK_XCD) take
// consecutive numbers, a bijection of [0, n): neighbouring problems, which share cache lines of
// the problem-major arrays, then meet in one XCD's L2 instead of being fetched by every XCD.
__device__ __forceinline__ int xcd_order(int b, int n) {
  int t = b / RK_XCD;  // rank among the workgroups of b's class, after the classes before it
  for (int c = 0; c < b % RK_XCD; ++c) t += (n - c + RK_XCD - 1) / RK_XCD;
  return t;
}

// A resident scene geometry as the kernels see it (copied by value into the term tables; the
// data stays in the context).  Volume: grid values [nz][ny][nx] at org + (i, j, k) h, sampled
// trilinearly inside [lo, hi], box distance + pad outside.  Mesh: nt triangles of RK_TRI_FLOATS
// floats each (vertices a, b, c; face normal; pseudonormals of the edges ab, bc, ca and of the
// vertices a, b, c), bounding box [lo, hi].
#define RK_TRI_FLOATS 30
struct SceneGeom {
  int kind;  // 0 volume, 1 mesh
  const float* data;
  int nx, ny, nz;  // volume grid; mesh: nx = triangles
  float org[3], inv_h;
  float lo[3], hi[3];
  float pad;
};

struct Terms {
  int robot, n_pose, n_pos, limit, smooth;
  int R, Rf, Rq;
  int pose_link[RK_MAXT];
  float pose_w[RK_MAXT][6];
  int pos_link[RK_MAXQ];
  float pos_w[RK_MAXQ][3];
  float limit_w[RK_NA_MAX];
  float smooth_w[RK_NA_MAX];
  // RotationTask rows (after the position rows): n_rot + n_pose <= RK_MAXT
  int n_rot, Rr;
  int rot_link[RK_MAXT];
  float rot_w[RK_MAXT][3];
  // register wave kernels: elimination order of the dofs (lane k holds dof wperm[k]; lanes
  // >= na are padding and map to themselves), its inverse, and per 4-column block s of the
  // factorisation the 4-row chunks c > s whose rows have structural nonzeros in the block's
  // columns (bit c): the other chunks' trailing updates add exact zeros and are skipped
  int wperm[RK_NA_MAX], wiperm[RK_NA_MAX];
  unsigned wchol[RK_NA_MAX / 4];
  // block-arrow form of the same order (supernodal half-wave kernels), when it exists: the dofs
  // fall into sn_nslot slots of <= SN_NB consecutive positions (sn_start, sn_size), each coupled
  // only to itself and to the sn_nr root positions starting at sn_r0 (the last ones); else
  // sn_nslot = 0
  int sn_nslot, sn_r0, sn_nr;
  int sn_start[SN_MAXSLOT], sn_size[SN_MAXSLOT];
  // (target, dof) pair layout of the pose-target Jacobian (supernodal kernels with pose targets):
  // lane p of a problem forms the column of dof position pr_pos[p] for pose target pr_t[p] (-1:
  // idle), the pr_h[p]-th target of its 16-lane block; a target's pairs are its slot's dofs, then
  // all root dofs if it moves with any.  J^T J of each block's targets is one block-diagonal
  // 16 x 16 matrix-core product.  Per slot: the lane of its first pair (-1: no target moves it)
  // and of its target's first root pair (-1: none); per target moving the root: its first root
  // pair.  pr_ok = 0: no such layout (each slot moved by at most one target, the pairs of each
  // target in one block of 16 lanes).
  int pr_ok, pr_kt, pr_nrc;
  int pr_t[RK_NA_MAX], pr_pos[RK_NA_MAX], pr_h[RK_NA_MAX];
  int sn_poff[SN_MAXSLOT], sn_proot[SN_MAXSLOT], pr_rc[RK_MAXT];
  // SceneCollisionTask rows (after the smoothness rows), one per robot sphere: r = w pen(sd) with
  // sd = scene distance of the sphere centre - radius; spheres sorted by the joint whose pose
  // moves their link (-1: the base link) so that FK meets them in order
  int nc;
  int c_joint[RK_COLL_MAXS];
  float c_c[RK_COLL_MAXS][3], c_r[RK_COLL_MAXS], c_w[RK_COLL_MAXS];
  float c_margin;
  int c_penalty;  // 0 smooth, 1 plain, 2 surface distance
  int ns, nb;     // the scene (padded to whole groups: 8 spheres, 4 boxes)
  float obs_s[RK_SCENE_MAXS][4];
  float obs_b[RK_SCENE_MAXB][6];
  int ng;         // then the scene's volumes / meshes
  SceneGeom geom[RK_SCENE_MAXG];
};

// Dof elimination order and block fill pattern of the damped normal equations of one term
// structure (host, at batch creation).  J^T J has a nonzero (a, b) iff some target's link moves
// with both dofs (limit / smoothness rows are diagonal).  A minimum-degree order (ties: lowest
// dof) eliminates the leaves of a kinematic tree first -- for a humanoid the legs, then the arms,
// then the waist they share -- which keeps the Cholesky factor free of fill between branches;
// the symbolic factorisation in that order gives the structurally nonzero (row chunk, column
// block) pairs.  The ordering only changes which exact zeros are skipped and the rounding
// order, not the system solved.
static void sn_structure(const Robot& r, Terms& th, const std::vector<unsigned>& adj0, std::vector<unsigned>& adj,
                         const std::vector<unsigned>& dep);
static void wave_elimination(const Robot& r, Terms& th) {
  const int na = r.na;
  th.sn_nslot = th.sn_r0 = th.sn_nr = 0;
  th.pr_ok = th.pr_kt = th.pr_nrc = 0;
  for (int s = 0; s < SN_MAXSLOT; ++s) th.sn_start[s] = th.sn_size[s] = 0;
  for (int k = 0; k < RK_NA_MAX; ++k) th.wperm[k] = th.wiperm[k] = k;
  for (int s4 = 0; s4 < RK_NA_MAX / 4; ++s4) th.wchol[s4] = ~0u;
  if (na > RK_NA_MAX) return;
  std::vector<int> links;
  for (int i = 0; i < th.n_pose; ++i) links.push_back(th.pose_link[i]);
  for (int i = 0; i < th.n_pos; ++i) links.push_back(th.pos_link[i]);
  for (int i = 0; i < th.n_rot; ++i) links.push_back(th.rot_link[i]);
  std::vector<unsigned> dep(links.size(), 0u);  // bit a: the target's link moves with dof a
  for (size_t t = 0; t < links.size(); ++t)
    for (int a = 0; a < na; ++a)
      for (int c = r.col_start[a]; c < r.col_start[a + 1]; ++c)
        if ((r.anc[links[t]] >> r.col_joint[c]) & 1ull) dep[t] |= 1u << a;
  std::vector<unsigned> adj(na, 0u);
  for (unsigned m : dep)
    for (int a = 0; a < na; ++a)
      if ((m >> a) & 1u) adj[a] |= m & ~(1u << a);
  const std::vector<unsigned> adj0 = adj;
  unsigned left = na >= 32 ? ~0u : ((1u << na) - 1u);
  for (int k = 0; k < na; ++k) {
    int best = -1, bdeg = 1 << 30;
    for (int a = 0; a < na; ++a)
      if ((left >> a) & 1u) {
        const int d = __builtin_popcount(adj[a] & left);
        if (d < bdeg) { bdeg = d; best = a; }
      }
    th.wperm[k] = best;
    const unsigned nb = adj[best] & left & ~(1u << best);
    for (int a = 0; a < na; ++a)
      if ((nb >> a) & 1u) adj[a] |= nb & ~(1u << a);  // fill: the neighbours become a clique
    left &= ~(1u << best);
  }
  for (int k = 0; k < na; ++k) th.wiperm[th.wperm[k]] = k;
  sn_structure(r, th, adj0, adj, dep);
  // filled pattern in the new order: row i has a nonzero in column k < i iff dofs wperm[i],
  // wperm[k] are adjacent in the filled graph (adj now holds it)
  for (int s4 = 0; s4 < RK_NA_MAX / 4; ++s4) {
    unsigned m = 0;
    for (int c = s4 + 1; c < RK_NA_MAX / 4; ++c)
      for (int i = 4 * c; i < 4 * c + 4 && i < na; ++i)
        for (int k = 4 * s4; k < 4 * s4 + 4 && k < na; ++k)
          if ((adj[th.wperm[i]] >> th.wperm[k]) & 1u) m |= 1u << c;
    th.wchol[s4] = m;
  }
}

// symbolic elimination of the coupling graph adj (bit b of adj[a]: dofs a, b share a row block)
// in the order perm: returns the filled graph (eliminating a joins its remaining neighbours)
static std::vector<unsigned> sn_fill(std::vector<unsigned> adj, const int* perm, int na) {
  unsigned left = na >= 32 ? ~0u : ((1u << na) - 1u);
  for (int k = 0; k < na; ++k) {
    const int a = perm[k];
    const unsigned nb = adj[a] & left & ~(1u << a);
    for (int b = 0; b < na; ++b)
      if ((nb >> b) & 1u) adj[b] |= nb & ~(1u << b);
    left &= ~(1u << a);
  }
  return adj;
}

// Block-arrow form of the normal equations (host, at batch creation), in a leaves-first order:
// dofs by decreasing depth of their joints in the kinematic tree (ties: lowest dof), which for a
// tree eliminates every dof before the dofs above it, without fill.  In the filled pattern every
// nonzero (i, k), i > k, joins position k to an ancestor i in the elimination tree (parent(k) =
// the first later position adjacent to k).  The root set holds every node with two or more
// children and all their ancestors; what remains of the tree is chains, each coupled only to
// itself and to the root.  For a humanoid: the two legs and the two arms (one chain each), the
// waist as the root.  Chains are packed first-fit by decreasing size into slots of <= SN_NB dofs
// and the order becomes [slot 0][slot 1]...[root] (chains and root keep their relative order: a
// topological order of the same tree, so the fill is unchanged).  If that form exists the dofs
// take this order (wperm) and adj becomes its filled graph; without it (a chain longer than
// SN_NB, a root larger than SN_NR, more than SN_MAXSLOT slots) nothing changes and sn_nslot
// stays 0: the supernodal kernels are not offered.
static void sn_pairs(Terms& th, const std::vector<unsigned>& dep);
static void sn_structure(const Robot& r, Terms& th, const std::vector<unsigned>& adj0, std::vector<unsigned>& adj,
                         const std::vector<unsigned>& dep) {
  th.sn_nslot = th.sn_r0 = th.sn_nr = 0;
  for (int s = 0; s < SN_MAXSLOT; ++s) th.sn_start[s] = th.sn_size[s] = 0;
  const int na = r.na;
  if (na <= 0 || na > RK_NA_MAX) return;
  std::vector<int> jdepth(r.nj, 0), ddepth(na, 0);
  for (int j = 0; j < r.nj; ++j) jdepth[j] = r.parent[j] >= 0 ? jdepth[r.parent[j]] + 1 : 0;
  for (int a = 0; a < na; ++a)
    for (int c = r.col_start[a]; c < r.col_start[a + 1]; ++c) ddepth[a] = std::max(ddepth[a], jdepth[r.col_joint[c]]);
  int order[RK_NA_MAX];
  for (int a = 0; a < na; ++a) order[a] = a;
  std::stable_sort(order, order + na, [&](int a, int b) { return ddepth[a] > ddepth[b]; });
  const std::vector<unsigned> filled = sn_fill(adj0, order, na);
  std::vector<int> parent(na, -1), nchild(na, 0);
  for (int k = 0; k < na; ++k)
    for (int i = k + 1; i < na; ++i)
      if ((filled[order[i]] >> order[k]) & 1u) {
        parent[k] = i;
        ++nchild[i];
        break;
      }
  std::vector<char> root(na, 0);
  for (int k = 0; k < na; ++k)
    if (nchild[k] >= 2)
      for (int a = k; a >= 0 && !root[a]; a = parent[a]) root[a] = 1;
  // chains: each non-root position belongs to the chain of its topmost non-root ancestor
  std::vector<int> top(na), chains;
  for (int k = na - 1; k >= 0; --k) {
    if (root[k]) continue;
    top[k] = (parent[k] >= 0 && !root[parent[k]]) ? top[parent[k]] : k;
    if (top[k] == k) chains.push_back(k);
  }
  std::vector<std::vector<int>> members(na);
  for (int k = 0; k < na; ++k)
    if (!root[k]) members[top[k]].push_back(k);
  int nr = 0;
  for (int k = 0; k < na; ++k) nr += root[k];
  if (nr > SN_NR) return;
  std::sort(chains.begin(), chains.end(), [&](int a, int b) {
    return members[a].size() != members[b].size() ? members[a].size() > members[b].size() : a < b;
  });
  std::vector<std::vector<int>> slots;
  const bool merge = (int)chains.size() > SN_MAXSLOT;  // one chain per slot while they fit
  for (int c : chains) {
    if ((int)members[c].size() > SN_NB) return;
    bool put = false;
    if (merge)
      for (auto& sl : slots)
      if (sl.size() + members[c].size() <= (size_t)SN_NB) {
        sl.insert(sl.end(), members[c].begin(), members[c].end());
        put = true;
        break;
      }
    if (!put) slots.push_back(members[c]);
  }
  if ((int)slots.size() > SN_MAXSLOT) return;
  int perm[RK_NA_MAX], pos = 0;
  for (size_t s = 0; s < slots.size(); ++s) {
    th.sn_start[s] = pos;
    th.sn_size[s] = (int)slots[s].size();
    for (int k : slots[s]) perm[pos++] = order[k];
  }
  th.sn_r0 = pos;
  for (int k = 0; k < na; ++k)
    if (root[k]) perm[pos++] = order[k];
  th.sn_nr = nr;
  th.sn_nslot = (int)slots.size();
  for (int k = 0; k < na; ++k) {
    th.wperm[k] = perm[k];
    th.wiperm[perm[k]] = k;
  }
  adj = sn_fill(adj0, perm, na);
  sn_pairs(th, dep);
}

// The (target, dof) pair layout of the pose targets over the block-arrow order (host, at batch
// creation; dep[t]: bit a = target t's link moves with dof a).  Targets are packed first-fit by
// decreasing pair count into the two 16-lane blocks of a problem.
static void sn_pairs(Terms& th, const std::vector<unsigned>& dep) {
  th.pr_ok = th.pr_kt = th.pr_nrc = 0;
  for (int p = 0; p < RK_NA_MAX; ++p) th.pr_t[p] = -1, th.pr_pos[p] = 0, th.pr_h[p] = 0;
  for (int k = 0; k < SN_MAXSLOT; ++k) th.sn_poff[k] = th.sn_proot[k] = -1;
  for (int t = 0; t < RK_MAXT; ++t) th.pr_rc[t] = -1;
  if (th.sn_nslot <= 0 || th.n_pose <= 0 || th.n_pos || th.n_rot || th.nc) return;
  const int nt = th.n_pose;
  std::vector<int> tslot(nt, -1), troot(nt, 0), tn(nt, 0), owner(th.sn_nslot, -1);
  for (int t = 0; t < nt; ++t) {
    for (int k = 0; k < th.sn_nslot; ++k)
      for (int i = 0; i < th.sn_size[k]; ++i)
        if ((dep[t] >> th.wperm[th.sn_start[k] + i]) & 1u) {
          if (tslot[t] >= 0 && tslot[t] != k) return;  // a target moved by two slots
          tslot[t] = k;
        }
    for (int w = 0; w < th.sn_nr; ++w)
      if ((dep[t] >> th.wperm[th.sn_r0 + w]) & 1u) troot[t] = 1;
    if (tslot[t] >= 0) {
      if (owner[tslot[t]] >= 0) return;  // a slot moved by two targets
      owner[tslot[t]] = t;
    }
    tn[t] = (tslot[t] >= 0 ? th.sn_size[tslot[t]] : 0) + (troot[t] ? th.sn_nr : 0);
  }
  std::vector<int> order(nt);
  for (int t = 0; t < nt; ++t) order[t] = t;
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return tn[a] > tn[b]; });
  int fill[2] = {0, 0}, kt[2] = {0, 0};
  std::vector<int> tlane(nt, -1), th_(nt, 0);
  for (int t : order) {
    if (tn[t] == 0) continue;
    int b = fill[0] + tn[t] <= 16 ? 0 : fill[1] + tn[t] <= 16 ? 1 : -1;
    if (b < 0) return;
    tlane[t] = 16 * b + fill[b];
    th_[t] = kt[b]++;
    fill[b] += tn[t];
  }
  for (int t = 0; t < nt; ++t) {
    if (tlane[t] < 0) continue;
    int p = tlane[t];
    if (tslot[t] >= 0) {
      const int k = tslot[t];
      th.sn_poff[k] = p;
      for (int i = 0; i < th.sn_size[k]; ++i, ++p) {
        th.pr_t[p] = t;
        th.pr_pos[p] = th.sn_start[k] + i;
        th.pr_h[p] = th_[t];
      }
    }
    if (troot[t]) {
      if (tslot[t] >= 0) th.sn_proot[tslot[t]] = p;
      th.pr_rc[th.pr_nrc++] = p;
      for (int w = 0; w < th.sn_nr; ++w, ++p) {
        th.pr_t[p] = t;
        th.pr_pos[p] = th.sn_r0 + w;
        th.pr_h[p] = th_[t];
      }
    }
  }
  th.pr_kt = std::max(kt[0], kt[1]);
  th.pr_ok = 1;
}

struct LMParams {
  float lm_lambda, lambda_factor, lambda_min, lambda_max, rho_min, gain_ratio_epsilon;
};

// Structure-of-arrays problem buffers of one batch: element (k, problem s) at [k * stride + s].
struct Problems {
  const Robot* robot;
  const Terms* terms;
  float* q;               // [na][stride]
  const float* q_seed;    // [na][stride] start iterate for solves that restart from the seeds, or null
  const float* pose_tgt;  // [n_pose*7][stride]
  const float* pos_tgt;   // [n_pos*3][stride]
  const float* rot_tgt;   // [n_rot*4][stride] target quaternions (w, x, y, z)
  const float* q_prev;    // [na][stride]
  float* cost;            // [stride]
  float* lam;             // [stride]
  int* nacc;              // [stride]
  float* trace;           // [(trace_iters+1)][stride] or null
  int* acc_trace;         // [trace_iters][stride] or null
  float* q_trace;         // [trace_iters][na][stride] iterate q_i before step i, or null
  float* lam_trace;       // [trace_iters][stride] damping used by step i, or null
  int trace_iters;
  int n, stride;
};

template <int NJ, int NA, int NP, int NQ>
struct Cfg {
  static constexpr int nj = NJ, na = NA, np = NP, nq = NQ;
  static constexpr int rp = 6 * NP + 3 * NQ;
  static constexpr int rmax = rp + 2 * NA;
};
using CfgS = Cfg<16, 8, 2, 2>;    // serial arms: nj <= 16, na <= 8, <= 2 pose / 2 position targets
using CfgL = Cfg<32, 32, 4, 12>;  // trees: nj <= 32, na <= 32, <= 4 pose / 12 position targets

template <class C>
struct alignas(16) Shared {
  float Tw[2][C::nj][8];  // world joint poses: [cur] at q, [1-cur] at the proposal
  float S[C::nj][6];      // motion-subspace columns at q
  float J[C::rp][C::na];  // pose + position rows of the weighted Jacobian
  float r[C::rmax];       // weighted residual rows (all terms)
  float jl[C::na], js[C::na];  // diagonal Jacobian of limit / smoothness rows
  float H[C::na][C::na + 1];   // J^T J (lower), then its damped Cholesky factor in place
  float g[C::na], dq[C::na], q[C::na], qp[C::na], qprev[C::na];
  float tg[C::np][36];  // per pose target: R(T_f^-1) 9, t(T_f^-1) 3, jlog A 9, B 9, weights 6
  float tp[C::np][8];   // pose targets
  float tq[C::nq][4];   // position targets
  float cost, pred, lam, pad0;
  int cur, acc, nacc, pad1;
};

__device__ __forceinline__ void tri_index(int k, int& i, int& j) {
  int r = (int)((sqrtf(8.f * (float)k + 1.f) - 1.f) * 0.5f);
  while ((r + 1) * (r + 2) / 2 <= k) ++r;
  while (r * (r + 1) / 2 > k) --r;
  i = r;
  j = k - r * (r + 1) / 2;
}

// ---------------------------------------------------------------------------------------------
// per-problem building blocks shared by the fused and plain forms
// ---------------------------------------------------------------------------------------------

// FrameTask (log mode) quantities of one target, evaluated from the half-angle of the error
// rotation.  E = T_f^-1 T_tgt; with n = |q_xyz(E)| and r = |q(E)|:
//   half = atan2(n, q_w)             (quaternion_to_axis_angle, rotation_conversions.py:226-246)
//   sin(half) = n / r, cos(half) = q_w / r,
//   sin(a) = 2 sin(half) cos(half),  1 - cos(a) = 2 sin(half)^2       (a = |log_so3(E)|)
// which replaces the sincos(a) of se3_log_map (se3_kernels.py:230-309) and the sincos(a / 2),
// sincos(a) of se3_jlog (se3_kernels.py:358-428) by identities: one atan2, no sin / cos calls.
// The thresholds (eps = 1e-4) and the branch structure are the reference's.  With FAST, square
// roots and reciprocals are the hardware's (v_sqrt_f32, v_rsq_f32, v_rcp_f32, 1 ulp): these run on
// one lane per target and the correctly rounded sequences dominated their instruction count.  The
// trajectory kernel's goal term keeps the correctly rounded forms (its long banded solves amplify
// the difference).
template <bool FAST>
__device__ __forceinline__ float tl_sqrt(float x) { return FAST ? __builtin_amdgcn_sqrtf(x) : sqrtf(x); }
template <bool FAST>
__device__ __forceinline__ float tl_rcp(float x) { return FAST ? __builtin_amdgcn_rcpf(x) : 1.f / x; }
template <bool FAST>
__device__ __forceinline__ float tl_rsq(float x) { return FAST ? __builtin_amdgcn_rsqf(x) : 1.f / sqrtf(x); }
struct TargetLog {
  Pose Tfi;      // T_f^-1
  float e[6];    // log(E) = (v, w)
  float a, sh, ch;  // |w|, sin(half), cos(half)
};

template <bool FAST>
__device__ __forceinline__ void target_log(const Pose& Tf, const Pose& Tt, TargetLog& L) {
  const float eps = 1e-4f;
  L.Tfi = inverse(Tf);
  const Pose E = compose(L.Tfi, Tt);
  const float* q = E.q;
  const float n2 = q[1] * q[1] + q[2] * q[2] + q[3] * q[3];
  const float n = tl_sqrt<FAST>(n2);
  const float half = atan2f(n, q[0]);
  const float angle = 2.f * half;
  const float rinv = tl_rsq<FAST>(n2 + q[0] * q[0]);
  L.sh = n * rinv;
  L.ch = q[0] * rinv;
  // 1 / s with s = sin(half) / angle
  const float inv_s = fabsf(angle) < 1e-6f ? tl_rcp<FAST>(0.5f - (angle * angle) * (1.f / 48.f))
                                            : angle * tl_rcp<FAST>(L.sh);
  float w[3] = {q[1] * inv_s, q[2] * inv_s, q[3] * inv_s};
  const float a = tl_sqrt<FAST>(w[0] * w[0] + w[1] * w[1] + w[2] * w[2]);
  float c2 = 1.f / 12.f;
  if (a >= eps) {
    const float omc = 2.f * L.sh * L.sh;
    if (fabsf(omc) >= eps)
      c2 = tl_rcp<FAST>(a * a) * (1.f - (a * L.ch) * tl_rcp<FAST>(2.f * L.sh));
  }
  float v[3];
  ikk_apply(-0.5f, c2, w, E.t, v);
  L.e[0] = v[0]; L.e[1] = v[1]; L.e[2] = v[2];
  L.e[3] = w[0]; L.e[4] = w[1]; L.e[5] = w[2];
  L.a = a;
}

// Jacobian blocks of the target (frame_task.py:391-456, 557-621): jlog(F), F = T_tgt^-1 T_f = E^-1
// (its rotation vector is -w), and Ad(T_f^-1), packed as R(T_f^-1) 9, t(T_f^-1) 3, A 9, B 9, w 6
// with jlog = [[A, B], [0, A]], A = so3_jac_left_inv(w_F)^T (so3_kernels.py:13-40).
template <bool FAST>
__device__ __forceinline__ void target_blocks(const Pose& Tt, const Pose& Tf, const TargetLog& L, const float* wt,
                                              float* tg) {
  const float eps = 1e-4f;
  const float w[3] = {-L.e[3], -L.e[4], -L.e[5]};
  float p[3];  // translation of F
  {
    const Pose F = compose(inverse(Tt), Tf);
    p[0] = F.t[0]; p[1] = F.t[1]; p[2] = F.t[2];
  }
  const float t2 = w[0] * w[0] + w[1] * w[1] + w[2] * w[2];
  const bool taylor = t2 < eps;
  const float th = L.a;
  const float coeff =
      taylor ? 1.f / 12.f : (1.f - th * L.ch * tl_rcp<FAST>(2.f * L.sh)) * tl_rcp<FAST>(t2);
  const float K2[9] = {w[0] * w[0] - t2, w[0] * w[1], w[0] * w[2],
                       w[1] * w[0], w[1] * w[1] - t2, w[1] * w[2],
                       w[2] * w[0], w[2] * w[1], w[2] * w[2] - t2};
  const float K[9] = {0.f, -w[2], w[1], w[2], 0.f, -w[0], -w[1], w[0], 0.f};
  float A[9], B[9];
#pragma unroll
  for (int i = 0; i < 9; ++i) A[i] = ((i % 4) == 0 ? 1.f : 0.f) + 0.5f * K[i] + coeff * K2[i];
  const float sk[9] = {0.f, -p[2], p[1], p[2], 0.f, -p[0], -p[1], p[0], 0.f};
  if (taylor) {
#pragma unroll
    for (int i = 0; i < 9; ++i) B[i] = 0.5f * sk[i];
  } else {
    const float th_inv = tl_rcp<FAST>(th);
    const float th2_inv = th_inv * th_inv;
    const float st = 2.f * L.sh * L.ch;
    const float inv_2_2ct = tl_rcp<FAST>(4.f * L.sh * L.sh);  // 1 / (2 (1 - cos th))
    const float beta = th2_inv - st * th_inv * inv_2_2ct;
    const float bdot = -2.f * th2_inv * th2_inv + (1.f + st * th_inv) * th2_inv * inv_2_2ct;
    const float wTp = w[0] * p[0] + w[1] * p[1] + w[2] * p[2];
    float v3[3], Cm[9];
#pragma unroll
    for (int i = 0; i < 3; ++i) v3[i] = bdot * wTp * w[i] - (t2 * bdot + 2.f * beta) * p[i];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int j = 0; j < 3; ++j)
        Cm[i * 3 + j] = v3[i] * w[j] + beta * (w[i] * p[j]) + (i == j ? wTp * beta : 0.f) + 0.5f * sk[i * 3 + j];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int j = 0; j < 3; ++j)
        B[i * 3 + j] = Cm[i * 3 + 0] * A[0 * 3 + j] + Cm[i * 3 + 1] * A[1 * 3 + j] + Cm[i * 3 + 2] * A[2 * 3 + j];
  }
  const Mat3 R = quat_to_mat(L.Tfi.q);
#pragma unroll
  for (int i = 0; i < 9; ++i) tg[i] = R.m[i];
  tg[9] = L.Tfi.t[0]; tg[10] = L.Tfi.t[1]; tg[11] = L.Tfi.t[2];
#pragma unroll
  for (int i = 0; i < 9; ++i) tg[12 + i] = A[i];
#pragma unroll
  for (int i = 0; i < 9; ++i) tg[21 + i] = B[i];
#pragma unroll
  for (int i = 0; i < 6; ++i) tg[30 + i] = wt[i];
}

// Pose residual of a target r = w * log(T_f^-1 T_tgt) (frame_task.py:505-554); with tg != nullptr
// also its Jacobian blocks.
template <bool FAST = true>
__device__ __forceinline__ void pose_target(const Pose& Tf, const Pose& Tt, const float* w, float* r6,
                                            float* tg = nullptr) {
  TargetLog L;
  target_log<FAST>(Tf, Tt, L);
#pragma unroll
  for (int i = 0; i < 6; ++i) r6[i] = w[i] * L.e[i];
  if (tg) target_blocks<FAST>(Tt, Tf, L, w, tg);
}

// the Jacobian blocks alone
template <bool FAST = true>
__device__ __forceinline__ void pose_blocks(const Pose& Tf, const Pose& Tt, const float* w, float* tg) {
  TargetLog L;
  target_log<FAST>(Tf, Tt, L);
  target_blocks<FAST>(Tt, Tf, L, w, tg);
}

// spatial column of link `link` for actuated dof a: sum of S_j * map[j][a] over ancestor joints
// in ascending j (frame_task.py:431-437).  S_at(j, row) reads the subspace.
template <class SAt>
__device__ __forceinline__ void spatial_column(const Robot* rb, int link, int a, SAt S_at, float* sc) {
#pragma unroll
  for (int r = 0; r < 6; ++r) sc[r] = 0.f;
  unsigned long long anc = rb->anc[link];
  for (int c = rb->col_start[a]; c < rb->col_start[a + 1]; ++c) {
    int j = rb->col_joint[c];
    if ((anc >> j) & 1ull) {
      float w = rb->col_w[c];
#pragma unroll
      for (int r = 0; r < 6; ++r) sc[r] += S_at(j, r) * w;
    }
  }
}

// pose-row Jacobian column: -w * jlog * Ad(T_f^-1) * sc, with the structured blocks in tg
__device__ __forceinline__ void pose_jac_column(const float* tg, const float* sc, float* out6) {
  const float* R = tg;
  const float* t = tg + 9;
  const float* A = tg + 12;
  const float* B = tg + 21;
  const float* w = tg + 30;
  float Rv[3], Rw[3], txRw[3];
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    Rv[i] = R[i * 3 + 0] * sc[0] + R[i * 3 + 1] * sc[1] + R[i * 3 + 2] * sc[2];
    Rw[i] = R[i * 3 + 0] * sc[3] + R[i * 3 + 1] * sc[4] + R[i * 3 + 2] * sc[5];
  }
  cross3(t, Rw, txRw);
  float bv[3] = {Rv[0] + txRw[0], Rv[1] + txRw[1], Rv[2] + txRw[2]};
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    float top = A[i * 3 + 0] * bv[0] + A[i * 3 + 1] * bv[1] + A[i * 3 + 2] * bv[2] + B[i * 3 + 0] * Rw[0] +
                B[i * 3 + 1] * Rw[1] + B[i * 3 + 2] * Rw[2];
    float bot = A[i * 3 + 0] * Rw[0] + A[i * 3 + 1] * Rw[1] + A[i * 3 + 2] * Rw[2];
    out6[i] = -w[i] * top;
    out6[3 + i] = -w[3 + i] * bot;
  }
}

// The same map as one 6 x 6 matrix for a kernel that forms many columns of one target in one
// lane (the serial-arm thread kernel, a column per dof): with AR = A R and the rows scaled by -w,
//   top = [-w A R | -w (A [t]x R + B R)] sc,   bot = [0 | -w A R] sc_w
// M holds the three nonzero 3 x 3 blocks (Mv, Mw, Mb), built once (about 135 operations), each
// column then 27 FMAs instead of about 60.  The rounding differs from pose_jac_column's.
__device__ __forceinline__ void pose_jac_matrix(const float* tg, float (&M)[27]) {
  const float* R = tg;
  const float* t = tg + 9;
  const float* A = tg + 12;
  const float* B = tg + 21;
  const float* w = tg + 30;
  float AR[9], tR[9];
#pragma unroll
  for (int k = 0; k < 3; ++k) {  // column k of [t]x R: t x (column k of R)
    const float rk[3] = {R[k], R[3 + k], R[6 + k]};
    float c[3];
    cross3(t, rk, c);
    tR[k] = c[0]; tR[3 + k] = c[1]; tR[6 + k] = c[2];
  }
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      AR[i * 3 + k] = A[i * 3 + 0] * R[k] + A[i * 3 + 1] * R[3 + k] + A[i * 3 + 2] * R[6 + k];
      const float mw = A[i * 3 + 0] * tR[k] + A[i * 3 + 1] * tR[3 + k] + A[i * 3 + 2] * tR[6 + k] +
                       B[i * 3 + 0] * R[k] + B[i * 3 + 1] * R[3 + k] + B[i * 3 + 2] * R[6 + k];
      M[i * 3 + k] = -w[i] * AR[i * 3 + k];
      M[9 + i * 3 + k] = -w[i] * mw;
      M[18 + i * 3 + k] = -w[3 + i] * AR[i * 3 + k];
    }
}
__device__ __forceinline__ void pose_jac_apply(const float (&M)[27], const float* sc, float* out6) {
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    out6[i] = M[i * 3 + 0] * sc[0] + M[i * 3 + 1] * sc[1] + M[i * 3 + 2] * sc[2] + M[9 + i * 3 + 0] * sc[3] +
              M[9 + i * 3 + 1] * sc[4] + M[9 + i * 3 + 2] * sc[5];
    out6[3 + i] = M[18 + i * 3 + 0] * sc[3] + M[18 + i * 3 + 1] * sc[4] + M[18 + i * 3 + 2] * sc[5];
  }
}

// position-row Jacobian column: -w * (v + w x p)  (position_task.py:253-262)
__device__ __forceinline__ void pos_jac_column(const float* sc, const float* p, const float* w, float* out3) {
  float wxp[3];
  cross3(sc + 3, p, wxp);
#pragma unroll
  for (int i = 0; i < 3; ++i) out3[i] = -w[i] * (sc[i] + wxp[i]);
}

// joint-limit row, "abs" mode (position_limit.py:508-534) and its diagonal Jacobian (:537-573)
__device__ __forceinline__ float limit_row(float x, float lo, float hi, float w) {
  float eu = fmaxf(x - hi, 0.f);
  float el = fmaxf(lo - x, 0.f);
  return w * (eu + el);
}
__device__ __forceinline__ float limit_diag(float x, float lo, float hi, float w) {
  return w * (x > hi ? 1.f : (x < lo ? -1.f : 0.f));
}

// RotationTask residual direction: axis-angle of q_link * q_target^-1 (rotation_task.py:323-349;
// the product is standardised to w >= 0 by quat_mul)
__device__ __forceinline__ void rotation_target(const Pose& Tf, const float* qt, float* aa) {
  const float qinv[4] = {qt[0], -qt[1], -qt[2], -qt[3]};
  float qe[4];
  quat_mul(Tf.q, qinv, qe);
  quat_to_axis_angle(qe, aa);
}

// cost = sum over terms (in row order) of 0.5 * sum of squares of its rows
__device__ __forceinline__ float term_cost(const Terms* tm, int na, const float* r) {
  float cost = 0.f;
  int off = 0;
  const int sizes[5] = {tm->Rf, tm->Rq, tm->Rr, tm->limit ? na : 0, tm->smooth ? na : 0};
#pragma unroll
  for (int t = 0; t < 5; ++t) {
    if (sizes[t] == 0) continue;
    float s = 0.f;
    for (int k = 0; k < sizes[t]; ++k) s += r[off + k] * r[off + k];
    cost = cost + 0.5f * s;
    off += sizes[t];
  }
  return cost;
}

// LM accept / damping update (lm_accept_reject_cost_kernel, lm_optimizer.py:337-362)
__device__ __forceinline__ bool lm_accept(const LMParams& lp, float& cost, float cost_prop, float pred, float& lam) {
  bool acc = pred > 0.f;
  if (acc) {
    float rho = (cost - cost_prop) / (pred + lp.gain_ratio_epsilon);
    acc = rho >= lp.rho_min;
  }
  if (acc) {
    lam = fmaxf(lam / lp.lambda_factor, lp.lambda_min);
    cost = cost_prop;
  } else {
    lam = fminf(fmaxf(lam * lp.lambda_factor, lp.lambda_min), lp.lambda_max);
  }
  return acc;
}

// ---------------------------------------------------------------------------------------------
// fused form: G lanes per problem, everything in LDS
// ---------------------------------------------------------------------------------------------

// FK level by level (joints of one depth in parallel): T_j = T_parent o (P_j o M(q_j)).
template <int G, class C>
__device__ __forceinline__ void fk_levels(const Robot* rb, const float* q, float (*Tw)[8], int lane) {
  const int nlev = rb->nlevels;
  for (int lv = 0; lv < nlev; ++lv) {
    const int b = rb->level_start[lv], e = rb->level_start[lv + 1];
    for (int k = b + lane; k < e; k += G) {
      int j = rb->order[k];
      Pose X = joint_local(rb, j, q);
      int p = rb->parent[j];
      if (p >= 0) {
        Pose Tp;
        load_pose(Tw[p], Tp);
        X = compose(Tp, X);
      }
      store_pose(Tw[j], X);
    }
    __syncthreads();
  }
}

template <class C>
__device__ __forceinline__ Pose link_pose(const Robot* rb, float (*T)[8], int link) {
  int j = rb->link_joint[link];
  Pose P;
  if (j >= 0) {
    load_pose(T[j], P);
  } else {
    P.t[0] = P.t[1] = P.t[2] = 0.f;
    P.q[0] = 1.f; P.q[1] = P.q[2] = P.q[3] = 0.f;
  }
  return P;
}

// residual rows of all terms at (T, q); with_jac also fills tg, jl, js.  No trailing sync.
template <int G, class C>
__device__ __forceinline__ void residual_rows(const Robot* rb, const Terms* tm, Shared<C>* sh, float (*T)[8],
                                              const float* q, int lane, bool with_jac) {
  const int na = rb->na;
  for (int t = lane; t < tm->n_pose; t += G) {
    Pose Tf = link_pose<C>(rb, T, tm->pose_link[t]);
    Pose Tt;
    load_pose(sh->tp[t], Tt);
    pose_target(Tf, Tt, tm->pose_w[t], sh->r + 6 * t, with_jac ? sh->tg[t] : nullptr);
  }
  for (int u = lane; u < tm->n_pos; u += G) {
    Pose Tf = link_pose<C>(rb, T, tm->pos_link[u]);
#pragma unroll
    for (int i = 0; i < 3; ++i) sh->r[tm->Rf + 3 * u + i] = tm->pos_w[u][i] * (sh->tq[u][i] - Tf.t[i]);
  }
  for (int u = lane; u < tm->n_rot; u += G) {
    Pose Tf = link_pose<C>(rb, T, tm->rot_link[u]);
    float aa[3];
    rotation_target(Tf, sh->tp[tm->n_pose + u] + 3, aa);
#pragma unroll
    for (int i = 0; i < 3; ++i) sh->r[tm->Rf + tm->Rq + 3 * u + i] = tm->rot_w[u][i] * aa[i];
  }
  const int off_l = tm->Rf + tm->Rq + tm->Rr;
  if (tm->limit)
    for (int d = lane; d < na; d += G) {
      sh->r[off_l + d] = limit_row(q[d], rb->lo[d], rb->hi[d], tm->limit_w[d]);
      if (with_jac) sh->jl[d] = limit_diag(q[d], rb->lo[d], rb->hi[d], tm->limit_w[d]);
    }
  const int off_s = off_l + (tm->limit ? na : 0);
  if (tm->smooth)
    for (int d = lane; d < na; d += G) {
      sh->r[off_s + d] = tm->smooth_w[d] * (q[d] - sh->qprev[d]);
      if (with_jac) sh->js[d] = tm->smooth_w[d];
    }
}

// H = J^T J (lower), g = J^T r, cost (each entry summed by one lane, rows in order)
template <int G, class C>
__device__ __forceinline__ void normal_equations(const Terms* tm, Shared<C>* sh, int na, int lane) {
  const int ntri = na * (na + 1) / 2;
  const int Rf = tm->Rf, Rq = tm->Rq, Rr = tm->Rr;
  const int off_l = Rf + Rq + Rr, off_s = off_l + (tm->limit ? na : 0);
  for (int k = lane; k <= ntri + na; k += G) {
    if (k < ntri) {
      int i, j;
      tri_index(k, i, j);
      float s = 0.f;
      for (int r = 0; r < Rf; ++r) s += sh->J[r][i] * sh->J[r][j];
      float h = s;
      if (Rq) {
        float s2 = 0.f;
        for (int r = Rf; r < Rf + Rq; ++r) s2 += sh->J[r][i] * sh->J[r][j];
        h = h + s2;
      }
      if (Rr) {
        float s3 = 0.f;
        for (int r = Rf + Rq; r < Rf + Rq + Rr; ++r) s3 += sh->J[r][i] * sh->J[r][j];
        h = h + s3;
      }
      if (i == j) {
        if (tm->limit) h = h + sh->jl[i] * sh->jl[i];
        if (tm->smooth) h = h + sh->js[i] * sh->js[i];
      }
      sh->H[i][j] = h;
    } else if (k < ntri + na) {
      int i = k - ntri;
      float s = 0.f;
      for (int r = 0; r < Rf; ++r) s += sh->J[r][i] * sh->r[r];
      float gi = s;
      if (Rq) {
        float s2 = 0.f;
        for (int r = Rf; r < Rf + Rq; ++r) s2 += sh->J[r][i] * sh->r[r];
        gi = gi + s2;
      }
      if (Rr) {
        float s3 = 0.f;
        for (int r = Rf + Rq; r < Rf + Rq + Rr; ++r) s3 += sh->J[r][i] * sh->r[r];
        gi = gi + s3;
      }
      if (tm->limit) gi = gi + sh->jl[i] * sh->r[off_l + i];
      if (tm->smooth) gi = gi + sh->js[i] * sh->r[off_s + i];
      sh->g[i] = gi;
    } else {
      sh->cost = term_cost(tm, na, sh->r);
    }
  }
}

// damped Cholesky (H + lam I) L L^T, delta = -(H + lam I)^-1 g, pred = 0.5 delta.(lam delta - g)
// (_cholesky_kernel dense branch, lm_optimizer.py:56-79).  Ends with a sync.
template <int G, class C>
__device__ __forceinline__ void chol_solve(Shared<C>* sh, int na, int lane) {
  if (na <= 12) {
    if (lane == 0) {
      const float lam = sh->lam;
      for (int j = 0; j < na; ++j) {
        float s = sh->H[j][j] + lam;
        for (int k = 0; k < j; ++k) s -= sh->H[j][k] * sh->H[j][k];
        float d = sqrtf(s);
        sh->H[j][j] = d;
        for (int i = j + 1; i < na; ++i) {
          float v = sh->H[i][j];
          for (int k = 0; k < j; ++k) v -= sh->H[i][k] * sh->H[j][k];
          sh->H[i][j] = v / d;
        }
      }
    }
  } else {
    for (int d = lane; d < na; d += G) sh->H[d][d] = sh->H[d][d] + sh->lam;
    __syncthreads();
    for (int k = 0; k < na; ++k) {
      if (lane == 0) sh->H[k][k] = sqrtf(sh->H[k][k]);
      __syncthreads();
      for (int i = k + 1 + lane; i < na; i += G) sh->H[i][k] = sh->H[i][k] / sh->H[k][k];
      __syncthreads();
      const int m = na - k - 1;
      const int nt = m * (m + 1) / 2;
      for (int t = lane; t < nt; t += G) {
        int ii, jj;
        tri_index(t, ii, jj);
        const int i = k + 1 + ii, j = k + 1 + jj;
        sh->H[i][j] -= sh->H[i][k] * sh->H[j][k];
      }
      __syncthreads();
    }
  }
  if (lane == 0) {
    const float lam = sh->lam;
    for (int i = 0; i < na; ++i) {
      float v = -sh->g[i];
      for (int k = 0; k < i; ++k) v -= sh->H[i][k] * sh->dq[k];
      sh->dq[i] = v / sh->H[i][i];
    }
    for (int i = na - 1; i >= 0; --i) {
      float v = sh->dq[i];
      for (int k = i + 1; k < na; ++k) v -= sh->H[k][i] * sh->dq[k];
      sh->dq[i] = v / sh->H[i][i];
    }
    float p = 0.f;
    for (int i = 0; i < na; ++i) p += sh->dq[i] * (lam * sh->dq[i] - sh->g[i]);
    sh->pred = 0.5f * p;
  }
  __syncthreads();
}

// Jacobian columns of the pose + position rows at the current point (after residual_rows)
template <int G, class C>
__device__ __forceinline__ void jacobian_columns(const Robot* rb, const Terms* tm, Shared<C>* sh, float (*T)[8],
                                                 int lane) {
  const int na = rb->na;
  const int nitems = (tm->n_pose + tm->n_pos + tm->n_rot) * na;
  auto S_at = [&](int j, int r) { return sh->S[j][r]; };
  for (int k = lane; k < nitems; k += G) {
    const int t = k / na, a = k - t * na;
    float sc[6];
    if (t < tm->n_pose) {
      spatial_column(rb, tm->pose_link[t], a, S_at, sc);
      float c6[6];
      pose_jac_column(sh->tg[t], sc, c6);
#pragma unroll
      for (int i = 0; i < 6; ++i) sh->J[6 * t + i][a] = c6[i];
    } else if (t < tm->n_pose + tm->n_pos) {
      const int u = t - tm->n_pose;
      const int link = tm->pos_link[u];
      spatial_column(rb, link, a, S_at, sc);
      Pose Tf = link_pose<C>(rb, T, link);
      float c3[3];
      pos_jac_column(sc, Tf.t, tm->pos_w[u], c3);
#pragma unroll
      for (int i = 0; i < 3; ++i) sh->J[tm->Rf + 3 * u + i][a] = c3[i];
    } else {
      // geometric Jacobian of the rotation rows: w * angular column (rotation_task.py:352-401)
      const int u = t - tm->n_pose - tm->n_pos;
      spatial_column(rb, tm->rot_link[u], a, S_at, sc);
#pragma unroll
      for (int i = 0; i < 3; ++i) sh->J[tm->Rf + tm->Rq + 3 * u + i][a] = tm->rot_w[u][i] * sc[3 + i];
    }
  }
}

// The whole fixed-iteration solve of the problems first .. first + 64/G - 1 of P.
template <int G, class C>
__device__ __forceinline__ void solve_tile(const Problems& P, int first, const LMParams& lp, int it0, int iters,
                                           char* lds) {
  constexpr int PPW = WAVE / G;
  const int lane = threadIdx.x & (G - 1);
  const int slot = threadIdx.x / G;
  Shared<C>* sh = reinterpret_cast<Shared<C>*>(lds) + slot;
  const Robot* rb = P.robot;
  const Terms* tm = P.terms;
  const int nj = rb->nj, na = rb->na;
  const int idx = first + slot;
  const bool valid = idx < P.n;
  const int s = valid ? idx : P.n - 1;
  const int S = P.stride;
  (void)PPW;

  for (int d = lane; d < na; d += G) {
    sh->q[d] = (it0 == 0 && P.q_seed) ? P.q_seed[d * S + s] : P.q[d * S + s];
    if (tm->smooth) sh->qprev[d] = P.q_prev[d * S + s];
  }
  for (int k = lane; k < tm->n_pose * 7; k += G) sh->tp[k / 7][k % 7] = P.pose_tgt[k * S + s];
  for (int k = lane; k < tm->n_pos * 3; k += G) sh->tq[k / 3][k % 3] = P.pos_tgt[k * S + s];
  for (int k = lane; k < tm->n_rot * 4; k += G) sh->tp[tm->n_pose + k / 4][3 + k % 4] = P.rot_tgt[k * S + s];
  if (lane == 0) {
    sh->lam = it0 == 0 ? lp.lm_lambda : P.lam[s];
    sh->nacc = it0 == 0 ? 0 : P.nacc[s];
    sh->cur = 0;
  }
  __syncthreads();
  fk_levels<G, C>(rb, sh->q, sh->Tw[0], lane);

  for (int it = 0; it < iters; ++it) {
    const int cur = sh->cur;
    float(*T)[8] = sh->Tw[cur];
    if (valid && P.q_trace && it0 + it < P.trace_iters) {
      const int gi = it0 + it;
      for (int d = lane; d < na; d += G) P.q_trace[(gi * na + d) * S + s] = sh->q[d];
      if (lane == 0) P.lam_trace[gi * S + s] = sh->lam;
    }
    for (int j = lane; j < nj; j += G) {
      Pose Tj;
      load_pose(T[j], Tj);
      subspace_column(rb, j, Tj, sh->S[j]);
    }
    residual_rows<G, C>(rb, tm, sh, T, sh->q, lane, true);
    __syncthreads();
    jacobian_columns<G, C>(rb, tm, sh, T, lane);
    __syncthreads();
    normal_equations<G, C>(tm, sh, na, lane);
    __syncthreads();
    chol_solve<G, C>(sh, na, lane);
    for (int d = lane; d < na; d += G) sh->qp[d] = sh->q[d] + sh->dq[d];
    __syncthreads();
    fk_levels<G, C>(rb, sh->qp, sh->Tw[1 - cur], lane);
    residual_rows<G, C>(rb, tm, sh, sh->Tw[1 - cur], sh->qp, lane, false);
    __syncthreads();
    if (lane == 0) {
      const float cp = term_cost(tm, na, sh->r);
      float cost = sh->cost, lam = sh->lam;
      const int gi = it0 + it;
      if (valid && P.trace && gi == 0) P.trace[s] = cost;
      const bool acc = lm_accept(lp, cost, cp, sh->pred, lam);
      sh->cost = cost;
      sh->lam = lam;
      sh->acc = acc;
      if (acc) {
        sh->cur = 1 - cur;
        sh->nacc += 1;
      }
      if (valid && P.trace && gi < P.trace_iters) {
        P.trace[(gi + 1) * S + s] = cost;
        P.acc_trace[gi * S + s] = acc;
      }
    }
    __syncthreads();
    if (sh->acc)
      for (int d = lane; d < na; d += G) sh->q[d] = sh->qp[d];
    __syncthreads();
  }
  if (valid) {
    for (int d = lane; d < na; d += G) P.q[d * S + s] = sh->q[d];
    if (lane == 0) {
      if (iters > 0) P.cost[s] = sh->cost;
      P.lam[s] = sh->lam;
      P.nacc[s] = sh->nacc;
    }
  }
  __syncthreads();
}

template <int G, class C>
__global__ __launch_bounds__(64) void k_fused(Problems P, LMParams lp, int it0, int iters) {
  // wave issue priority 1 (trajectory waves stay at 0): where IK and trajectory launches share a
  // SIMD, as in the request server, the short IK solves issue first
  __builtin_amdgcn_s_setprio(1);
  extern __shared__ __align__(16) char lds[];
  solve_tile<G, C>(P, xcd_order(blockIdx.x, gridDim.x) * (WAVE / G), lp, it0, iters, lds);
}

#include "rk_scene.h"
#include "rk_thread.h"
#include "rk_quad.h"
#include "rk_wave.h"
#include "rk_traj.h"

// kernel variants: (threads per problem, data placement)
//   V_S*/V_L*: G lanes per problem, LDS-staged (small / large LDS configuration)
//   V_T*:      one lane per problem, register-resident, serial arms with NA dofs
//   V_W*:      one wavefront per problem, register-resident rows (trees, <= 64 joints, <= 32 dofs)
//   V_H*:      two problems per wavefront, 32 lanes each, otherwise as V_W* (trees, <= 32 joints)
//   V_P*:      as V_T*, with a position + rotation target pair instead of a pose target
//   V_TC*/V_PC*: as V_T* / V_P* (6 or 7 dofs), with SceneCollisionTask rows
enum Variant {
  V_S4 = 0, V_S8, V_S16, V_L16, V_L32, V_L64, V_T5, V_T6, V_T7, V_T8, V_W40, V_W012, V_H40, V_H012,
  V_P5, V_P6, V_P7, V_P8, V_TC6, V_TC7, V_PC6, V_PC7, V_HS40, V_HS012, V_Q6, V_Q7, NVARIANT
};
__host__ __device__ constexpr bool variant_quad(int v) { return v == V_Q6 || v == V_Q7; }
__host__ __device__ constexpr bool variant_wave(int v) { return v == V_W40 || v == V_W012; }
__host__ __device__ constexpr bool variant_super(int v) { return v == V_HS40 || v == V_HS012; }
__host__ __device__ constexpr bool variant_half(int v) { return v == V_H40 || v == V_H012 || variant_super(v); }
__host__ __device__ constexpr bool variant_thread(int v) {
  return (v >= V_T5 && v <= V_T8) || (v >= V_P5 && v <= V_P8) || (v >= V_TC6 && v <= V_PC7);
}
__host__ __device__ constexpr int variant_G(int v) {
  return variant_thread(v) ? 1
         : variant_quad(v) ? 4
         : variant_wave(v) ? 64
         : variant_half(v) ? 32
         : v == V_S4       ? 4
         : v == V_S8       ? 8
         : v == V_S16      ? 16
         : v == V_L16      ? 16
         : v == V_L32      ? 32
                           : 64;
}
__host__ __device__ constexpr bool variant_small(int v) { return v <= V_S16; }
// problems per workgroup (one wavefront)
__host__ __device__ constexpr int variant_ppw(int v) { return variant_thread(v) ? WAVE : WAVE / variant_G(v); }
// LDS bytes per workgroup of a variant
__host__ __device__ constexpr size_t variant_lds(int v) {
  return variant_thread(v) || variant_quad(v) ? 0
         : v == V_W40      ? wave_lds_bytes<4, 0, 64>()
         : v == V_W012     ? wave_lds_bytes<0, 12, 64>()
         : v == V_H40 || v == V_HS40   ? wave_lds_bytes<4, 0, 32>()
         : v == V_H012 || v == V_HS012 ? wave_lds_bytes<0, 12, 32>()
         : variant_small(v) ? (WAVE / variant_G(v)) * sizeof(Shared<CfgS>)
                            : (WAVE / variant_G(v)) * sizeof(Shared<CfgL>);
}

struct Bucket {
  Problems P;
  int variant;
  int tile_base;  // first global tile of this bucket
  int tiles;
  int pad;
};

// one tile (one wavefront's worth of problems) of bucket bk
template <int V>
__device__ __forceinline__ void run_tile(const Bucket* bk, int local, const LMParams& lp, int it0, int iters,
                                         char* lds) {
  const Problems P = bk->P;
  if constexpr (V == V_S4) solve_tile<4, CfgS>(P, local * 16, lp, it0, iters, lds);
  else if constexpr (V == V_S8) solve_tile<8, CfgS>(P, local * 8, lp, it0, iters, lds);
  else if constexpr (V == V_S16) solve_tile<16, CfgS>(P, local * 4, lp, it0, iters, lds);
  else if constexpr (V == V_L16) solve_tile<16, CfgL>(P, local * 4, lp, it0, iters, lds);
  else if constexpr (V == V_L32) solve_tile<32, CfgL>(P, local * 2, lp, it0, iters, lds);
  else if constexpr (V == V_L64) solve_tile<64, CfgL>(P, local, lp, it0, iters, lds);
  else if constexpr (V == V_T5) thread_solve<5, 7>(P, local * WAVE + threadIdx.x, lp, it0, iters);
  else if constexpr (V == V_T6) thread_solve<6, 8>(P, local * WAVE + threadIdx.x, lp, it0, iters);
  else if constexpr (V == V_T7) thread_solve<7, 9>(P, local * WAVE + threadIdx.x, lp, it0, iters);
  else if constexpr (V == V_T8) thread_solve<8, 10>(P, local * WAVE + threadIdx.x, lp, it0, iters);
  else if constexpr (V == V_P5) thread_solve<5, 7, true>(P, local * WAVE + threadIdx.x, lp, it0, iters);
  else if constexpr (V == V_P6) thread_solve<6, 8, true>(P, local * WAVE + threadIdx.x, lp, it0, iters);
  else if constexpr (V == V_P7) thread_solve<7, 9, true>(P, local * WAVE + threadIdx.x, lp, it0, iters);
  else if constexpr (V == V_P8) thread_solve<8, 10, true>(P, local * WAVE + threadIdx.x, lp, it0, iters);
  else if constexpr (V == V_TC6) thread_solve<6, 8, false, true>(P, local * WAVE + threadIdx.x, lp, it0, iters);
  else if constexpr (V == V_TC7) thread_solve<7, 9, false, true>(P, local * WAVE + threadIdx.x, lp, it0, iters);
  else if constexpr (V == V_PC6) thread_solve<6, 8, true, true>(P, local * WAVE + threadIdx.x, lp, it0, iters);
  else if constexpr (V == V_PC7) thread_solve<7, 9, true, true>(P, local * WAVE + threadIdx.x, lp, it0, iters);
  else if constexpr (V == V_W40) wave_solve<4, 0, 64>(P, local, lp, it0, iters, lds);
  else if constexpr (V == V_W012) wave_solve<0, 12, 64>(P, local, lp, it0, iters, lds);
  else if constexpr (V == V_H40) wave_solve<4, 0, 32>(P, local, lp, it0, iters, lds);
  else if constexpr (V == V_Q6) quad_solve<6, 8>(P, local, lp, it0, iters);
  else if constexpr (V == V_Q7) quad_solve<7, 9>(P, local, lp, it0, iters);
  else if constexpr (V == V_HS40) wave_solve<4, 0, 32, true>(P, local, lp, it0, iters, lds);
  else if constexpr (V == V_HS012) wave_solve<0, 12, 32, true>(P, local, lp, it0, iters, lds);
  else wave_solve<0, 12, 32>(P, local, lp, it0, iters, lds);
}

// Persistent uneven-batch kernel of one variant: each workgroup (one wavefront) takes a first
// tile, then pulls tiles of the variant's buckets from its queue until the queue is empty.  A
// group launches one such kernel per variant it holds, concurrently (forked streams, one graph);
// a kernel per variant gives each its own register allocation instead of the worst case of all.
// First tiles are XCD-aware: the nblocks workgroups take tiles [0, nblocks), and the workgroups
// that share an XCD (relative index r with the same r mod RK_XCD) take consecutive tiles, so
// neighbouring problems, which share cache lines of the problem-major arrays, meet in one XCD's
// L2 instead of being fetched by every XCD (the humanoid kernel's HBM fetch 2.9 -> 1.2 MB per
// solve).  The queue hands out tiles from nblocks on.
// counter[0] is the tile queue, counter[1] counts workgroups that left it; the last one out
// (every other workgroup has taken its final tile) puts both back to zero for the next launch,
// so no reset kernel precedes a launch.  Both start at zero when the group is created.
// register budget of a variant's persistent kernel: waves per SIMD the allocation must allow
__host__ __device__ constexpr int variant_waves_per_eu(int v) {
  return variant_wave(v) || variant_half(v) || variant_quad(v) ? 2 : 1;
}

// the tile loop of workgroup r of the nblocks workgroups that serve variant V
template <int V>
__device__ __forceinline__ void persist_loop(const Bucket* buckets, int nb, int total_tiles, int* counter, int nblocks,
                                             int r, const LMParams& lp, int it0, int iters, char* lds, int& tile_s) {
  int tile = xcd_order(r, nblocks);  // the first tile
  for (;;) {
    if (tile >= total_tiles) break;
    int b = 0;
    while (b + 1 < nb && tile >= buckets[b + 1].tile_base) ++b;
    run_tile<V>(buckets + b, tile - buckets[b].tile_base, lp, it0, iters, lds);
    if (threadIdx.x == 0) tile_s = nblocks + atomicAdd(counter, 1);
    __syncthreads();
    tile = tile_s;
    __syncthreads();
  }
  if (threadIdx.x == 0 && atomicAdd(counter + 1, 1) == nblocks - 1) {
    atomicExch(counter, 0);
    atomicExch(counter + 1, 0);
  }
}

template <int V>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(variant_waves_per_eu(V)))) void k_persistent(const Bucket* __restrict__ buckets, int nb, int total_tiles,
                                                   int* counter, LMParams lp, int it0, int iters) {
  // static LDS of the variant's size: its offsets fold into the LDS instructions (a dynamic
  // base made every address a live scalar)
  __shared__ __align__(16) char lds[variant_lds(V) > 0 ? variant_lds(V) : 16];
  __shared__ int tile_s;
  persist_loop<V>(buckets, nb, total_tiles, counter, (int)gridDim.x, (int)blockIdx.x, lp, it0, iters, lds, tile_s);
}

// The register-resident variants (thread-per-problem up to 7 dofs, wave and half-wave) fit one
// kernel's budget (<= 256 registers, <= 15.2 KB LDS, 2 wavefronts per SIMD): a group made of
// them runs as ONE launch, consecutive block ranges serving the variants' queues, instead of a
// kernel per variant on forked streams (the fork, the join and the streams' start skew cost
// ~50 us per solve on the uneven benchmark).
__host__ __device__ constexpr bool variant_mixable(int v) {
  return (v >= V_T5 && v <= V_T7) || variant_wave(v) || variant_half(v) || variant_quad(v);
}
constexpr size_t mix_lds() {
  size_t m = 16;
  for (int v = 0; v < NVARIANT; ++v)
    if (variant_mixable(v) && variant_lds(v) > m) m = variant_lds(v);
  return m;
}
struct MixArgs {
  int nslots;
  int variant[NVARIANT], block0[NVARIANT], nblocks[NVARIANT], first[NVARIANT], nb[NVARIANT], tiles[NVARIANT],
      ctr[NVARIANT];
};

__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(2))) void k_persistent_reg(const Bucket* __restrict__ buckets, MixArgs ma, int* counter,
                                                                                        LMParams lp, int it0, int iters) {
  __shared__ __align__(16) char lds[mix_lds()];
  __shared__ int tile_s;
  int k = 0;
  while (k + 1 < ma.nslots && (int)blockIdx.x >= ma.block0[k + 1]) ++k;
  const Bucket* bk = buckets + ma.first[k];
  int* const ctr = counter + RK_QSTRIDE * ma.ctr[k];
  const int nb = ma.nb[k], tiles = ma.tiles[k], nbl = ma.nblocks[k], r = (int)blockIdx.x - ma.block0[k];
  switch (ma.variant[k]) {
    case V_T5: persist_loop<V_T5>(bk, nb, tiles, ctr, nbl, r, lp, it0, iters, lds, tile_s); break;
    case V_T6: persist_loop<V_T6>(bk, nb, tiles, ctr, nbl, r, lp, it0, iters, lds, tile_s); break;
    case V_T7: persist_loop<V_T7>(bk, nb, tiles, ctr, nbl, r, lp, it0, iters, lds, tile_s); break;
    case V_W40: persist_loop<V_W40>(bk, nb, tiles, ctr, nbl, r, lp, it0, iters, lds, tile_s); break;
    case V_W012: persist_loop<V_W012>(bk, nb, tiles, ctr, nbl, r, lp, it0, iters, lds, tile_s); break;
    case V_H40: persist_loop<V_H40>(bk, nb, tiles, ctr, nbl, r, lp, it0, iters, lds, tile_s); break;
    case V_HS40: persist_loop<V_HS40>(bk, nb, tiles, ctr, nbl, r, lp, it0, iters, lds, tile_s); break;
    case V_Q6: persist_loop<V_Q6>(bk, nb, tiles, ctr, nbl, r, lp, it0, iters, lds, tile_s); break;
    case V_Q7: persist_loop<V_Q7>(bk, nb, tiles, ctr, nbl, r, lp, it0, iters, lds, tile_s); break;
    case V_HS012: persist_loop<V_HS012>(bk, nb, tiles, ctr, nbl, r, lp, it0, iters, lds, tile_s); break;
    default: persist_loop<V_H012>(bk, nb, tiles, ctr, nbl, r, lp, it0, iters, lds, tile_s); break;
  }
}

// ---------------------------------------------------------------------------------------------
// plain form: the reference's decomposition (one thread per element, J in HBM)
// ---------------------------------------------------------------------------------------------
struct PlainBufs {
  float* Tj;    // [nj*7][S] world joint poses
  float* Sw;    // [nj*6][S] motion subspace
  float* J;     // [R*na][S] weighted Jacobian
  float* r;     // [R][S] residual at q
  float* rp;    // [R][S] residual at the proposal
  float* qp;    // [na][S]
  float* dq;    // [na][S]
  float* pred;  // [S]
  int* acc;     // [S]
};

__device__ __forceinline__ Pose plain_pose(const float* Tj, int j, int S, int b) {
  Pose P;
  if (j < 0) {
    P.t[0] = P.t[1] = P.t[2] = 0.f;
    P.q[0] = 1.f; P.q[1] = P.q[2] = P.q[3] = 0.f;
    return P;
  }
  P.t[0] = Tj[(j * 7 + 0) * S + b]; P.t[1] = Tj[(j * 7 + 1) * S + b]; P.t[2] = Tj[(j * 7 + 2) * S + b];
  P.q[0] = Tj[(j * 7 + 3) * S + b]; P.q[1] = Tj[(j * 7 + 4) * S + b];
  P.q[2] = Tj[(j * 7 + 5) * S + b]; P.q[3] = Tj[(j * 7 + 6) * S + b];
  return P;
}

// FK (+ motion subspace) one thread per problem, joints in depth order
__global__ void k_plain_fk(Problems P, const float* __restrict__ qsrc, float* Tj, float* Sw) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= P.n) return;
  const Robot* rb = P.robot;
  const int S = P.stride, na = rb->na;
  float q[RK_NA_MAX];
  for (int d = 0; d < na; ++d) q[d] = qsrc[d * S + b];
  for (int k = 0; k < rb->nj; ++k) {
    const int j = rb->order[k];
    Pose X = joint_local(rb, j, q);
    const int p = rb->parent[j];
    if (p >= 0) X = compose(plain_pose(Tj, p, S, b), X);
    const float v[7] = {X.t[0], X.t[1], X.t[2], X.q[0], X.q[1], X.q[2], X.q[3]};
    for (int c = 0; c < 7; ++c) Tj[(j * 7 + c) * S + b] = v[c];
    if (Sw) {
      float s6[6];
      subspace_column(rb, j, X, s6);
      for (int r = 0; r < 6; ++r) Sw[(j * 6 + r) * S + b] = s6[r];
    }
  }
}

__global__ void k_fill(float* p, long n, float v) {
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

// pose residual rows, thread per (problem, target)
__global__ void k_plain_pose_res(Problems P, const float* Tj, float* r) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  const int t = blockIdx.y;
  if (b >= P.n) return;
  const Terms* tm = P.terms;
  const int S = P.stride;
  Pose Tf = plain_pose(Tj, P.robot->link_joint[tm->pose_link[t]], S, b);
  Pose Tt;
  float tv[7];
  for (int c = 0; c < 7; ++c) tv[c] = P.pose_tgt[(t * 7 + c) * S + b];
  load_pose(tv, Tt);
  float r6[6];
  pose_target(Tf, Tt, tm->pose_w[t], r6);
  for (int i = 0; i < 6; ++i) r[(6 * t + i) * S + b] = r6[i];
}

// pose Jacobian, thread per (problem, actuated dof), loops targets (frame_task.py:557-621)
__global__ void k_plain_pose_jac(Problems P, const float* Tj, const float* Sw, float* J) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  const int a = blockIdx.y;
  if (b >= P.n) return;
  const Robot* rb = P.robot;
  const Terms* tm = P.terms;
  const int S = P.stride, na = rb->na;
  auto S_at = [&](int j, int r) { return Sw[(j * 6 + r) * S + b]; };
  for (int t = 0; t < tm->n_pose; ++t) {
    Pose Tf = plain_pose(Tj, rb->link_joint[tm->pose_link[t]], S, b);
    float tv[7];
    for (int c = 0; c < 7; ++c) tv[c] = P.pose_tgt[(t * 7 + c) * S + b];
    Pose Tt;
    load_pose(tv, Tt);
    float tg[36];
    pose_blocks(Tf, Tt, tm->pose_w[t], tg);
    float sc[6], c6[6];
    spatial_column(rb, tm->pose_link[t], a, S_at, sc);
    pose_jac_column(tg, sc, c6);
    for (int i = 0; i < 6; ++i) J[((6 * t + i) * na + a) * S + b] = c6[i];
  }
}

__global__ void k_plain_pos_res(Problems P, const float* Tj, float* r) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  const int u = blockIdx.y;
  if (b >= P.n) return;
  const Terms* tm = P.terms;
  const int S = P.stride;
  Pose Tf = plain_pose(Tj, P.robot->link_joint[tm->pos_link[u]], S, b);
  for (int i = 0; i < 3; ++i)
    r[(tm->Rf + 3 * u + i) * S + b] = tm->pos_w[u][i] * (P.pos_tgt[(u * 3 + i) * S + b] - Tf.t[i]);
}

__global__ void k_plain_pos_jac(Problems P, const float* Tj, const float* Sw, float* J) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  const int a = blockIdx.y;
  if (b >= P.n) return;
  const Robot* rb = P.robot;
  const Terms* tm = P.terms;
  const int S = P.stride, na = rb->na;
  auto S_at = [&](int j, int r) { return Sw[(j * 6 + r) * S + b]; };
  for (int u = 0; u < tm->n_pos; ++u) {
    Pose Tf = plain_pose(Tj, rb->link_joint[tm->pos_link[u]], S, b);
    float sc[6], c3[3];
    spatial_column(rb, tm->pos_link[u], a, S_at, sc);
    pos_jac_column(sc, Tf.t, tm->pos_w[u], c3);
    for (int i = 0; i < 3; ++i) J[((tm->Rf + 3 * u + i) * na + a) * S + b] = c3[i];
  }
}

// RotationTask rows r = w axis_angle(q_link q_tgt^-1) (rotation_task.py:323-349), thread per
// (problem, target)
__global__ void k_plain_rot_res(Problems P, const float* Tj, float* r) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  const int u = blockIdx.y;
  if (b >= P.n) return;
  const Terms* tm = P.terms;
  const int S = P.stride;
  const Pose Tf = plain_pose(Tj, P.robot->link_joint[tm->rot_link[u]], S, b);
  float qt[4], aa[3];
  for (int c = 0; c < 4; ++c) qt[c] = P.rot_tgt[(u * 4 + c) * S + b];
  rotation_target(Tf, qt, aa);
  for (int i = 0; i < 3; ++i) r[(tm->Rf + tm->Rq + 3 * u + i) * S + b] = tm->rot_w[u][i] * aa[i];
}

// RotationTask Jacobian: w * the angular part of the link's spatial column (rotation_task.py:352-401),
// thread per (problem, dof)
__global__ void k_plain_rot_jac(Problems P, const float* Sw, float* J) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  const int a = blockIdx.y;
  if (b >= P.n) return;
  const Robot* rb = P.robot;
  const Terms* tm = P.terms;
  const int S = P.stride, na = rb->na;
  auto S_at = [&](int j, int r) { return Sw[(j * 6 + r) * S + b]; };
  for (int u = 0; u < tm->n_rot; ++u) {
    float sc[6];
    spatial_column(rb, tm->rot_link[u], a, S_at, sc);
    for (int i = 0; i < 3; ++i) J[((tm->Rf + tm->Rq + 3 * u + i) * na + a) * S + b] = tm->rot_w[u][i] * sc[3 + i];
  }
}

// SceneCollisionTask rows (residuals only: zero-iteration scoring), thread per problem
__global__ void k_plain_coll_res(Problems P, const float* Tj, float* r) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= P.n) return;
  const Terms* tm = P.terms;
  const int S = P.stride, na = P.robot->na;
  const int off = tm->Rf + tm->Rq + tm->Rr + (tm->limit ? na : 0) + (tm->smooth ? na : 0);
  for (int k = 0; k < tm->nc; ++k) {
    const Pose T = plain_pose(Tj, tm->c_joint[k], S, b);
    float c[3];
    quat_apply(T.q, tm->c_c[k], c);
    for (int i = 0; i < 3; ++i) c[i] += T.t[i];
    int grp;
    float pen, dpen;
    coll_penalty(scene_sdf(tm, c, grp) - tm->c_r[k], tm->c_margin, tm->c_penalty, pen, dpen);
    r[(off + k) * S + b] = tm->c_w[k] * pen;
  }
}

// limit + smoothness rows (and their diagonal Jacobian), thread per (problem, dof)
__global__ void k_plain_diag_rows(Problems P, const float* qsrc, float* r, float* J) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  const int d = blockIdx.y;
  if (b >= P.n) return;
  const Robot* rb = P.robot;
  const Terms* tm = P.terms;
  const int S = P.stride, na = rb->na;
  const float x = qsrc[d * S + b];
  int off = tm->Rf + tm->Rq + tm->Rr;
  if (tm->limit) {
    r[(off + d) * S + b] = limit_row(x, rb->lo[d], rb->hi[d], tm->limit_w[d]);
    if (J) J[((off + d) * na + d) * S + b] = limit_diag(x, rb->lo[d], rb->hi[d], tm->limit_w[d]);
    off += na;
  }
  if (tm->smooth) {
    r[(off + d) * S + b] = tm->smooth_w[d] * (x - P.q_prev[d * S + b]);
    if (J) J[((off + d) * na + d) * S + b] = tm->smooth_w[d];
  }
}

// the solve kernel of the plain form (the reference's _solve_kernel, lm_optimizer.py:145-189):
// J, r from HBM -> H, g, cost -> damped Cholesky -> delta, pred.  One problem per G lanes.
template <int G, class C>
__global__ __launch_bounds__(64) void k_plain_solve(Problems P, PlainBufs B) {
  extern __shared__ __align__(16) char lds[];
  const int lane = threadIdx.x & (G - 1);
  const int slot = threadIdx.x / G;
  Shared<C>* sh = reinterpret_cast<Shared<C>*>(lds) + slot;
  const Robot* rb = P.robot;
  const Terms* tm = P.terms;
  const int na = rb->na, S = P.stride;
  const int idx = blockIdx.x * (WAVE / G) + slot;
  const bool valid = idx < P.n;
  const int s = valid ? idx : P.n - 1;
  const int Rp = tm->Rf + tm->Rq + tm->Rr;
  const int R = tm->R;
  for (int k = lane; k < Rp * na; k += G) sh->J[k / na][k % na] = B.J[k * S + s];
  for (int k = lane; k < R; k += G) sh->r[k] = B.r[k * S + s];
  const int off_l = Rp, off_s = Rp + (tm->limit ? na : 0);
  for (int d = lane; d < na; d += G) {
    if (tm->limit) sh->jl[d] = B.J[((off_l + d) * na + d) * S + s];
    if (tm->smooth) sh->js[d] = B.J[((off_s + d) * na + d) * S + s];
  }
  if (lane == 0) sh->lam = P.lam[s];
  __syncthreads();
  normal_equations<G, C>(tm, sh, na, lane);
  __syncthreads();
  chol_solve<G, C>(sh, na, lane);
  if (valid) {
    for (int d = lane; d < na; d += G) B.dq[d * S + s] = sh->dq[d];
    if (lane == 0) {
      B.pred[s] = sh->pred;
      P.cost[s] = sh->cost;
    }
  }
}

__global__ void k_plain_trace_state(Problems P, int it) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= P.n || it >= P.trace_iters) return;
  const int S = P.stride, na = P.robot->na;
  for (int d = 0; d < na; ++d) P.q_trace[(it * na + d) * S + b] = P.q[d * S + b];
  P.lam_trace[it * S + b] = P.lam[b];
}

__global__ void k_plain_integrate(Problems P, PlainBufs B) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  const int d = blockIdx.y;
  if (b >= P.n) return;
  const int S = P.stride;
  B.qp[d * S + b] = P.q[d * S + b] + B.dq[d * S + b];
}

// cost at the current iterate from its residual rows (strided layout): rk_batch_solve with zero
// iterations (LMOptimizer._start scores the start, opt/lm_optimizer.py:636-640)
__global__ void k_plain_cost(Problems P, const float* __restrict__ r) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= P.n) return;
  const int S = P.stride, na = P.robot->na;
  const Terms* tm = P.terms;
  float c = 0.f;
  int off = 0;
  const int sizes[6] = {tm->Rf, tm->Rq, tm->Rr, tm->limit ? na : 0, tm->smooth ? na : 0, tm->nc};
  for (int t = 0; t < 6; ++t) {
    if (!sizes[t]) continue;
    float s = 0.f;
    for (int k = 0; k < sizes[t]; ++k) {
      const float v = r[(off + k) * S + b];
      s += v * v;
    }
    c = c + 0.5f * s;
    off += sizes[t];
  }
  P.cost[b] = c;
}

__global__ void k_plain_accept(Problems P, PlainBufs B, LMParams lp, int it) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= P.n) return;
  const int S = P.stride, na = P.robot->na;
  const Terms* tm = P.terms;
  // proposed cost: per-term sums of the proposal's rows (strided layout)
  float cp = 0.f;
  int off = 0;
  const int sizes[5] = {tm->Rf, tm->Rq, tm->Rr, tm->limit ? na : 0, tm->smooth ? na : 0};
  for (int t = 0; t < 5; ++t) {
    if (!sizes[t]) continue;
    float s = 0.f;
    for (int k = 0; k < sizes[t]; ++k) {
      const float v = B.rp[(off + k) * S + b];
      s += v * v;
    }
    cp = cp + 0.5f * s;
    off += sizes[t];
  }
  float cost = P.cost[b], lam = P.lam[b];
  if (P.trace && it == 0) P.trace[b] = cost;
  const bool acc = lm_accept(lp, cost, cp, B.pred[b], lam);
  P.cost[b] = cost;
  P.lam[b] = lam;
  B.acc[b] = acc;
  if (acc) P.nacc[b] += 1;
  if (P.trace && it < P.trace_iters) {
    P.trace[(it + 1) * S + b] = cost;
    P.acc_trace[it * S + b] = acc;
  }
}

__global__ void k_plain_accept_state(Problems P, PlainBufs B) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  const int d = blockIdx.y;
  if (b >= P.n) return;
  const int S = P.stride;
  if (B.acc[b]) P.q[d * S + b] = B.qp[d * S + b];
}

__global__ void k_init_lm(Problems P, float lam0) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= P.n) return;
  P.lam[b] = lam0;
  P.nacc[b] = 0;
}

// ---------------------------------------------------------------------------------------------
// utility kernels: FK / link Jacobian / SE(3) ops / IK metrics
// ---------------------------------------------------------------------------------------------
__device__ void fk_thread(const Robot* rb, const float* q, Pose* Tw) {
  for (int k = 0; k < rb->nj; ++k) {
    const int j = rb->order[k];
    Pose X = joint_local(rb, j, q);
    const int p = rb->parent[j];
    if (p >= 0) X = compose(Tw[p], X);
    Tw[j] = X;
  }
}

// column k of Ad(T) = [[R, t^ R], [0, R]] (se3_kernels.py:323-346): the twist e_k seen through T
__device__ __forceinline__ void adjoint_column(const Pose& T, int k, float* out) {
  const Mat3 R = quat_to_mat(T.q);
  float v[3] = {0.f, 0.f, 0.f}, w[3] = {0.f, 0.f, 0.f}, Rv[3], Rw[3], txRw[3];
  if (k < 3) v[k] = 1.f;
  else w[k - 3] = 1.f;
  mat3_vec(R, v, Rv);
  mat3_vec(R, w, Rw);
  cross3(T.t, Rw, txRw);
  for (int i = 0; i < 3; ++i) {
    out[i] = Rv[i] + txRw[i];
    out[3 + i] = Rw[i];
  }
}

// world poses of every joint of problem b with the robot's base at T_world_base (null: the
// identity): the chain is composed from the root, then the base on the left, as the accumulating
// FK composes each link's chain and then T_world_base (robot_kernels.py:327-377)
__device__ __forceinline__ Pose fk_with_base(const Robot* rb, const float* q, const float* base, Pose* Tw) {
  fk_thread(rb, q, Tw);
  Pose Bp;
  if (base) {
    load_pose(base, Bp);
    for (int j = 0; j < rb->nj; ++j) Tw[j] = compose(Bp, Tw[j]);
  } else {
    Bp.t[0] = Bp.t[1] = Bp.t[2] = 0.f;
    Bp.q[0] = 1.f; Bp.q[1] = Bp.q[2] = Bp.q[3] = 0.f;
  }
  return Bp;
}

// T_link [B][nl][7]; base [B][7] or null (robot.py:131-271: a link without a parent joint sits at
// the base)
__global__ void k_fk_host_layout(const Robot* rb, int B, const float* q, const float* base, float* T_link) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= B) return;
  Pose Tw[RK_NJ_MAX];
  const Pose Bp = fk_with_base(rb, q + (size_t)b * rb->na, base ? base + (size_t)b * 7 : nullptr, Tw);
  for (int l = 0; l < rb->nl; ++l) {
    const int j = rb->link_joint[l];
    store_pose(T_link + ((size_t)b * rb->nl + l) * 7, j >= 0 ? Tw[j] : Bp);
  }
}

// J [B][6][fb + na], fb = 6 floating-base columns first when base is given (robot_kernels.py:667-737):
// base column k is Ad(T_world_base) e_k (spatial) or Ad(T_link^-1 T_world_base) e_k (body); joint
// column a sums the world motion-subspace columns of the link's ancestor joints driving dof a
// (spatial), seen from the link frame through Ad(T_link^-1) (body)
__global__ void k_link_jacobian(const Robot* rb, int B, const float* q, const float* base, int link, int frame,
                                float* J) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= B) return;
  Pose Tw[RK_NJ_MAX];
  const Pose Bp = fk_with_base(rb, q + (size_t)b * rb->na, base ? base + (size_t)b * 7 : nullptr, Tw);
  const int na = rb->na, fb = base ? 6 : 0, ncol = fb + na;
  const int jl = rb->link_joint[link];
  const Pose Tf = jl >= 0 ? Tw[jl] : Bp;
  const Pose Tfi = inverse(Tf);
  const Mat3 R = quat_to_mat(Tfi.q);
  for (int k = 0; k < fb; ++k) {
    float col[6];
    adjoint_column(frame == 1 ? compose(Tfi, Bp) : Bp, k, col);
    for (int i = 0; i < 6; ++i) J[((size_t)b * 6 + i) * ncol + k] = col[i];
  }
  auto S_at = [&](int j, int r) {
    float s6[6];
    subspace_column(rb, j, Tw[j], s6);
    return s6[r];
  };
  for (int a = 0; a < na; ++a) {
    float sc[6], col[6];
    spatial_column(rb, link, a, S_at, sc);
    if (frame == 1) {
      float Rv[3], Rw[3], txRw[3];
      mat3_vec(R, sc, Rv);
      mat3_vec(R, sc + 3, Rw);
      cross3(Tfi.t, Rw, txRw);
      for (int i = 0; i < 3; ++i) {
        col[i] = Rv[i] + txRw[i];
        col[3 + i] = Rw[i];
      }
    } else {
      for (int i = 0; i < 6; ++i) col[i] = sc[i];
    }
    for (int i = 0; i < 6; ++i) J[((size_t)b * 6 + i) * ncol + fb + a] = col[i];
  }
}


// plain streaming copy (16 B per lane per step): a kernel of exactly known HBM traffic, used to
// calibrate the FETCH_SIZE / WRITE_SIZE counters
__global__ void k_copy_probe(const float4* __restrict__ src, float4* __restrict__ dst, long n4) {
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n4; i += (long)gridDim.x * blockDim.x)
    dst[i] = src[i];
}

// a scene of primitives and geometries as scene_sdf reads it, for rk_scene_query
struct SceneTable {
  int ns, nb;
  float obs_s[RK_SCENE_MAXS][4];
  float obs_b[RK_SCENE_MAXB][6];
  int ng;
  SceneGeom geom[RK_SCENE_MAXG];
};
typedef const __attribute__((address_space(4))) SceneTable SceneK;

// one lane per point: the scene's signed distance and the deciding element's normal
__global__ void k_scene_query(const SceneTable* sc_, int n, const float* pts, float* dist, float* normal) {
  SceneK* sc = (SceneK*)sc_;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const bool live = i < n;
  float p[3];
#pragma unroll
  for (int k = 0; k < 3; ++k) p[k] = live ? pts[3 * i + k] : 0.f;
  int grp;
  const float d = scene_sdf(sc, p, grp);
  float nn[3];
  scene_normal(sc, p, scene_winner(sc, p, grp), nn);
  if (live) {
    dist[i] = d;
    if (normal)
#pragma unroll
      for (int k = 0; k < 3; ++k) normal[3 * i + k] = nn[k];
  }
}

__global__ void k_se3_op(int op, int n, const float* a, const float* bb, float* out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  Pose A, Bp;
  load_pose(a + 7 * i, A);
  if (op == 0) {
    load_pose(bb + 7 * i, Bp);
    store_pose(out + 7 * i, compose(A, Bp));
  } else if (op == 1) {
    store_pose(out + 7 * i, inverse(A));
  } else if (op == 2) {
    se3_log(A, out + 6 * i);
  } else if (op == 3) {
    Mat3 MA, MB;
    se3_jlog_blocks(A, MA, MB);
    float* o = out + 36 * i;
    for (int k = 0; k < 36; ++k) o[k] = 0.f;
    for (int r = 0; r < 3; ++r)
      for (int c = 0; c < 3; ++c) {
        o[r * 6 + c] = MA.m[r * 3 + c];
        o[r * 6 + 3 + c] = MB.m[r * 3 + c];
        o[(r + 3) * 6 + 3 + c] = MA.m[r * 3 + c];
      }
  } else if (op == 4) {
    Mat3 R = quat_to_mat(A.q);
    float* o = out + 36 * i;
    for (int k = 0; k < 36; ++k) o[k] = 0.f;
    for (int c = 0; c < 3; ++c) {
      float col[3] = {R.m[c], R.m[3 + c], R.m[6 + c]};
      float tx[3];
      cross3(A.t, col, tx);
      for (int r = 0; r < 3; ++r) {
        o[r * 6 + c] = R.m[r * 3 + c];
        o[r * 6 + 3 + c] = tx[r];
        o[(r + 3) * 6 + 3 + c] = R.m[r * 3 + c];
      }
    }
  } else if (op == 5) {
    float p[3];
    quat_apply(A.q, bb + 3 * i, p);
    for (int k = 0; k < 3; ++k) out[3 * i + k] = p[k] + A.t[k];
  }
}

// _ik_metrics_kernel (helpers/ik/ik.py:59-82)
__global__ void k_ik_metrics(Problems P, float pos_thr, float ang_thr, float* pos_err, float* ang_err, int* success) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= P.n) return;
  const Robot* rb = P.robot;
  const Terms* tm = P.terms;
  const int S = P.stride;
  float q[RK_NA_MAX];
  for (int d = 0; d < rb->na; ++d) q[d] = P.q[d * S + b];
  Pose Tw[RK_NJ_MAX];
  fk_thread(rb, q, Tw);
  int reached = 1;
  for (int t = 0; t < tm->n_pose; ++t) {
    const int j = rb->link_joint[tm->pose_link[t]];
    Pose Tf;
    if (j >= 0) Tf = Tw[j];
    else {
      Tf.t[0] = Tf.t[1] = Tf.t[2] = 0.f;
      Tf.q[0] = 1.f; Tf.q[1] = Tf.q[2] = Tf.q[3] = 0.f;
    }
    float tv[7];
    for (int c = 0; c < 7; ++c) tv[c] = P.pose_tgt[(t * 7 + c) * S + b];
    Pose Tt;
    load_pose(tv, Tt);
    Pose E = compose(inverse(Tf), Tt);
    const float pe = sqrtf(E.t[0] * E.t[0] + E.t[1] * E.t[1] + E.t[2] * E.t[2]);
    const float v = sqrtf(E.q[1] * E.q[1] + E.q[2] * E.q[2] + E.q[3] * E.q[3]);
    const float ae = 2.f * asinf(fminf(v, 1.f));
    pos_err[(size_t)b * tm->n_pose + t] = pe;
    ang_err[(size_t)b * tm->n_pose + t] = ae;
    if (pe >= pos_thr || ae >= ang_thr) reached = 0;
  }
  success[b] = reached;
}

}  // namespace rk

// =============================================================================================
// host side
// =============================================================================================
using namespace rk;

static thread_local std::string g_err;

static int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}

// ---------------------------------------------------------------- allocation-free metrics staging
// rk_batch_metrics_noalloc stages its per-problem results in this file-scope device array instead
// of allocating them per call: position errors [RK_NOALLOC_MAX_PROBLEMS * RK_MAXT], angle errors
// [same], success flags [RK_NOALLOC_MAX_PROBLEMS] (2.25 MiB, held by every device the library runs
// on until the process ends).  In host code `scratch` is only a stand-in for the symbol, never a
// device address: each device's copy is found with hipGetSymbolAddress when the first context is
// created on that device (with the device current) and kept in g_scratch_dev, and only that
// pointer is passed to kernels and copies.  The metrics kernel writes every entry it later reads,
// so nothing needs clearing between calls.
constexpr int RK_MAX_DEVICES = 64;
constexpr size_t RK_NOALLOC_PAIRS = (size_t)RK_NOALLOC_MAX_PROBLEMS * RK_MAXT;
static __device__ __align__(256) float scratch[2 * RK_NOALLOC_PAIRS + RK_NOALLOC_MAX_PROBLEMS];
static std::mutex g_scratch_mu;
static float* g_scratch_dev[RK_MAX_DEVICES] = {};

// set-up of the current device's staging pointer (once per device; the device must be current)
static int scratch_setup(int device) {
  if (device < 0 || device >= RK_MAX_DEVICES) return fail(RK_EINVAL, "device ordinal beyond the library's table");
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (g_scratch_dev[device]) return RK_OK;
  void* p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !p)
    return fail(RK_EHIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
  g_scratch_dev[device] = static_cast<float*>(p);
  return RK_OK;
}

static float* scratch_of(int device) {
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  return device >= 0 && device < RK_MAX_DEVICES ? g_scratch_dev[device] : nullptr;
}

#define HIP_TRY(expr)                                                                      \
  do {                                                                                     \
    hipError_t e_ = (expr);                                                                \
    if (e_ != hipSuccess)                                                                  \
      return fail(RK_EHIP, std::string(#expr) + ": " + hipGetErrorString(e_));             \
  } while (0)

// ---------------------------------------------------------------- population: best-k select + gather
struct SelectArgs {
  const float *cost, *q, *pose, *pos, *rot, *qprev;
  float *q_out, *seed_out, *pose_out, *pos_out, *rot_out, *qprev_out, *cost_out;
  int* sel;
  int cap_in, cap_out, width_in, width_out, na, n_pose7, n_pos3, n_rot4;
};

// One wavefront per problem; lane j holds candidate j (width_in <= 64).  A candidate's rank is
// the number of candidates before it in the order (cost ascending, NaN as +inf, then candidate
// index), the order of _select_best_k_kernel (opt/population_solver.py:24-44): the keys are
// distinct, so the ranks are a permutation and rank r < width_out names output row r.  The
// selected rows (q, also written as the next stage's seeds, every target field and the cost) are then
// copied, lanes over (row, field) pairs (problem-major SoA, stride = capacity).
__global__ __launch_bounds__(64) void k_select_gather(SelectArgs a) {
  __shared__ int pick[WAVE];
  const int b = blockIdx.x, j = threadIdx.x;
  float c = j < a.width_in ? a.cost[b * a.width_in + j] : INFINITY;
  if (c != c) c = INFINITY;
  int rank = 0;
  for (int i = 0; i < a.width_in; ++i) {
    const float ci = __shfl(c, i);
    rank += (ci < c || (ci == c && i < j)) ? 1 : 0;
  }
  if (j < a.width_in && rank < a.width_out) pick[rank] = j;
  __syncthreads();
  // lanes over (kept row r, field f) pairs, so the loads of every kept row are in flight at once
  const int src0 = b * a.width_in, dst0 = b * a.width_out;
  for (int e = j; e < a.width_out; e += WAVE) {
    const int si = src0 + pick[e];
    a.sel[dst0 + e] = si;
    a.cost_out[dst0 + e] = a.cost[si];
  }
  for (int e = j; e < a.width_out * a.na; e += WAVE) {
    const int r = e / a.na, f = e - r * a.na;
    const int si = src0 + pick[r], di = dst0 + r;
    const float v = a.q[(size_t)f * a.cap_in + si];
    a.q_out[(size_t)f * a.cap_out + di] = v;
    a.seed_out[(size_t)f * a.cap_out + di] = v;
    a.qprev_out[(size_t)f * a.cap_out + di] = a.qprev[(size_t)f * a.cap_in + si];
  }
  for (int e = j; e < a.width_out * a.n_pose7; e += WAVE) {
    const int r = e / a.n_pose7, f = e - r * a.n_pose7;
    a.pose_out[(size_t)f * a.cap_out + dst0 + r] = a.pose[(size_t)f * a.cap_in + src0 + pick[r]];
  }
  for (int e = j; e < a.width_out * a.n_pos3; e += WAVE) {
    const int r = e / a.n_pos3, f = e - r * a.n_pos3;
    a.pos_out[(size_t)f * a.cap_out + dst0 + r] = a.pos[(size_t)f * a.cap_in + src0 + pick[r]];
  }
  for (int e = j; e < a.width_out * a.n_rot4; e += WAVE) {
    const int r = e / a.n_rot4, f = e - r * a.n_rot4;
    a.rot_out[(size_t)f * a.cap_out + dst0 + r] = a.rot[(size_t)f * a.cap_in + src0 + pick[r]];
  }
}

struct rk_context {
  int device = 0;
  hipStream_t stream = nullptr;
  std::mutex mu;
  std::vector<Robot> robots;
  Robot* robots_dev = nullptr;
  int cu_count = 0;
  int64_t mem_bytes = 0;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  hipEvent_t marks[8] = {};  // rk_context_mark slots, created on first use
  std::map<int, int> tuned;  // batch signature -> threads per problem chosen by the autotuner
  // scene geometries (rk_geom_upload_*): the kernels' view and the device bytes behind each
  std::vector<SceneGeom> geoms;
  std::vector<std::pair<void*, size_t>> geom_mem;
};

struct rk_batch {
  rk_context* ctx = nullptr;
  Terms th{};
  Terms* td = nullptr;
  int capacity = 0, n = 0, trace_iters = 0, na = 0;
  int it_done = 0;  // iterations run since the last rk_batch_set / seeded solve (RK_CONTINUE offset)
  int traj_T = 0;   // > 0: a trajectory batch (na = traj_T * robot dofs), solved by k_traj
  int traj_na = 0;  // robot dofs of a trajectory batch
  bool traj_lim = false;  // its problems carry bound rows (TrajectoryLimitTask) or scene geometries
  TrajTerms* tt = nullptr;
  float *q = nullptr, *q_seed = nullptr, *pose_tgt = nullptr, *pos_tgt = nullptr, *q_prev = nullptr;
  float* rot_tgt = nullptr;
  float *cost = nullptr, *lam = nullptr, *trace = nullptr;
  int *nacc = nullptr, *acc_trace = nullptr;
  float *q_trace = nullptr, *lam_trace = nullptr;
  PlainBufs pb{};
  bool plain_alloc = false;
  int* sel = nullptr;  // rk_batch_select: source row of each row (allocated on first select into the batch)
  int64_t bytes = 0;
};

#define RK_GROUP_TRAJ_MAX 4  // trajectory arm classes: 6 / 7 dofs, with or without bound rows / geometries
#define RK_VAR_TRAJ 1000     // last_var code of trajectory class c: RK_VAR_TRAJ + c

struct rk_group {
  rk_context* ctx = nullptr;
  std::vector<rk_batch*> batches;
  std::vector<int> tpp;  // chosen threads per problem, 0 = default
  int wg_per_cu = 0, iters_per_launch = 0;
  Bucket* buckets_dev = nullptr;    // bucket table of eager solves (rewritten by each one)
  Bucket* buckets_graph = nullptr;  // bucket table the captured graph reads (written only at capture)
  int* counter = nullptr;  // per variant: tile queue and exit count (self-resetting)
  // trajectory batches (any horizon): one k_traj_multi launch per arm class on its own stream
  std::vector<int> ik, traj;           // indices of the IK / trajectory batches in `batches`
  TrajBucket* traj_dev = nullptr;      // trajectory table of eager solves
  TrajBucket* traj_graph = nullptr;    // trajectory table the captured graph reads
  // kernels of a solve: the IK variants, then the trajectory classes (RK_GROUP_TRAJ_MAX at most)
  hipStream_t aux[NVARIANT + RK_GROUP_TRAJ_MAX] = {};
  hipEvent_t ev_fork = nullptr, ev_join[NVARIANT + RK_GROUP_TRAJ_MAX] = {};
  hipEvent_t ev_k0[NVARIANT + RK_GROUP_TRAJ_MAX] = {}, ev_k1[NVARIANT + RK_GROUP_TRAJ_MAX] = {};
  std::vector<int> last_var;                               // variants of the last solve, launch order
  std::vector<unsigned> last_mask;                         // batches (bit = index) of each of those kernels
  hipGraphExec_t graph = nullptr;
  int graph_iters = -1;
  std::vector<float> graph_key;
  int64_t bytes = 0;
};

static int dmalloc(rk_context* ctx, void** p, size_t bytes, int64_t* acct) {
  if (bytes == 0) bytes = 4;
  hipError_t e = hipMalloc(p, bytes);
  if (e != hipSuccess) return fail(RK_ENOMEM, std::string("hipMalloc: ") + hipGetErrorString(e));
  ctx->mem_bytes += (int64_t)bytes;
  if (acct) *acct += (int64_t)bytes;
  return RK_OK;
}

static void dfree(rk_context* ctx, void* p, size_t bytes) {
  if (!p) return;
  hipFree(p);
  ctx->mem_bytes -= (int64_t)std::max<size_t>(bytes, 4);
}

static Problems make_problems(rk_batch* b, bool from_seed = false) {
  Problems P;
  P.robot = b->ctx->robots_dev + b->th.robot;
  P.terms = b->td;
  P.q = b->q;
  P.q_seed = from_seed ? b->q_seed : nullptr;
  P.pose_tgt = b->pose_tgt;
  P.pos_tgt = b->pos_tgt;
  P.rot_tgt = b->rot_tgt;
  P.q_prev = b->q_prev;
  P.cost = b->cost;
  P.lam = b->lam;
  P.nacc = b->nacc;
  P.trace = b->trace;
  P.acc_trace = b->acc_trace;
  P.q_trace = b->q_trace;
  P.lam_trace = b->lam_trace;
  P.trace_iters = b->trace_iters;
  P.n = b->n;
  P.stride = b->capacity;
  return P;
}

static bool fits_small(const Terms& th, const Robot& r) {
  return r.nj <= CfgS::nj && r.na <= CfgS::na && th.n_pose + th.n_rot <= CfgS::np && th.n_pos <= CfgS::nq && !th.nc;
}

// register thread-per-problem kernels: serial arms with 5..8 dofs, <= 2 fixed tail joints; one
// pose target, or <= 1 position and <= 1 rotation target whose links share a parent joint
static bool fits_thread(const Terms& th, const Robot& r) {
  if (!(r.chain && r.na >= 5 && r.na <= 8 && r.nj <= r.na + 2)) return false;
  if (th.n_pose) return th.n_pose == 1 && th.n_pos == 0 && th.n_rot == 0;
  if (th.n_pos > 1 || th.n_rot > 1) return false;
  return !(th.n_pos && th.n_rot && r.link_joint[th.pos_link[0]] != r.link_joint[th.rot_link[0]]);
}

// register wave kernels: trees with <= 32 dofs, each dof driven by <= 2 joints; <= 4 pose
// targets and no position targets, or <= 12 position targets and no pose targets.  hw = 64: one
// problem per wavefront, <= 64 joints; hw = 32: two per wavefront, <= 32 joints.
static int wave_kind(const Terms& th, const Robot& r, int hw) {
  if (r.nj > hw || r.na > WV_NA || th.n_rot || th.nc) return -1;
  for (int a = 0; a < r.na; ++a)
    if (r.col_start[a + 1] - r.col_start[a] > 2) return -1;
  if (th.n_pos == 0 && th.n_pose <= 4) return hw == 64 ? V_W40 : V_H40;
  if (th.n_pose == 0 && th.n_pos <= 12) return hw == 64 ? V_W012 : V_H012;
  return -1;
}

// quad kernels: 6- and 7-dof serial arms with one pose target (no position / rotation targets, no
// collision rows)
static bool fits_quad(const Terms& th, const Robot& r) {
  return fits_thread(th, r) && (r.na == 6 || r.na == 7) && th.n_pose == 1 && th.nc == 0;
}

// supernodal half-wave kernels: the half-wave kinds whose normal equations have the block-arrow
// form (sn_structure found one); with pose targets also the (target, dof) pair layout (sn_pairs)
static int super_kind(const Terms& th, const Robot& r) {
  const int k = wave_kind(th, r, 32);
  if (k < 0 || th.sn_nslot <= 0) return -1;
  if (k == V_H40 && !th.pr_ok) return -1;  // pose targets: the pair layout is required
  return k == V_H40 ? V_HS40 : V_HS012;
}

static int signature(const Terms& th) {
  return ((((th.robot * 8 + th.n_pose) * 16 + th.n_pos) * 4 + th.limit * 2 + th.smooth) * 8 + th.n_rot) * 2 +
         (th.nc ? 1 : 0);
}

// threads per problem -> variant.  G = 0: the tuned choice for this term structure if the
// autotuner made one on this context, else the default (registers where they fit; a serial arm
// takes a quad of lanes per problem while one lane per problem would leave SIMDs idle: fewer
// than 64 problems per SIMD, n > 0 the problem count, 0 unknown).
static int pick_variant(const rk_context* ctx, const Terms& th, int G, int* v, int n = 0) {
  const Robot& r = ctx->robots[th.robot];
  const bool small = fits_small(th, r);
  if (G == 0) {
    auto it = ctx->tuned.find(signature(th));
    if (it != ctx->tuned.end()) G = it->second;
    else
      G = fits_thread(th, r)               ? (fits_quad(th, r) && n > 0 && n < WAVE * 4 * ctx->cu_count ? RK_TPP_QUAD : 1)
          : super_kind(th, r) >= 0         ? RK_TPP_SUPER
          : wave_kind(th, r, 32) >= 0      ? RK_TPP_HALF
          : wave_kind(th, r, 64) >= 0      ? RK_TPP_WAVE
          : small                          ? 8
                                           : 64;
  }
  const bool pr = th.n_pose == 0 && (th.n_pos || th.n_rot);
  if (th.nc) {  // SceneCollisionTask rows: the register kernels of 6- and 7-dof serial arms
    if ((G == 1 || G == 0) && fits_thread(th, r) && (r.na == 6 || r.na == 7)) {
      *v = (pr ? V_PC6 : V_TC6) + (r.na - 6);
      return RK_OK;
    }
    return fail(RK_EINVAL, "collision rows run the one-lane-per-problem register kernel (threads_per_problem 1) of "
                           "6- and 7-dof serial arms");
  }
  if (G == 1 && fits_thread(th, r)) {
    *v = (pr ? V_P5 : V_T5) + (r.na - 5);
    return RK_OK;
  }
  if ((G == RK_TPP_WAVE || G == RK_TPP_HALF) && wave_kind(th, r, -G) >= 0) {
    *v = wave_kind(th, r, -G);
    return RK_OK;
  }
  if (G == RK_TPP_SUPER && super_kind(th, r) >= 0) {
    *v = super_kind(th, r);
    return RK_OK;
  }
  if (G == RK_TPP_QUAD && fits_quad(th, r)) {
    *v = r.na == 6 ? V_Q6 : V_Q7;
    return RK_OK;
  }
  if (small && (G == 4 || G == 8 || G == 16)) {
    *v = G == 4 ? V_S4 : G == 8 ? V_S8 : V_S16;
    return RK_OK;
  }
  if (G == 16 || G == 32 || G == 64) {
    *v = G == 16 ? V_L16 : G == 32 ? V_L32 : V_L64;
    return RK_OK;
  }
  return fail(RK_EINVAL,
              "threads_per_problem must be 1 (serial arms), -4 (a quad of lanes per problem: 6- and 7-dof serial "
              "arms with one pose target), -64 / -32 (register wave / half-wave: trees), -33 (supernodal "
              "half-wave: trees whose normal equations have the block-arrow form), 4, 8, 16 (small robots) or "
              "16, 32, 64");
}
static int pick_variant(const rk_batch* b, int G, int* v) { return pick_variant(b->ctx, b->th, G, v, b->n); }

static void legal_tpps(const rk_batch* b, std::vector<int>& out) {
  const Robot& r = b->ctx->robots[b->th.robot];
  out.clear();
  if (fits_thread(b->th, r)) out.push_back(1);
  if (wave_kind(b->th, r, 64) >= 0) out.push_back(RK_TPP_WAVE);
  if (wave_kind(b->th, r, 32) >= 0) out.push_back(RK_TPP_HALF);
  if (super_kind(b->th, r) >= 0) out.push_back(RK_TPP_SUPER);
  if (fits_quad(b->th, r)) out.push_back(RK_TPP_QUAD);
  if (fits_small(b->th, r)) out.insert(out.end(), {4, 8, 16, 32, 64});
  else out.insert(out.end(), {16, 32, 64});
}

template <int G, class C>
static int launch_fused(hipStream_t st, const Problems& P, const LMParams& lp, int it0, int iters) {
  const size_t lds = (WAVE / G) * sizeof(Shared<C>);
  static bool attr_set = false;
  if (!attr_set) {
    HIP_TRY(hipFuncSetAttribute((const void*)k_fused<G, C>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    attr_set = true;
  }
  const int grid = (P.n + (WAVE / G) - 1) / (WAVE / G);
  hipLaunchKernelGGL((k_fused<G, C>), dim3(grid), dim3(WAVE), lds, st, P, lp, it0, iters);
  HIP_TRY(hipGetLastError());
  return RK_OK;
}

static int launch_variant(hipStream_t st, int v, const Problems& P, const LMParams& lp, int it0, int iters) {
  switch (v) {
    case V_S4: return launch_fused<4, CfgS>(st, P, lp, it0, iters);
    case V_S8: return launch_fused<8, CfgS>(st, P, lp, it0, iters);
    case V_S16: return launch_fused<16, CfgS>(st, P, lp, it0, iters);
    case V_L16: return launch_fused<16, CfgL>(st, P, lp, it0, iters);
    case V_L32: return launch_fused<32, CfgL>(st, P, lp, it0, iters);
    case V_L64: return launch_fused<64, CfgL>(st, P, lp, it0, iters);
    default: break;
  }
  const dim3 grid((P.n + WAVE - 1) / WAVE);
  switch (v) {
    case V_T5: hipLaunchKernelGGL((k_thread<5, 7>), grid, dim3(WAVE), 0, st, P, lp, it0, iters); break;
    case V_T6: hipLaunchKernelGGL((k_thread<6, 8>), grid, dim3(WAVE), 0, st, P, lp, it0, iters); break;
    case V_T7: hipLaunchKernelGGL((k_thread<7, 9>), grid, dim3(WAVE), 0, st, P, lp, it0, iters); break;
    case V_T8: hipLaunchKernelGGL((k_thread<8, 10>), grid, dim3(WAVE), 0, st, P, lp, it0, iters); break;
    case V_P5: hipLaunchKernelGGL((k_thread<5, 7, true>), grid, dim3(WAVE), 0, st, P, lp, it0, iters); break;
    case V_P6: hipLaunchKernelGGL((k_thread<6, 8, true>), grid, dim3(WAVE), 0, st, P, lp, it0, iters); break;
    case V_P7: hipLaunchKernelGGL((k_thread<7, 9, true>), grid, dim3(WAVE), 0, st, P, lp, it0, iters); break;
    case V_P8: hipLaunchKernelGGL((k_thread<8, 10, true>), grid, dim3(WAVE), 0, st, P, lp, it0, iters); break;
    case V_TC6: hipLaunchKernelGGL((k_thread<6, 8, false, true>), grid, dim3(WAVE), 0, st, P, lp, it0, iters); break;
    case V_TC7: hipLaunchKernelGGL((k_thread<7, 9, false, true>), grid, dim3(WAVE), 0, st, P, lp, it0, iters); break;
    case V_PC6: hipLaunchKernelGGL((k_thread<6, 8, true, true>), grid, dim3(WAVE), 0, st, P, lp, it0, iters); break;
    case V_PC7: hipLaunchKernelGGL((k_thread<7, 9, true, true>), grid, dim3(WAVE), 0, st, P, lp, it0, iters); break;
    case V_W40: hipLaunchKernelGGL((k_wave<4, 0, 64>), dim3(P.n), dim3(WAVE), 0, st, P, lp, it0, iters); break;
    case V_W012: hipLaunchKernelGGL((k_wave<0, 12, 64>), dim3(P.n), dim3(WAVE), 0, st, P, lp, it0, iters); break;
    case V_H40:
      hipLaunchKernelGGL((k_wave<4, 0, 32>), dim3((P.n + 1) / 2), dim3(WAVE), 0, st, P, lp, it0, iters);
      break;
    case V_HS40:
      hipLaunchKernelGGL((k_wave<4, 0, 32, true>), dim3((P.n + 1) / 2), dim3(WAVE), 0, st, P, lp, it0, iters);
      break;
    case V_Q6: hipLaunchKernelGGL((k_quad<6, 8>), dim3((P.n + 15) / 16), dim3(WAVE), 0, st, P, lp, it0, iters); break;
    case V_Q7: hipLaunchKernelGGL((k_quad<7, 9>), dim3((P.n + 15) / 16), dim3(WAVE), 0, st, P, lp, it0, iters); break;
    case V_HS012:
      hipLaunchKernelGGL((k_wave<0, 12, 32, true>), dim3((P.n + 1) / 2), dim3(WAVE), 0, st, P, lp, it0, iters);
      break;
    default:
      hipLaunchKernelGGL((k_wave<0, 12, 32>), dim3((P.n + 1) / 2), dim3(WAVE), 0, st, P, lp, it0, iters);
      break;
  }
  HIP_TRY(hipGetLastError());
  return RK_OK;
}

// trajectory kernel of an arm with `na` dofs (two wavefronts per problem); lim: the problems carry
// bound rows (the kernel instantiation with them); shared: the launch runs beside latency-bound IK
// launches (the request server), which take the form with v_readlane band updates
static int launch_traj(hipStream_t st, int na, const TrajTerms* tt, const Problems& P, const LMParams& lp, int it0,
                       int iters, bool lim, bool shared = false) {
  const dim3 g(P.n), b(2 * WAVE);
  if (na == 6 && lim && shared)
    hipLaunchKernelGGL((k_traj<6, 8, true, false>), g, b, 0, st, P, tt, lp, it0, iters);
  else if (na == 6 && lim)
    hipLaunchKernelGGL((k_traj<6, 8, true>), g, b, 0, st, P, tt, lp, it0, iters);
  else if (na == 6 && shared)
    hipLaunchKernelGGL((k_traj<6, 8, false, false>), g, b, 0, st, P, tt, lp, it0, iters);
  else if (na == 6)
    hipLaunchKernelGGL((k_traj<6, 8>), g, b, 0, st, P, tt, lp, it0, iters);
  else if (lim && shared)
    hipLaunchKernelGGL((k_traj<7, 9, true, false>), g, b, 0, st, P, tt, lp, it0, iters);
  else if (lim)
    hipLaunchKernelGGL((k_traj<7, 9, true>), g, b, 0, st, P, tt, lp, it0, iters);
  else if (shared)
    hipLaunchKernelGGL((k_traj<7, 9, false, false>), g, b, 0, st, P, tt, lp, it0, iters);
  else
    hipLaunchKernelGGL((k_traj<7, 9>), g, b, 0, st, P, tt, lp, it0, iters);
  HIP_TRY(hipGetLastError());
  return RK_OK;
}
static int launch_traj(rk_batch* b, const Problems& P, const LMParams& lp, int it0, int iters) {
  return launch_traj(b->ctx->stream, b->traj_na, b->tt, P, lp, it0, iters, b->traj_lim);
}
static void launch_traj_metrics(hipStream_t st, int na, const TrajTerms* tt, const Problems& P, float* pe, float* ae,
                                float* sd) {
  if (na == 6)
    hipLaunchKernelGGL((k_traj_metrics<6, 8>), dim3(P.n), dim3(WAVE), 0, st, P, tt, pe, ae, sd);
  else
    hipLaunchKernelGGL((k_traj_metrics<7, 9>), dim3(P.n), dim3(WAVE), 0, st, P, tt, pe, ae, sd);
}

static LMParams to_lp(const rk_lm_params* p) {
  LMParams lp;
  lp.lm_lambda = p->lm_lambda;
  lp.lambda_factor = p->lambda_factor;
  lp.lambda_min = p->lambda_min;
  lp.lambda_max = p->lambda_max;
  lp.rho_min = p->rho_min;
  lp.gain_ratio_epsilon = p->gain_ratio_epsilon;
  return lp;
}

static int alloc_plain(rk_batch* b) {
  if (b->plain_alloc) return RK_OK;
  rk_context* ctx = b->ctx;
  const Robot& r = ctx->robots[b->th.robot];
  const size_t S = b->capacity;
  int rc;
  if ((rc = dmalloc(ctx, (void**)&b->pb.Tj, sizeof(float) * r.nj * 7 * S, &b->bytes))) return rc;
  if ((rc = dmalloc(ctx, (void**)&b->pb.Sw, sizeof(float) * r.nj * 6 * S, &b->bytes))) return rc;
  if ((rc = dmalloc(ctx, (void**)&b->pb.J, sizeof(float) * b->th.R * r.na * S, &b->bytes))) return rc;
  if ((rc = dmalloc(ctx, (void**)&b->pb.r, sizeof(float) * b->th.R * S, &b->bytes))) return rc;
  if ((rc = dmalloc(ctx, (void**)&b->pb.rp, sizeof(float) * b->th.R * S, &b->bytes))) return rc;
  if ((rc = dmalloc(ctx, (void**)&b->pb.qp, sizeof(float) * r.na * S, &b->bytes))) return rc;
  if ((rc = dmalloc(ctx, (void**)&b->pb.dq, sizeof(float) * r.na * S, &b->bytes))) return rc;
  if ((rc = dmalloc(ctx, (void**)&b->pb.pred, sizeof(float) * S, &b->bytes))) return rc;
  if ((rc = dmalloc(ctx, (void**)&b->pb.acc, sizeof(int) * S, &b->bytes))) return rc;
  b->plain_alloc = true;
  return RK_OK;
}

template <int G, class C>
static int launch_plain_solve(rk_context* ctx, const Problems& P, const PlainBufs& B) {
  const size_t lds = (WAVE / G) * sizeof(Shared<C>);
  static bool attr_set = false;
  if (!attr_set) {
    HIP_TRY(hipFuncSetAttribute((const void*)k_plain_solve<G, C>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                (int)lds));
    attr_set = true;
  }
  const int grid = (P.n + (WAVE / G) - 1) / (WAVE / G);
  hipLaunchKernelGGL((k_plain_solve<G, C>), dim3(grid), dim3(WAVE), lds, ctx->stream, P, B);
  return RK_OK;
}

// the reference's GPU decomposition, one host-loop iteration = 13 launches for IK terms
// cost at the current q (or the seeds) without iterating, through the per-term residual kernels
static int eval_cost(rk_batch* b, bool from_seed) {
  int rc;
  if ((rc = alloc_plain(b))) return rc;
  rk_context* ctx = b->ctx;
  if (from_seed)
    HIP_TRY(hipMemcpyAsync(b->q, b->q_seed, sizeof(float) * b->na * b->capacity, hipMemcpyDeviceToDevice,
                           ctx->stream));
  Problems P = make_problems(b);
  const int na = ctx->robots[b->th.robot].na;
  const dim3 blk(256), g1((b->n + 255) / 256);
  hipLaunchKernelGGL(k_plain_fk, g1, blk, 0, ctx->stream, P, (const float*)b->q, b->pb.Tj, (float*)nullptr);
  if (b->th.n_pose)
    hipLaunchKernelGGL(k_plain_pose_res, dim3(g1.x, b->th.n_pose), blk, 0, ctx->stream, P, (const float*)b->pb.Tj,
                       b->pb.r);
  if (b->th.n_pos)
    hipLaunchKernelGGL(k_plain_pos_res, dim3(g1.x, b->th.n_pos), blk, 0, ctx->stream, P, (const float*)b->pb.Tj,
                       b->pb.r);
  if (b->th.n_rot)
    hipLaunchKernelGGL(k_plain_rot_res, dim3(g1.x, b->th.n_rot), blk, 0, ctx->stream, P, (const float*)b->pb.Tj,
                       b->pb.r);
  if (b->th.limit || b->th.smooth)
    hipLaunchKernelGGL(k_plain_diag_rows, dim3(g1.x, na), blk, 0, ctx->stream, P, (const float*)b->q, b->pb.r,
                       (float*)nullptr);
  if (b->th.nc)
    hipLaunchKernelGGL(k_plain_coll_res, g1, blk, 0, ctx->stream, P, (const float*)b->pb.Tj, b->pb.r);
  hipLaunchKernelGGL(k_plain_cost, g1, blk, 0, ctx->stream, P, (const float*)b->pb.r);
  HIP_TRY(hipGetLastError());
  return RK_OK;
}

static int solve_plain(rk_batch* b, int it0, int iters, const LMParams& lp, bool from_seed) {
  int rc;
  if (b->th.nc) return fail(RK_EINVAL, "collision rows run the fused register kernel (not the plain per-term launch)");
  if ((rc = alloc_plain(b))) return rc;
  rk_context* ctx = b->ctx;
  if (from_seed)
    HIP_TRY(hipMemcpyAsync(b->q, b->q_seed, sizeof(float) * b->na * b->capacity, hipMemcpyDeviceToDevice,
                           ctx->stream));
  Problems P = make_problems(b);
  const Robot& rob = ctx->robots[b->th.robot];
  const int na = rob.na, S = b->capacity;
  const dim3 blk(256);
  const dim3 g1((b->n + 255) / 256);
  const bool small = fits_small(b->th, rob);
  if (it0 == 0) hipLaunchKernelGGL(k_init_lm, g1, blk, 0, ctx->stream, P, lp.lm_lambda);
  for (int it = it0; it < it0 + iters; ++it) {
    if (P.q_trace) hipLaunchKernelGGL(k_plain_trace_state, g1, blk, 0, ctx->stream, P, it);
    hipLaunchKernelGGL(k_plain_fk, g1, blk, 0, ctx->stream, P, (const float*)b->q, b->pb.Tj, b->pb.Sw);
    const long nJ = (long)b->th.R * na * S;
    hipLaunchKernelGGL(k_fill, dim3((nJ + 255) / 256), blk, 0, ctx->stream, b->pb.J, nJ, 0.f);
    if (b->th.n_pose) {
      hipLaunchKernelGGL(k_plain_pose_res, dim3(g1.x, b->th.n_pose), blk, 0, ctx->stream, P, (const float*)b->pb.Tj,
                         b->pb.r);
      hipLaunchKernelGGL(k_plain_pose_jac, dim3(g1.x, na), blk, 0, ctx->stream, P, (const float*)b->pb.Tj,
                         (const float*)b->pb.Sw, b->pb.J);
    }
    if (b->th.n_pos) {
      hipLaunchKernelGGL(k_plain_pos_res, dim3(g1.x, b->th.n_pos), blk, 0, ctx->stream, P, (const float*)b->pb.Tj,
                         b->pb.r);
      hipLaunchKernelGGL(k_plain_pos_jac, dim3(g1.x, na), blk, 0, ctx->stream, P, (const float*)b->pb.Tj,
                         (const float*)b->pb.Sw, b->pb.J);
    }
    if (b->th.n_rot) {
      hipLaunchKernelGGL(k_plain_rot_res, dim3(g1.x, b->th.n_rot), blk, 0, ctx->stream, P, (const float*)b->pb.Tj,
                         b->pb.r);
      hipLaunchKernelGGL(k_plain_rot_jac, dim3(g1.x, na), blk, 0, ctx->stream, P, (const float*)b->pb.Sw, b->pb.J);
    }
    if (b->th.limit || b->th.smooth)
      hipLaunchKernelGGL(k_plain_diag_rows, dim3(g1.x, na), blk, 0, ctx->stream, P, (const float*)b->q, b->pb.r,
                         b->pb.J);
    if (small) {
      if ((rc = launch_plain_solve<8, CfgS>(ctx, P, b->pb))) return rc;
    } else {
      if ((rc = launch_plain_solve<64, CfgL>(ctx, P, b->pb))) return rc;
    }
    hipLaunchKernelGGL(k_plain_integrate, dim3(g1.x, na), blk, 0, ctx->stream, P, b->pb);
    hipLaunchKernelGGL(k_plain_fk, g1, blk, 0, ctx->stream, P, (const float*)b->pb.qp, b->pb.Tj, (float*)nullptr);
    if (b->th.n_pose)
      hipLaunchKernelGGL(k_plain_pose_res, dim3(g1.x, b->th.n_pose), blk, 0, ctx->stream, P, (const float*)b->pb.Tj,
                         b->pb.rp);
    if (b->th.n_pos)
      hipLaunchKernelGGL(k_plain_pos_res, dim3(g1.x, b->th.n_pos), blk, 0, ctx->stream, P, (const float*)b->pb.Tj,
                         b->pb.rp);
    if (b->th.n_rot)
      hipLaunchKernelGGL(k_plain_rot_res, dim3(g1.x, b->th.n_rot), blk, 0, ctx->stream, P, (const float*)b->pb.Tj,
                         b->pb.rp);
    if (b->th.limit || b->th.smooth)
      hipLaunchKernelGGL(k_plain_diag_rows, dim3(g1.x, na), blk, 0, ctx->stream, P, (const float*)b->pb.qp,
                         b->pb.rp, (float*)nullptr);
    hipLaunchKernelGGL(k_plain_accept, g1, blk, 0, ctx->stream, P, b->pb, lp, it);
    hipLaunchKernelGGL(k_plain_accept_state, dim3(g1.x, na), blk, 0, ctx->stream, P, b->pb);
  }
  HIP_TRY(hipGetLastError());
  return RK_OK;
}

// ---------------------------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------------------------
extern "C" {

const char* rk_last_error(void) { return g_err.c_str(); }
int32_t rk_abi_version(void) { return 1; }

int rk_context_create(int32_t device, rk_context** out) {
  if (!out) return fail(RK_EINVAL, "out is null");
  int ndev = 0;
  HIP_TRY(hipGetDeviceCount(&ndev));
  if (device < 0 || device >= ndev) return fail(RK_EINVAL, "device ordinal out of range");
  HIP_TRY(hipSetDevice(device));
  rk_context* c = new rk_context();
  c->device = device;
  hipError_t e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
  if (e != hipSuccess) {
    delete c;
    return fail(RK_EHIP, std::string("hipStreamCreate: ") + hipGetErrorString(e));
  }
  hipDeviceProp_t prop;
  HIP_TRY(hipGetDeviceProperties(&prop, device));
  c->cu_count = prop.multiProcessorCount;
  HIP_TRY(hipEventCreate(&c->ev0));
  HIP_TRY(hipEventCreate(&c->ev1));
  int rc = dmalloc(c, (void**)&c->robots_dev, sizeof(Robot) * RK_MAX_ROBOTS, nullptr);
  if (!rc) rc = scratch_setup(device);
  if (rc) {
    if (c->robots_dev) dfree(c, c->robots_dev, sizeof(Robot) * RK_MAX_ROBOTS);
    hipStreamDestroy(c->stream);
    delete c;
    return rc;
  }
  *out = c;
  return RK_OK;
}

int rk_context_destroy(rk_context* ctx) {
  if (!ctx) return RK_OK;
  hipSetDevice(ctx->device);
  hipStreamSynchronize(ctx->stream);
  hipFree(ctx->robots_dev);
  for (auto& m : ctx->geom_mem) hipFree(m.first);
  hipEventDestroy(ctx->ev0);
  hipEventDestroy(ctx->ev1);
  for (hipEvent_t e : ctx->marks)
    if (e) hipEventDestroy(e);
  hipStreamDestroy(ctx->stream);
  delete ctx;
  return RK_OK;
}

int rk_context_sync(rk_context* ctx) {
  if (!ctx) return fail(RK_EINVAL, "null context");
  HIP_TRY(hipStreamSynchronize(ctx->stream));
  return RK_OK;
}

int rk_context_mark(rk_context* ctx, int32_t slot) {
  if (!ctx || slot < 0 || slot >= 8) return fail(RK_EINVAL, "null context or mark slot outside 0..7");
  HIP_TRY(hipSetDevice(ctx->device));
  if (!ctx->marks[slot]) HIP_TRY(hipEventCreate(&ctx->marks[slot]));
  HIP_TRY(hipEventRecord(ctx->marks[slot], ctx->stream));
  return RK_OK;
}

int rk_context_elapsed(rk_context* ctx, int32_t a, int32_t b, float* ms) {
  if (!ctx || !ms || a < 0 || a >= 8 || b < 0 || b >= 8) return fail(RK_EINVAL, "bad argument");
  if (!ctx->marks[a] || !ctx->marks[b]) return fail(RK_EINVAL, "mark slot not recorded");
  HIP_TRY(hipEventSynchronize(ctx->marks[b]));
  HIP_TRY(hipEventElapsedTime(ms, ctx->marks[a], ctx->marks[b]));
  return RK_OK;
}

int rk_context_mem(rk_context* ctx, int64_t* bytes) {
  if (!ctx || !bytes) return fail(RK_EINVAL, "null argument");
  *bytes = ctx->mem_bytes;
  return RK_OK;
}

int rk_robot_upload(rk_context* ctx, const rk_robot_desc* d, int32_t* robot_id) {
  if (!ctx || !d || !robot_id) return fail(RK_EINVAL, "null argument");
  std::lock_guard<std::mutex> lk(ctx->mu);
  if ((int)ctx->robots.size() >= RK_MAX_ROBOTS) return fail(RK_ENOMEM, "too many robots on this context");
  if (d->nj < 1 || d->nj > RK_NJ_MAX || d->nl < 1 || d->nl > RK_NL_MAX || d->na < 1 || d->na > RK_NA_MAX)
    return fail(RK_EINVAL, "robot size out of range (nj, nl <= 64, 1 <= na <= 32)");
  // validate, then renumber joints in depth order (parents before children) so that the
  // register kernels can index joints with compile-time indices
  for (int j = 0; j < d->nj; ++j) {
    if (d->parent_joint[j] < -1 || d->parent_joint[j] >= d->nj || d->act_idx[j] >= d->na ||
        d->mimic_idx[j] >= d->na || d->joint_type[j] < 0 || d->joint_type[j] > 2)
      return fail(RK_EINVAL, "invalid joint index or type");
  }
  for (int l = 0; l < d->nl; ++l)
    if (d->link_parent_joint[l] < -1 || d->link_parent_joint[l] >= d->nj)
      return fail(RK_EINVAL, "invalid link parent joint");
  std::vector<int> depth(d->nj, -1);
  for (int pass = 0; pass < d->nj + 1; ++pass) {
    bool changed = false;
    for (int j = 0; j < d->nj; ++j) {
      if (depth[j] >= 0) continue;
      const int p = d->parent_joint[j];
      if (p < 0) depth[j] = 0, changed = true;
      else if (depth[p] >= 0) depth[j] = depth[p] + 1, changed = true;
    }
    if (!changed) break;
  }
  int maxd = 0;
  for (int j = 0; j < d->nj; ++j) {
    if (depth[j] < 0) return fail(RK_EINVAL, "circular joint parents");
    maxd = std::max(maxd, depth[j]);
  }
  std::vector<int> new2old, old2new(d->nj);
  for (int lv = 0; lv <= maxd; ++lv)
    for (int j = 0; j < d->nj; ++j)
      if (depth[j] == lv) new2old.push_back(j);
  for (int n = 0; n < d->nj; ++n) old2new[new2old[n]] = n;
  Robot r;
  memset(&r, 0, sizeof(r));
  r.nj = d->nj;
  r.nl = d->nl;
  r.na = d->na;
  for (int n = 0; n < d->nj; ++n) {
    const int j = new2old[n];
    r.joint_type[n] = d->joint_type[j];
    r.parent[n] = d->parent_joint[j] < 0 ? -1 : old2new[d->parent_joint[j]];
    r.act[n] = d->act_idx[j];
    r.mimic[n] = d->mimic_idx[j];
    r.mul[n] = d->mimic_mul[j];
    r.off[n] = d->mimic_off[j];
    for (int k = 0; k < 3; ++k) r.axis[n][k] = d->joint_axis[3 * j + k];
    for (int k = 0; k < 7; ++k) r.pT[n][k] = d->parent_T[7 * j + k];
    {  // q_P (0, a), not standardised: it is one half of a linear combination, not a rotation
      const float* p = &r.pT[n][3];
      const float* a = r.axis[n];
      r.pqa[n][0] = -(p[1] * a[0] + p[2] * a[1] + p[3] * a[2]);
      r.pqa[n][1] = p[0] * a[0] + p[2] * a[2] - p[3] * a[1];
      r.pqa[n][2] = p[0] * a[1] - p[1] * a[2] + p[3] * a[0];
      r.pqa[n][3] = p[0] * a[2] + p[1] * a[1] - p[2] * a[0];
    }
    r.order[n] = n;
  }
  for (int l = 0; l < d->nl; ++l) r.link_joint[l] = d->link_parent_joint[l] < 0 ? -1 : old2new[d->link_parent_joint[l]];
  for (int a = 0; a < d->na; ++a) {
    r.lo[a] = d->limits[2 * a];
    r.hi[a] = d->limits[2 * a + 1];
  }
  r.nlevels = maxd + 1;
  {
    int k = 0;
    for (int lv = 0; lv <= maxd; ++lv) {
      r.level_start[lv] = k;
      while (k < d->nj && depth[new2old[k]] == lv) ++k;
    }
    r.level_start[maxd + 1] = k;
  }
  r.nrounds = 0;
  while ((1 << r.nrounds) < maxd + 1) ++r.nrounds;
  // serial arm: joint j's parent is j-1, joints 0..na-1 actuated 1:1, the rest fixed, no mimic
  r.chain = 1;
  for (int j = 0; j < r.nj; ++j) {
    if (r.parent[j] != j - 1 || r.mimic[j] != -1) r.chain = 0;
    if (j < r.na && (r.act[j] != j || r.joint_type[j] == 0)) r.chain = 0;
    if (j >= r.na && (r.act[j] != -1 || r.joint_type[j] != 0)) r.chain = 0;
  }
  for (int l = 0; l < d->nl; ++l) {
    unsigned long long m = 0;
    for (int j = r.link_joint[l]; j != -1; j = r.parent[j]) m |= 1ull << j;
    r.anc[l] = m;
  }
  int c = 0;
  for (int a = 0; a < d->na; ++a) {
    r.col_start[a] = c;
    for (int j = 0; j < d->nj; ++j) {
      float w = 0.f;
      if (r.act[j] >= 0) w = r.act[j] == a ? 1.f : 0.f;
      else if (r.mimic[j] >= 0) w = r.mimic[j] == a ? r.mul[j] : 0.f;
      if (w != 0.f) {
        if (c >= RK_NJ_MAX) return fail(RK_EINVAL, "too many joint-to-dof entries");
        r.col_joint[c] = j;
        r.col_w[c] = w;
        ++c;
      }
    }
  }
  r.col_start[d->na] = c;
  const int id = (int)ctx->robots.size();
  HIP_TRY(hipSetDevice(ctx->device));
  HIP_TRY(hipMemcpyAsync(ctx->robots_dev + id, &r, sizeof(Robot), hipMemcpyHostToDevice, ctx->stream));
  HIP_TRY(hipStreamSynchronize(ctx->stream));
  ctx->robots.push_back(r);
  *robot_id = id;
  return RK_OK;
}

static int check_robot(rk_context* ctx, int32_t robot) {
  if (!ctx) return fail(RK_EINVAL, "null context");
  if (robot < 0 || robot >= (int)ctx->robots.size()) return fail(RK_EINVAL, "unknown robot id");
  return RK_OK;
}

int rk_fk_base(rk_context* ctx, int32_t robot, int32_t B, const float* q, const float* T_world_base, float* T_link) {
  int rc;
  if ((rc = check_robot(ctx, robot))) return rc;
  if (B <= 0 || !q || !T_link) return fail(RK_EINVAL, "bad batch or null array");
  std::lock_guard<std::mutex> lk(ctx->mu);
  HIP_TRY(hipSetDevice(ctx->device));
  const Robot& r = ctx->robots[robot];
  const size_t qb = sizeof(float) * B * r.na, bb = T_world_base ? sizeof(float) * B * 7 : 0,
               tb = sizeof(float) * B * r.nl * 7;
  char* d = nullptr;
  HIP_TRY(hipMalloc(&d, qb + bb + tb));
  float* dq = reinterpret_cast<float*>(d);
  float* dbase = T_world_base ? reinterpret_cast<float*>(d + qb) : nullptr;
  float* dT = reinterpret_cast<float*>(d + qb + bb);
  hipMemcpyAsync(dq, q, qb, hipMemcpyHostToDevice, ctx->stream);
  if (dbase) hipMemcpyAsync(dbase, T_world_base, bb, hipMemcpyHostToDevice, ctx->stream);
  hipLaunchKernelGGL(k_fk_host_layout, dim3((B + 127) / 128), dim3(128), 0, ctx->stream, ctx->robots_dev + robot, B,
                     (const float*)dq, (const float*)dbase, dT);
  hipMemcpyAsync(T_link, dT, tb, hipMemcpyDeviceToHost, ctx->stream);
  hipError_t e = hipStreamSynchronize(ctx->stream);
  hipFree(d);
  if (e != hipSuccess) return fail(RK_EHIP, std::string("rk_fk: ") + hipGetErrorString(e));
  return RK_OK;
}

int rk_fk(rk_context* ctx, int32_t robot, int32_t B, const float* q, float* T_link) {
  return rk_fk_base(ctx, robot, B, q, nullptr, T_link);
}

int rk_link_jacobian_base(rk_context* ctx, int32_t robot, int32_t B, const float* q, const float* T_world_base,
                          int32_t link, int32_t frame, float* J) {
  int rc;
  if ((rc = check_robot(ctx, robot))) return rc;
  const Robot& r = ctx->robots[robot];
  if (B <= 0 || !q || !J || link < 0 || link >= r.nl || (frame != 0 && frame != 1))
    return fail(RK_EINVAL, "bad argument to rk_link_jacobian");
  std::lock_guard<std::mutex> lk(ctx->mu);
  HIP_TRY(hipSetDevice(ctx->device));
  const int ncol = (T_world_base ? 6 : 0) + r.na;
  const size_t qb = sizeof(float) * B * r.na, bb = T_world_base ? sizeof(float) * B * 7 : 0,
               jb = sizeof(float) * B * 6 * ncol;
  char* d = nullptr;
  HIP_TRY(hipMalloc(&d, qb + bb + jb));
  float* dq = reinterpret_cast<float*>(d);
  float* dbase = T_world_base ? reinterpret_cast<float*>(d + qb) : nullptr;
  float* dJ = reinterpret_cast<float*>(d + qb + bb);
  hipMemcpyAsync(dq, q, qb, hipMemcpyHostToDevice, ctx->stream);
  if (dbase) hipMemcpyAsync(dbase, T_world_base, bb, hipMemcpyHostToDevice, ctx->stream);
  hipLaunchKernelGGL(k_link_jacobian, dim3((B + 127) / 128), dim3(128), 0, ctx->stream, ctx->robots_dev + robot, B,
                     (const float*)dq, (const float*)dbase, link, frame, dJ);
  hipMemcpyAsync(J, dJ, jb, hipMemcpyDeviceToHost, ctx->stream);
  hipError_t e = hipStreamSynchronize(ctx->stream);
  hipFree(d);
  if (e != hipSuccess) return fail(RK_EHIP, std::string("rk_link_jacobian: ") + hipGetErrorString(e));
  return RK_OK;
}

int rk_link_jacobian(rk_context* ctx, int32_t robot, int32_t B, const float* q, int32_t link, int32_t frame,
                     float* J) {
  return rk_link_jacobian_base(ctx, robot, B, q, nullptr, link, frame, J);
}

int rk_copy_probe(rk_context* ctx, int64_t bytes, float* kernel_ms) {
  if (!ctx || bytes < 16 || bytes % 16) return fail(RK_EINVAL, "bytes must be a positive multiple of 16");
  std::lock_guard<std::mutex> lk(ctx->mu);
  HIP_TRY(hipSetDevice(ctx->device));
  float4 *src = nullptr, *dst = nullptr;
  HIP_TRY(hipMalloc(&src, bytes));
  HIP_TRY(hipMalloc(&dst, bytes));
  const long n4 = bytes / 16;
  const int grid = (int)std::min<long>((n4 + 255) / 256, 4L * 256 * ctx->cu_count);
  hipLaunchKernelGGL(k_copy_probe, dim3(grid), dim3(256), 0, ctx->stream, (const float4*)src, dst, n4);
  HIP_TRY(hipEventRecord(ctx->ev0, ctx->stream));
  hipLaunchKernelGGL(k_copy_probe, dim3(grid), dim3(256), 0, ctx->stream, (const float4*)src, dst, n4);
  HIP_TRY(hipEventRecord(ctx->ev1, ctx->stream));
  hipError_t e = hipEventSynchronize(ctx->ev1);
  if (e == hipSuccess && kernel_ms) e = hipEventElapsedTime(kernel_ms, ctx->ev0, ctx->ev1);
  hipFree(src);
  hipFree(dst);
  if (e != hipSuccess) return fail(RK_EHIP, std::string("rk_copy_probe: ") + hipGetErrorString(e));
  return RK_OK;
}

int rk_se3_op(rk_context* ctx, int32_t op, int32_t n, const float* a, const float* b, float* out) {
  if (!ctx || n <= 0 || !a || !out || op < 0 || op > 5 || ((op == 0 || op == 5) && !b))
    return fail(RK_EINVAL, "bad argument to rk_se3_op");
  std::lock_guard<std::mutex> lk(ctx->mu);
  HIP_TRY(hipSetDevice(ctx->device));
  const int ow = op == 0 || op == 1 ? 7 : op == 2 ? 6 : op == 5 ? 3 : 36;
  const int bw = op == 5 ? 3 : 7;
  float *da = nullptr, *db = nullptr, *dout = nullptr;
  HIP_TRY(hipMalloc(&da, sizeof(float) * n * 7));
  HIP_TRY(hipMalloc(&db, sizeof(float) * n * 7));
  HIP_TRY(hipMalloc(&dout, sizeof(float) * n * ow));
  hipMemcpyAsync(da, a, sizeof(float) * n * 7, hipMemcpyHostToDevice, ctx->stream);
  if (b) hipMemcpyAsync(db, b, sizeof(float) * n * bw, hipMemcpyHostToDevice, ctx->stream);
  hipLaunchKernelGGL(k_se3_op, dim3((n + 127) / 128), dim3(128), 0, ctx->stream, op, n, (const float*)da,
                     (const float*)db, dout);
  hipMemcpyAsync(out, dout, sizeof(float) * n * ow, hipMemcpyDeviceToHost, ctx->stream);
  hipError_t e = hipStreamSynchronize(ctx->stream);
  hipFree(da);
  hipFree(db);
  hipFree(dout);
  if (e != hipSuccess) return fail(RK_EHIP, std::string("rk_se3_op: ") + hipGetErrorString(e));
  return RK_OK;
}

static int scene_geoms(rk_context* ctx, int n, const int32_t* ids, int* ng, SceneGeom* out);

static int geom_store(rk_context* ctx, const SceneGeom& g, const float* host, size_t floats, int32_t* id) {
  if ((int)ctx->geoms.size() >= RK_GEOM_MAX) return fail(RK_EINVAL, "at most 64 geometries per context");
  void* dev = nullptr;
  int rc;
  if ((rc = dmalloc(ctx, &dev, sizeof(float) * floats, nullptr))) return rc;
  hipError_t e = hipMemcpy(dev, host, sizeof(float) * floats, hipMemcpyHostToDevice);
  if (e != hipSuccess) {
    dfree(ctx, dev, sizeof(float) * floats);
    return fail(RK_EHIP, std::string("geometry upload: ") + hipGetErrorString(e));
  }
  SceneGeom gd = g;
  gd.data = (const float*)dev;
  ctx->geoms.push_back(gd);
  ctx->geom_mem.push_back({dev, sizeof(float) * floats});
  *id = (int32_t)ctx->geoms.size() - 1;
  return RK_OK;
}

int rk_geom_upload_volume(rk_context* ctx, const rk_volume_desc* d, int32_t* geom_id) {
  if (!ctx || !d || !geom_id || !d->values) return fail(RK_EINVAL, "null argument");
  if (d->nx < 2 || d->ny < 2 || d->nz < 2) return fail(RK_EINVAL, "a volume needs at least 2 grid points per axis");
  const size_t nv = (size_t)d->nx * d->ny * d->nz;
  if (nv > ((size_t)1 << 26)) return fail(RK_EINVAL, "at most 2^26 grid values");
  if (!(d->voxel_size > 0.f) || !std::isfinite(d->padding)) return fail(RK_EINVAL, "voxel_size must be positive");
  const int n[3] = {d->nx, d->ny, d->nz};
  SceneGeom g{};
  g.kind = 0;
  g.nx = d->nx;
  g.ny = d->ny;
  g.nz = d->nz;
  g.inv_h = 1.f / d->voxel_size;
  g.pad = d->padding;
  for (int k = 0; k < 3; ++k) {
    const float top = d->origin[k] + (float)(n[k] - 1) * d->voxel_size, tol = 1e-4f * d->voxel_size;
    if (!(d->aabb_min[k] <= d->aabb_max[k]) || d->aabb_min[k] < d->origin[k] - tol || d->aabb_max[k] > top + tol)
      return fail(RK_EINVAL, "the volume's box must be ordered and lie inside its grid");
    g.org[k] = d->origin[k];
    g.lo[k] = d->aabb_min[k];
    g.hi[k] = d->aabb_max[k];
  }
  std::lock_guard<std::mutex> lk(ctx->mu);
  HIP_TRY(hipSetDevice(ctx->device));
  return geom_store(ctx, g, d->values, nv, geom_id);
}

// the triangle table of a mesh: vertices, face normal, and the angle-weighted pseudonormals of
// its edges (sum of the two adjacent face normals) and vertices (incident angles times face
// normals), normalised, computed in double (Baerentzen and Aanaes, "Signed distance computation
// using the angle weighted pseudonormal", IEEE TVCG 11(3), 2005)
int rk_geom_upload_mesh(rk_context* ctx, const rk_mesh_desc* d, int32_t* geom_id) {
  if (!ctx || !d || !geom_id || !d->vertices || !d->faces) return fail(RK_EINVAL, "null argument");
  const int nv = d->n_vertices, nf = d->n_faces;
  if (nv < 3 || nf < 1 || nf > (1 << 20)) return fail(RK_EINVAL, "a mesh needs >= 3 vertices and 1..2^20 faces");
  for (int i = 0; i < 3 * nf; ++i)
    if (d->faces[i] < 0 || d->faces[i] >= nv) return fail(RK_EINVAL, "face vertex index out of range");
  auto V = [&](int v, int k) { return (double)d->vertices[3 * v + k]; };
  auto unit = [](double* x) {
    const double l = std::sqrt(x[0] * x[0] + x[1] * x[1] + x[2] * x[2]);
    if (l > 0) for (int k = 0; k < 3; ++k) x[k] /= l;
  };
  std::vector<double> fn(3 * (size_t)nf), vn(3 * (size_t)nv, 0.0);
  std::map<std::pair<int, int>, std::array<double, 3>> en;
  double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
  for (int f = 0; f < nf; ++f) {
    const int* F = d->faces + 3 * f;
    double e1[3], e2[3], nn[3];
    for (int k = 0; k < 3; ++k) {
      e1[k] = V(F[1], k) - V(F[0], k);
      e2[k] = V(F[2], k) - V(F[0], k);
    }
    nn[0] = e1[1] * e2[2] - e1[2] * e2[1];
    nn[1] = e1[2] * e2[0] - e1[0] * e2[2];
    nn[2] = e1[0] * e2[1] - e1[1] * e2[0];
    unit(nn);
    for (int k = 0; k < 3; ++k) fn[3 * f + k] = nn[k];
    for (int c = 0; c < 3; ++c) {
      const int v0 = F[c], v1 = F[(c + 1) % 3], v2 = F[(c + 2) % 3];
      double a[3], b[3];
      for (int k = 0; k < 3; ++k) {
        a[k] = V(v1, k) - V(v0, k);
        b[k] = V(v2, k) - V(v0, k);
        lo[k] = std::min(lo[k], V(v0, k));
        hi[k] = std::max(hi[k], V(v0, k));
      }
      const double la = std::sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2]);
      const double lb = std::sqrt(b[0] * b[0] + b[1] * b[1] + b[2] * b[2]);
      double cs = la > 0 && lb > 0 ? (a[0] * b[0] + a[1] * b[1] + a[2] * b[2]) / (la * lb) : 1.0;
      const double ang = std::acos(std::max(-1.0, std::min(1.0, cs)));
      for (int k = 0; k < 3; ++k) vn[3 * v0 + k] += ang * nn[k];
      auto& e = en[{std::min(v0, v1), std::max(v0, v1)}];
      for (int k = 0; k < 3; ++k) e[k] += nn[k];
    }
  }
  std::vector<float> tab((size_t)nf * RK_TRI_FLOATS);
  for (int f = 0; f < nf; ++f) {
    const int* F = d->faces + 3 * f;
    float* T = tab.data() + (size_t)f * RK_TRI_FLOATS;
    for (int c = 0; c < 3; ++c)
      for (int k = 0; k < 3; ++k) T[3 * c + k] = d->vertices[3 * F[c] + k];
    for (int k = 0; k < 3; ++k) T[9 + k] = (float)fn[3 * f + k];
    for (int c = 0; c < 3; ++c) {
      std::array<double, 3> e = en[{std::min(F[c], F[(c + 1) % 3]), std::max(F[c], F[(c + 1) % 3])}];
      unit(e.data());
      double v[3] = {vn[3 * F[c]], vn[3 * F[c] + 1], vn[3 * F[c] + 2]};
      unit(v);
      for (int k = 0; k < 3; ++k) {
        T[12 + 3 * c + k] = (float)e[k];
        T[21 + 3 * c + k] = (float)v[k];
      }
    }
  }
  SceneGeom g{};
  g.kind = 1;
  g.nx = nf;
  for (int k = 0; k < 3; ++k) {
    g.lo[k] = (float)lo[k];
    g.hi[k] = (float)hi[k];
  }
  std::lock_guard<std::mutex> lk(ctx->mu);
  HIP_TRY(hipSetDevice(ctx->device));
  return geom_store(ctx, g, tab.data(), tab.size(), geom_id);
}

int rk_scene_query(rk_context* ctx, int32_t n_obs_spheres, const float* obs_spheres, int32_t n_obs_boxes,
                   const float* obs_boxes, int32_t n_geoms, const int32_t* geoms, int32_t n, const float* points,
                   float* dist, float* normal) {
  if (!ctx || n < 0 || (n && (!points || !dist))) return fail(RK_EINVAL, "null argument");
  if (n_obs_spheres < 0 || n_obs_spheres > RK_SCENE_MAXS || n_obs_boxes < 0 || n_obs_boxes > RK_SCENE_MAXB ||
      (n_obs_spheres && !obs_spheres) || (n_obs_boxes && !obs_boxes))
    return fail(RK_EINVAL, "at most 64 obstacle spheres and 16 obstacle boxes");
  if (n == 0) return RK_OK;
  std::lock_guard<std::mutex> lk(ctx->mu);
  HIP_TRY(hipSetDevice(ctx->device));
  SceneTable st;
  memset(&st, 0, sizeof(st));
  st.ns = n_obs_spheres;
  st.nb = n_obs_boxes;
  for (int i = 0; i < RK_SCENE_MAXS; ++i)
    for (int c = 0; c < 4; ++c) st.obs_s[i][c] = i < st.ns ? obs_spheres[4 * i + c] : (c < 3 ? RK_SDF_FAR : 0.f);
  for (int i = 0; i < RK_SCENE_MAXB; ++i)
    for (int c = 0; c < 6; ++c) st.obs_b[i][c] = i < st.nb ? obs_boxes[6 * i + c] : (c < 3 ? RK_SDF_FAR : 0.f);
  int rc;
  if ((rc = scene_geoms(ctx, n_geoms, geoms, &st.ng, st.geom))) return rc;
  SceneTable* dst = nullptr;
  float *dp = nullptr, *dd = nullptr, *dn = nullptr;
  const size_t pb = sizeof(float) * 3 * (size_t)n;
  hipError_t e = hipMalloc(&dst, sizeof(SceneTable));
  if (e == hipSuccess) e = hipMalloc(&dp, pb);
  if (e == hipSuccess) e = hipMalloc(&dd, sizeof(float) * (size_t)n);
  if (e == hipSuccess) e = hipMalloc(&dn, pb);
  if (e == hipSuccess) e = hipMemcpyAsync(dst, &st, sizeof(SceneTable), hipMemcpyHostToDevice, ctx->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(dp, points, pb, hipMemcpyHostToDevice, ctx->stream);
  if (e == hipSuccess) {
    hipLaunchKernelGGL(k_scene_query, dim3((n + 127) / 128), dim3(128), 0, ctx->stream, (const SceneTable*)dst, n,
                       (const float*)dp, dd, dn);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipMemcpyAsync(dist, dd, sizeof(float) * (size_t)n, hipMemcpyDeviceToHost, ctx->stream);
  if (e == hipSuccess && normal) e = hipMemcpyAsync(normal, dn, pb, hipMemcpyDeviceToHost, ctx->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  hipFree(dst);
  hipFree(dp);
  hipFree(dd);
  hipFree(dn);
  if (e != hipSuccess) return fail(RK_EHIP, std::string("rk_scene_query: ") + hipGetErrorString(e));
  return RK_OK;
}

}  // extern "C"

// the scene geometries of a descriptor, by id, copied into a term table
static int scene_geoms(rk_context* ctx, int n, const int32_t* ids, int* ng, SceneGeom* out) {
  if (n < 0 || n > RK_SCENE_MAXG) return fail(RK_EINVAL, "at most 4 volumes / meshes per scene");
  if (n && !ids) return fail(RK_EINVAL, "null geometry ids");
  for (int k = 0; k < n; ++k) {
    if (ids[k] < 0 || ids[k] >= (int)ctx->geoms.size()) return fail(RK_EINVAL, "unknown geometry id");
    out[k] = ctx->geoms[ids[k]];
  }
  *ng = n;
  return RK_OK;
}

// validated term structure of a descriptor (rk_batch_create, rk_server_add_terms)
static int make_terms(rk_context* ctx, const rk_terms_desc* t, Terms* out) {
  int rc;
  if (!t || !out) return fail(RK_EINVAL, "null argument");
  if ((rc = check_robot(ctx, t->robot))) return rc;
  const Robot& r = ctx->robots[t->robot];
  if (t->n_pose < 0 || t->n_pose > RK_MAXT || t->n_pos < 0 || t->n_pos > RK_MAXQ)
    return fail(RK_EINVAL, "at most 4 pose targets and 12 position targets per problem");
  if (t->n_rot < 0 || t->n_pose + t->n_rot > RK_MAXT)
    return fail(RK_EINVAL, "at most 4 pose and rotation targets together per problem");
  if (t->n_pose + t->n_pos + t->n_rot + (t->limit ? 1 : 0) + (t->smooth ? 1 : 0) + (t->n_coll > 0 ? 1 : 0) == 0)
    return fail(RK_EINVAL, "a batch needs at least one residual term");
  if (t->n_coll < 0 || t->n_coll > RK_COLL_MAXS || (t->n_coll && (t->n_obs_spheres < 0 ||
      t->n_obs_spheres > RK_SCENE_MAXS || t->n_obs_boxes < 0 || t->n_obs_boxes > RK_SCENE_MAXB)))
    return fail(RK_EINVAL, "at most 64 robot spheres, 64 obstacle spheres and 16 obstacle boxes");
  if (t->n_coll && (t->coll_penalty < 0 || t->coll_penalty > 2))
    return fail(RK_EINVAL, "coll_penalty: 0 smooth, 1 plain, 2 surface distance");
  if (r.nj > CfgL::nj) return fail(RK_EINVAL, "the LM kernels take robots with nj <= 32");
  Terms th;
  memset(&th, 0, sizeof(th));
  th.robot = t->robot;
  th.n_pose = t->n_pose;
  th.n_pos = t->n_pos;
  th.limit = t->limit ? 1 : 0;
  th.smooth = t->smooth ? 1 : 0;
  for (int i = 0; i < t->n_pose; ++i) {
    th.pose_link[i] = t->pose_links[i];
    if (th.pose_link[i] < 0 || th.pose_link[i] >= r.nl) return fail(RK_EINVAL, "pose link out of range");
    for (int k = 0; k < 6; ++k) th.pose_w[i][k] = t->pose_weights[6 * i + k];
  }
  for (int i = 0; i < t->n_pos; ++i) {
    th.pos_link[i] = t->pos_links[i];
    if (th.pos_link[i] < 0 || th.pos_link[i] >= r.nl) return fail(RK_EINVAL, "position link out of range");
    for (int k = 0; k < 3; ++k) th.pos_w[i][k] = t->pos_weights[3 * i + k];
  }
  for (int d = 0; d < r.na; ++d) {
    th.limit_w[d] = th.limit ? t->limit_weights[d] : 0.f;
    th.smooth_w[d] = th.smooth ? t->smooth_weights[d] : 0.f;
  }
  th.n_rot = t->n_rot;
  for (int i = 0; i < t->n_rot; ++i) {
    th.rot_link[i] = t->rot_links[i];
    if (th.rot_link[i] < 0 || th.rot_link[i] >= r.nl) return fail(RK_EINVAL, "rotation link out of range");
    for (int k = 0; k < 3; ++k) th.rot_w[i][k] = t->rot_weights[3 * i + k];
  }
  // SceneCollisionTask: spheres sorted (stably) by the joint whose pose moves their link, the scene
  // padded with far-away primitives to whole groups (8 spheres, 4 boxes)
  th.nc = t->n_coll;
  if (th.nc) {
    std::vector<int> order(th.nc);
    for (int k = 0; k < th.nc; ++k) {
      order[k] = k;
      if (t->coll_links[k] < 0 || t->coll_links[k] >= r.nl) return fail(RK_EINVAL, "collision sphere link out of range");
    }
    std::stable_sort(order.begin(), order.end(),
                     [&](int a, int b) { return r.link_joint[t->coll_links[a]] < r.link_joint[t->coll_links[b]]; });
    for (int i = 0; i < th.nc; ++i) {
      const int k = order[i];
      th.c_joint[i] = r.link_joint[t->coll_links[k]];
      for (int c = 0; c < 3; ++c) th.c_c[i][c] = t->coll_centers[3 * k + c];
      th.c_r[i] = t->coll_radii[k];
      th.c_w[i] = t->coll_weights ? t->coll_weights[k] : 1.f;
    }
    th.c_margin = t->coll_margin;
    th.c_penalty = t->coll_penalty;
    th.ns = t->n_obs_spheres;
    th.nb = t->n_obs_boxes;
    for (int i = 0; i < RK_SCENE_MAXS; ++i)
      for (int c = 0; c < 4; ++c) th.obs_s[i][c] = i < th.ns ? t->obs_spheres[4 * i + c] : (c < 3 ? RK_SDF_FAR : 0.f);
    for (int i = 0; i < RK_SCENE_MAXB; ++i)
      for (int c = 0; c < 6; ++c) th.obs_b[i][c] = i < th.nb ? t->obs_boxes[6 * i + c] : (c < 3 ? RK_SDF_FAR : 0.f);
    if ((rc = scene_geoms(ctx, t->n_geoms, t->geoms, &th.ng, th.geom))) return rc;
  }
  th.Rf = 6 * th.n_pose;
  th.Rq = 3 * th.n_pos;
  th.Rr = 3 * th.n_rot;
  th.R = th.Rf + th.Rq + th.Rr + (th.limit ? r.na : 0) + (th.smooth ? r.na : 0) + th.nc;
  wave_elimination(r, th);
  *out = th;
  return RK_OK;
}

extern "C" {

int rk_batch_create(rk_context* ctx, const rk_terms_desc* t, int32_t capacity, int32_t trace_iters, rk_batch** out) {
  int rc;
  if (!t || !out) return fail(RK_EINVAL, "null argument");
  if (capacity <= 0) return fail(RK_EINVAL, "capacity must be positive");
  if (trace_iters < 0) return fail(RK_EINVAL, "trace_iters must be >= 0");
  Terms th;
  if ((rc = make_terms(ctx, t, &th))) return rc;
  const Robot& r = ctx->robots[t->robot];
  std::lock_guard<std::mutex> lk(ctx->mu);
  HIP_TRY(hipSetDevice(ctx->device));
  rk_batch* b = new rk_batch();
  b->ctx = ctx;
  b->th = th;
  b->capacity = capacity;
  b->trace_iters = trace_iters;
  b->na = r.na;
  const size_t S = capacity;
  auto A = [&](void** p, size_t bytes) { return dmalloc(ctx, p, bytes, &b->bytes); };
  if ((rc = A((void**)&b->td, sizeof(Terms))) || (rc = A((void**)&b->q, sizeof(float) * r.na * S)) ||
      (rc = A((void**)&b->q_seed, sizeof(float) * r.na * S)) ||
      (rc = A((void**)&b->pose_tgt, sizeof(float) * std::max(1, th.n_pose * 7) * S)) ||
      (rc = A((void**)&b->pos_tgt, sizeof(float) * std::max(1, th.n_pos * 3) * S)) ||
      (rc = A((void**)&b->rot_tgt, sizeof(float) * std::max(1, th.n_rot * 4) * S)) ||
      (rc = A((void**)&b->q_prev, sizeof(float) * r.na * S)) || (rc = A((void**)&b->cost, sizeof(float) * S)) ||
      (rc = A((void**)&b->lam, sizeof(float) * S)) || (rc = A((void**)&b->nacc, sizeof(int) * S))) {
    rk_batch_destroy(b);
    return rc;
  }
  if (trace_iters > 0) {
    if ((rc = A((void**)&b->trace, sizeof(float) * (trace_iters + 1) * S)) ||
        (rc = A((void**)&b->acc_trace, sizeof(int) * trace_iters * S)) ||
        (rc = A((void**)&b->q_trace, sizeof(float) * trace_iters * r.na * S)) ||
        (rc = A((void**)&b->lam_trace, sizeof(float) * trace_iters * S))) {
      rk_batch_destroy(b);
      return rc;
    }
  }
  hipMemcpyAsync(b->td, &th, sizeof(Terms), hipMemcpyHostToDevice, ctx->stream);
  hipMemsetAsync(b->q_prev, 0, sizeof(float) * r.na * S, ctx->stream);
  hipMemsetAsync(b->cost, 0, sizeof(float) * S, ctx->stream);
  HIP_TRY(hipStreamSynchronize(ctx->stream));
  *out = b;
  return RK_OK;
}

int rk_batch_destroy(rk_batch* b) {
  if (!b) return RK_OK;
  rk_context* ctx = b->ctx;
  hipSetDevice(ctx->device);
  hipStreamSynchronize(ctx->stream);
  void* ptrs[] = {b->rot_tgt, b->tt, b->td, b->q, b->q_seed, b->pose_tgt, b->pos_tgt, b->q_prev, b->cost, b->lam, b->nacc, b->trace,
                  b->acc_trace, b->q_trace, b->lam_trace, b->sel, b->pb.Tj, b->pb.Sw, b->pb.J, b->pb.r, b->pb.rp, b->pb.qp, b->pb.dq, b->pb.pred,
                  b->pb.acc};
  for (void* p : ptrs)
    if (p) hipFree(p);
  ctx->mem_bytes -= b->bytes;
  delete b;
  return RK_OK;
}

}  // extern "C"

// validated device description of a trajectory problem set (rk_traj_batch_create, rk_server_add_traj)
static int make_traj_terms(rk_context* ctx, const rk_traj_desc* d, TrajTerms* out) {
  int rc;
  if (!ctx || !d) return fail(RK_EINVAL, "null argument");
  if ((rc = check_robot(ctx, d->robot))) return rc;
  const Robot& r = ctx->robots[d->robot];
  if (!r.chain || (r.na != 6 && r.na != 7) || r.nj > r.na + 2)
    return fail(RK_EINVAL, "trajectory batches take serial arms with 6 or 7 dofs (<= 2 fixed tail joints)");
  if (d->frames < 3 || d->frames > RK_TRAJ_MAXT) return fail(RK_EINVAL, "frames must be 3..64");
  if (d->goal_link < 0 || d->goal_link >= r.nl) return fail(RK_EINVAL, "goal link out of range");
  if (d->n_spheres < 0 || d->n_spheres > RK_TRAJ_MAXS || d->n_obs_spheres < 0 || d->n_obs_spheres > RK_SCENE_MAXS ||
      d->n_obs_boxes < 0 || d->n_obs_boxes > RK_SCENE_MAXB)
    return fail(RK_EINVAL, "at most 64 robot spheres, 64 obstacle spheres and 16 obstacle boxes");
  if (!(d->dt > 0.f)) return fail(RK_EINVAL, "dt must be positive");
  TrajTerms& tt = *out;
  memset(&tt, 0, sizeof(tt));
  tt.robot = d->robot;
  tt.T = d->frames;
  tt.goal_link = d->goal_link;
  for (int i = 0; i < 6; ++i) tt.goal_w[i] = d->goal_weights ? d->goal_weights[i] : 1.f;
  tt.wv = d->vel_weight / d->dt;
  tt.wa = d->acc_weight / (d->dt * d->dt);
  tt.coll_w = d->coll_weight;
  tt.margin = d->coll_margin;
  // spheres sorted by the (depth-renumbered) joint that moves their link, base link first
  std::vector<int> order(d->n_spheres);
  for (int k = 0; k < d->n_spheres; ++k) {
    order[k] = k;
    if (d->sphere_links[k] < 0 || d->sphere_links[k] >= r.nl) return fail(RK_EINVAL, "sphere link out of range");
  }
  std::stable_sort(order.begin(), order.end(),
                   [&](int a, int b) { return r.link_joint[d->sphere_links[a]] < r.link_joint[d->sphere_links[b]]; });
  tt.nsph = d->n_spheres;
  for (int i = 0; i < d->n_spheres; ++i) {
    const int k = order[i];
    tt.sph_joint[i] = r.link_joint[d->sphere_links[k]];
    for (int c = 0; c < 3; ++c) tt.sph_c[i][c] = d->sphere_centers[3 * k + c];
    tt.sph_r[i] = d->sphere_radii[k];
  }
  tt.ns = d->n_obs_spheres;
  tt.nb = d->n_obs_boxes;
  // the kernels read the scene in groups of 8 spheres / 4 boxes: the rest of the last group is
  // far-away padding that never wins the minimum
  for (int i = 0; i < RK_SCENE_MAXS; ++i)
    for (int c = 0; c < 4; ++c) tt.obs_s[i][c] = i < tt.ns ? d->obs_spheres[4 * i + c] : (c < 3 ? RK_SDF_FAR : 0.f);
  for (int i = 0; i < RK_SCENE_MAXB; ++i)
    for (int c = 0; c < 6; ++c) tt.obs_b[i][c] = i < tt.nb ? d->obs_boxes[6 * i + c] : (c < 3 ? RK_SDF_FAR : 0.f);
  if ((rc = scene_geoms(ctx, d->n_geoms, d->geoms, &tt.ng, tt.geom))) return rc;
  // bound rows (TrajectoryLimitTask, forward stencils of orders 0..2): weights and bounds per order and dof
  if (d->limit_orders & ~7) return fail(RK_EINVAL, "limit_orders: bits 0 (position), 1 (velocity), 2 (acceleration)");
  tt.lim_orders = d->limit_orders;
  for (int o = 0; o < 3; ++o) {
    tt.lim_scale[o] = 1.f / powf(d->dt, (float)o);
    if (!((tt.lim_orders >> o) & 1)) continue;
    if (!d->limit_weights || !d->limit_lower || !d->limit_upper)
      return fail(RK_EINVAL, "limit_orders needs limit_weights, limit_lower and limit_upper");
    if (d->frames <= o) return fail(RK_EINVAL, "a forward bound of order o needs more than o frames");
    for (int k = 0; k < r.na; ++k) {
      tt.lim_w[o][k] = d->limit_weights[o * r.na + k];
      tt.lim_lo[o][k] = d->limit_lower[o * r.na + k];
      tt.lim_hi[o][k] = d->limit_upper[o * r.na + k];
      if (!(tt.lim_lo[o][k] <= tt.lim_hi[o][k])) return fail(RK_EINVAL, "limit bounds must be ordered, non-NaN");
    }
  }
  if (d->cg_max_iter < 0 || d->cg_max_iter > 4096) return fail(RK_EINVAL, "cg_max_iter must be 0..4096");
  tt.cg_iters = d->cg_max_iter;
  return RK_OK;
}

// the goal-only term header of a trajectory problem set (the goal pose is its one pose target)
static Terms traj_goal_terms(const rk_traj_desc* d, const TrajTerms& tt) {
  Terms th;
  memset(&th, 0, sizeof(th));
  th.robot = d->robot;
  th.n_pose = 1;
  th.pose_link[0] = d->goal_link;
  for (int k = 0; k < 6; ++k) th.pose_w[0][k] = tt.goal_w[k];
  th.Rf = 6;
  th.R = 6;
  return th;
}

extern "C" {

int rk_traj_batch_create(rk_context* ctx, const rk_traj_desc* d, int32_t capacity, int32_t trace_iters,
                         rk_batch** out) {
  int rc;
  if (!ctx || !d || !out) return fail(RK_EINVAL, "null argument");
  if (capacity <= 0 || trace_iters < 0) return fail(RK_EINVAL, "capacity must be positive, trace_iters >= 0");
  TrajTerms tt;
  if ((rc = make_traj_terms(ctx, d, &tt))) return rc;
  const Robot& r = ctx->robots[d->robot];
  // the batch: na = frames * dofs per problem (SoA [k][stride]), one 7-float goal per problem
  const Terms th = traj_goal_terms(d, tt);
  const int D = d->frames * r.na;
  std::lock_guard<std::mutex> lk(ctx->mu);
  HIP_TRY(hipSetDevice(ctx->device));
  rk_batch* b = new rk_batch();
  b->ctx = ctx;
  b->th = th;
  b->capacity = capacity;
  b->trace_iters = trace_iters;
  b->na = D;
  b->traj_T = d->frames;
  b->traj_na = r.na;
  b->traj_lim = tt.lim_orders != 0 || tt.ng > 0;
  const size_t S = capacity;
  auto A = [&](void** p, size_t bytes) { return dmalloc(ctx, p, bytes, &b->bytes); };
  if ((rc = A((void**)&b->tt, sizeof(TrajTerms))) || (rc = A((void**)&b->td, sizeof(Terms))) ||
      (rc = A((void**)&b->q, sizeof(float) * D * S)) || (rc = A((void**)&b->q_seed, sizeof(float) * D * S)) ||
      (rc = A((void**)&b->pose_tgt, sizeof(float) * 7 * S)) || (rc = A((void**)&b->pos_tgt, sizeof(float) * S)) ||
      (rc = A((void**)&b->q_prev, sizeof(float) * D * S)) || (rc = A((void**)&b->cost, sizeof(float) * S)) ||
      (rc = A((void**)&b->lam, sizeof(float) * S)) || (rc = A((void**)&b->nacc, sizeof(int) * S))) {
    rk_batch_destroy(b);
    return rc;
  }
  if (trace_iters > 0) {
    if ((rc = A((void**)&b->trace, sizeof(float) * (trace_iters + 1) * S)) ||
        (rc = A((void**)&b->acc_trace, sizeof(int) * trace_iters * S)) ||
        (rc = A((void**)&b->q_trace, sizeof(float) * (size_t)trace_iters * D * S)) ||
        (rc = A((void**)&b->lam_trace, sizeof(float) * trace_iters * S))) {
      rk_batch_destroy(b);
      return rc;
    }
  }
  hipMemcpyAsync(b->tt, &tt, sizeof(TrajTerms), hipMemcpyHostToDevice, ctx->stream);
  hipMemcpyAsync(b->td, &th, sizeof(Terms), hipMemcpyHostToDevice, ctx->stream);
  HIP_TRY(hipStreamSynchronize(ctx->stream));
  *out = b;
  return RK_OK;
}

int rk_traj_metrics(rk_batch* b, float* pos_err, float* ang_err, float* min_sd) {
  if (!b || !pos_err || !ang_err || !min_sd) return fail(RK_EINVAL, "null argument");
  if (!b->traj_T) return fail(RK_EINVAL, "not a trajectory batch");
  rk_context* ctx = b->ctx;
  std::lock_guard<std::mutex> lk(ctx->mu);
  HIP_TRY(hipSetDevice(ctx->device));
  const int B = b->n;
  float* dbuf = nullptr;
  HIP_TRY(hipMalloc(&dbuf, sizeof(float) * 3 * B));
  const Problems P = make_problems(b);
  launch_traj_metrics(ctx->stream, b->traj_na, b->tt, P, dbuf, dbuf + B, dbuf + 2 * B);
  hipError_t e = hipGetLastError();
  std::vector<float> h(3 * B);
  if (e == hipSuccess) e = hipMemcpyAsync(h.data(), dbuf, sizeof(float) * 3 * B, hipMemcpyDeviceToHost, ctx->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  hipFree(dbuf);
  if (e != hipSuccess) return fail(RK_EHIP, std::string("rk_traj_metrics: ") + hipGetErrorString(e));
  memcpy(pos_err, h.data(), sizeof(float) * B);
  memcpy(ang_err, h.data() + B, sizeof(float) * B);
  memcpy(min_sd, h.data() + 2 * B, sizeof(float) * B);
  return RK_OK;
}

// host row-major [B][k] -> device SoA [k][stride]
static int upload_soa(rk_context* ctx, float* dst, const float* src, int B, int k, int stride) {
  std::vector<float> tmp((size_t)k * B);
  for (int i = 0; i < B; ++i)
    for (int c = 0; c < k; ++c) tmp[(size_t)c * B + i] = src[(size_t)i * k + c];
  for (int c = 0; c < k; ++c)
    HIP_TRY(hipMemcpyAsync(dst + (size_t)c * stride, tmp.data() + (size_t)c * B, sizeof(float) * B,
                           hipMemcpyHostToDevice, ctx->stream));
  HIP_TRY(hipStreamSynchronize(ctx->stream));
  return RK_OK;
}

}  // extern "C"

template <class T>
static int download_soa(rk_context* ctx, T* dst, const T* src, int B, int k, int stride) {
  std::vector<T> tmp((size_t)k * B);
  for (int c = 0; c < k; ++c)
    HIP_TRY(hipMemcpyAsync(tmp.data() + (size_t)c * B, src + (size_t)c * stride, sizeof(T) * B,
                           hipMemcpyDeviceToHost, ctx->stream));
  HIP_TRY(hipStreamSynchronize(ctx->stream));
  for (int i = 0; i < B; ++i)
    for (int c = 0; c < k; ++c) dst[(size_t)i * k + c] = tmp[(size_t)c * B + i];
  return RK_OK;
}

extern "C" {

int rk_batch_set(rk_batch* b, int32_t B, const float* q, const float* pose_targets, const float* pos_targets,
                 const float* q_prev) {
  if (!b) return fail(RK_EINVAL, "null batch");
  if (B > b->capacity)
    return fail(RK_ECAPACITY, "batch size " + std::to_string(B) + " exceeds the " + std::to_string(b->capacity) +
                                  " this batch was built for");
  if (B <= 0) return fail(RK_EINVAL, "batch size must be positive");
  rk_context* ctx = b->ctx;
  std::lock_guard<std::mutex> lk(ctx->mu);
  HIP_TRY(hipSetDevice(ctx->device));
  int rc;
  b->n = B;
  b->it_done = 0;
  if (q) {
    if ((rc = upload_soa(ctx, b->q, q, B, b->na, b->capacity))) return rc;
    HIP_TRY(hipMemcpyAsync(b->q_seed, b->q, sizeof(float) * b->na * b->capacity, hipMemcpyDeviceToDevice,
                           ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
  }
  if (pose_targets && b->th.n_pose &&
      (rc = upload_soa(ctx, b->pose_tgt, pose_targets, B, b->th.n_pose * 7, b->capacity)))
    return rc;
  if (pos_targets && b->th.n_pos && (rc = upload_soa(ctx, b->pos_tgt, pos_targets, B, b->th.n_pos * 3, b->capacity)))
    return rc;
  if (q_prev && (rc = upload_soa(ctx, b->q_prev, q_prev, B, b->na, b->capacity))) return rc;
  return RK_OK;
}

static int batch_select(rk_batch* src, const rk_batch* scores, int32_t width_in, rk_batch* dst, int32_t width_out,
                        int32_t* src_rows);

int rk_batch_select(rk_batch* src, int32_t width_in, rk_batch* dst, int32_t width_out, int32_t* src_rows) {
  return batch_select(src, src, width_in, dst, width_out, src_rows);
}

int rk_batch_select_scored(rk_batch* src, const rk_batch* scores, int32_t width_in, rk_batch* dst, int32_t width_out,
                           int32_t* src_rows) {
  if (!scores) return fail(RK_EINVAL, "null score batch");
  if (scores == dst) return fail(RK_EINVAL, "the score batch cannot be the target");
  if (scores->ctx != src->ctx || scores->n != src->n || scores->traj_T)
    return fail(RK_EINVAL, "the score batch must hold one score per source row (same context and size)");
  return batch_select(src, scores, width_in, dst, width_out, src_rows);
}

int rk_batch_copy_q(rk_batch* src, rk_batch* dst) {
  if (!src || !dst || src == dst) return fail(RK_EINVAL, "two distinct batches needed");
  if (src->ctx != dst->ctx || src->na != dst->na || src->traj_T != dst->traj_T)
    return fail(RK_EINVAL, "batches of different contexts or shapes");
  if (src->n > dst->capacity)
    return fail(RK_ECAPACITY, "batch size " + std::to_string(src->n) + " exceeds the " + std::to_string(dst->capacity) +
                                  " this batch was built for");
  rk_context* ctx = src->ctx;
  std::lock_guard<std::mutex> lk(ctx->mu);
  HIP_TRY(hipSetDevice(ctx->device));
  if (src->n > 0)  // SoA: na rows of n floats, each batch with its own row pitch (its capacity)
    HIP_TRY(hipMemcpy2DAsync(dst->q, sizeof(float) * dst->capacity, src->q, sizeof(float) * src->capacity,
                             sizeof(float) * src->n, src->na, hipMemcpyDeviceToDevice, ctx->stream));
  dst->n = src->n;
  dst->it_done = 0;
  return RK_OK;
}

static int batch_select(rk_batch* src, const rk_batch* scores, int32_t width_in, rk_batch* dst, int32_t width_out,
                        int32_t* src_rows) {
  if (!src || !dst) return fail(RK_EINVAL, "null batch");
  if (src == dst) return fail(RK_EINVAL, "select needs two batches (the source is read while the target is written)");
  if (src->ctx != dst->ctx) return fail(RK_EINVAL, "batches of different contexts");
  if (src->traj_T || dst->traj_T) return fail(RK_EINVAL, "trajectory batches have no candidate select");
  if (width_in < 1 || width_in > WAVE || width_out < 1 || width_out > width_in)
    return fail(RK_EINVAL, "need 1 <= width_out <= width_in <= 64");
  if (src->n <= 0 || src->n % width_in)
    return fail(RK_EINVAL, "the source batch size " + std::to_string(src->n) + " is not a multiple of width_in " +
                               std::to_string(width_in));
  const Terms &a = src->th, &c = dst->th;
  if (a.robot != c.robot || a.n_pose != c.n_pose || a.n_pos != c.n_pos || a.n_rot != c.n_rot || src->na != dst->na)
    return fail(RK_EINVAL, "source and target batches have different robots or term structures");
  const int n_inst = src->n / width_in, n_out = n_inst * width_out;
  if (n_out > dst->capacity)
    return fail(RK_ECAPACITY, "batch size " + std::to_string(n_out) + " exceeds the " + std::to_string(dst->capacity) +
                                  " this batch was built for");
  rk_context* ctx = src->ctx;
  std::lock_guard<std::mutex> lk(ctx->mu);
  HIP_TRY(hipSetDevice(ctx->device));
  int rc;
  if (!dst->sel && (rc = dmalloc(ctx, (void**)&dst->sel, sizeof(int) * dst->capacity, &dst->bytes))) return rc;
  SelectArgs sa{scores->cost,  src->q,         src->pose_tgt, src->pos_tgt, src->rot_tgt, src->q_prev,
                dst->q,        dst->q_seed,    dst->pose_tgt, dst->pos_tgt, dst->rot_tgt, dst->q_prev,
                dst->cost,     dst->sel,      src->capacity,  dst->capacity, width_in,     width_out,    src->na,
                a.n_pose * 7,  a.n_pos * 3,    a.n_rot * 4};
  hipLaunchKernelGGL(k_select_gather, dim3(n_inst), dim3(WAVE), 0, ctx->stream, sa);
  HIP_TRY(hipGetLastError());
  dst->n = n_out;
  dst->it_done = 0;
  if (src_rows) {
    HIP_TRY(hipMemcpyAsync(src_rows, dst->sel, sizeof(int) * n_out, hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
  }
  return RK_OK;
}

int rk_batch_set_rotations(rk_batch* b, int32_t B, const float* rot_targets) {
  if (!b || !rot_targets) return fail(RK_EINVAL, "null argument");
  if (!b->th.n_rot) return fail(RK_EINVAL, "the batch has no rotation targets");
  if (B != b->n) return fail(RK_EINVAL, "rotation targets must cover the live batch (rk_batch_set first)");
  rk_context* ctx = b->ctx;
  std::lock_guard<std::mutex> lk(ctx->mu);
  HIP_TRY(hipSetDevice(ctx->device));
  return upload_soa(ctx, b->rot_tgt, rot_targets, B, b->th.n_rot * 4, b->capacity);
}

int rk_batch_solve(rk_batch* b, int32_t iters, const rk_lm_params* p, int32_t mode, int32_t tpp, float* kernel_ms) {
  if (!b || !p) return fail(RK_EINVAL, "null argument");
  if (b->n <= 0) return fail(RK_EINVAL, "batch has no problems (call rk_batch_set)");
  if (iters < 0) return fail(RK_EINVAL, "iters must be >= 0");
  if (b->trace && iters > b->trace_iters) return fail(RK_EINVAL, "iters exceeds the batch's trace length");
  rk_context* ctx = b->ctx;
  std::lock_guard<std::mutex> lk(ctx->mu);
  HIP_TRY(hipSetDevice(ctx->device));
  const LMParams lp = to_lp(p);
  int rc = RK_OK;
  const bool from_seed = (mode & RK_FROM_SEED) != 0;
  const bool cont = (mode & RK_CONTINUE) != 0;
  const bool nowait = (mode & RK_NOWAIT) != 0;
  // the start mark of the synchronous call's kernel time; an enqueued call records none (an event
  // between back-to-back launches cost about 1 % of an uneven solve)
  if (!nowait) HIP_TRY(hipEventRecord(ctx->ev0, ctx->stream));
  mode &= ~(RK_FROM_SEED | RK_CONTINUE | RK_NOWAIT);
  if (cont && from_seed) return fail(RK_EINVAL, "RK_CONTINUE and RK_FROM_SEED exclude each other");
  const int it0 = cont ? b->it_done : 0;
  if (b->trace && it0 + iters > b->trace_iters) return fail(RK_EINVAL, "iterations exceed the batch's trace length");
  if (b->traj_T) {
    if (mode != RK_MODE_FUSED) return fail(RK_EINVAL, "trajectory batches run the fused kernel only");
    rc = launch_traj(b, make_problems(b, from_seed), lp, it0, iters);
  } else if (iters == 0 && (mode == RK_MODE_PLAIN || mode == RK_MODE_FUSED)) {
    rc = eval_cost(b, from_seed);
  } else if (mode == RK_MODE_PLAIN) {
    rc = solve_plain(b, it0, iters, lp, from_seed);
  } else if (mode == RK_MODE_FUSED) {
    int v;
    if ((rc = pick_variant(b, tpp, &v))) return rc;
    rc = launch_variant(ctx->stream, v, make_problems(b, from_seed), lp, it0, iters);
  } else {
    return fail(RK_EINVAL, "mode must be RK_MODE_FUSED or RK_MODE_PLAIN");
  }
  if (rc) return rc;
  b->it_done = it0 + iters;
  if (nowait) return RK_OK;  // enqueued; a later synchronous call waits
  HIP_TRY(hipEventRecord(ctx->ev1, ctx->stream));
  HIP_TRY(hipEventSynchronize(ctx->ev1));
  if (kernel_ms) HIP_TRY(hipEventElapsedTime(kernel_ms, ctx->ev0, ctx->ev1));
  return RK_OK;
}

int rk_batch_get(rk_batch* b, float* q, float* cost, float* lambda, float* cost_trace, int32_t* accept_trace) {
  if (!b) return fail(RK_EINVAL, "null batch");
  rk_context* ctx = b->ctx;
  std::lock_guard<std::mutex> lk(ctx->mu);
  HIP_TRY(hipSetDevice(ctx->device));
  const int B = b->n;
  int rc;
  if (q && (rc = download_soa(ctx, q, b->q, B, b->na, b->capacity))) return rc;
  if (cost) {
    HIP_TRY(hipMemcpyAsync(cost, b->cost, sizeof(float) * B, hipMemcpyDeviceToHost, ctx->stream));
  }
  if (lambda) {
    HIP_TRY(hipMemcpyAsync(lambda, b->lam, sizeof(float) * B, hipMemcpyDeviceToHost, ctx->stream));
  }
  HIP_TRY(hipStreamSynchronize(ctx->stream));
  if (cost_trace) {
    if (!b->trace) return fail(RK_EINVAL, "batch was created without a trace");
    if ((rc = download_soa(ctx, cost_trace, b->trace, B, b->trace_iters + 1, b->capacity))) return rc;
  }
  if (accept_trace) {
    if (!b->acc_trace) return fail(RK_EINVAL, "batch was created without a trace");
    if ((rc = download_soa<int>(ctx, accept_trace, b->acc_trace, B, b->trace_iters, b->capacity)))
      return rc;
  }
  return RK_OK;
}

int rk_batch_advance(rk_batch* b) {
  if (!b) return fail(RK_EINVAL, "null batch");
  if (b->n <= 0) return fail(RK_EINVAL, "batch has no problems (call rk_batch_set)");
  rk_context* ctx = b->ctx;
  std::lock_guard<std::mutex> lk(ctx->mu);
  HIP_TRY(hipSetDevice(ctx->device));
  const size_t bytes = sizeof(float) * b->na * b->capacity;
  HIP_TRY(hipMemcpyAsync(b->q_prev, b->q, bytes, hipMemcpyDeviceToDevice, ctx->stream));
  HIP_TRY(hipMemcpyAsync(b->q_seed, b->q, bytes, hipMemcpyDeviceToDevice, ctx->stream));
  return RK_OK;
}

int rk_batch_get_iterates(rk_batch* b, float* q_trace, float* lam_trace) {
  if (!b || !q_trace || !lam_trace) return fail(RK_EINVAL, "null argument");
  if (!b->q_trace) return fail(RK_EINVAL, "batch was created without a trace");
  rk_context* ctx = b->ctx;
  std::lock_guard<std::mutex> lk(ctx->mu);
  HIP_TRY(hipSetDevice(ctx->device));
  int rc;
  if ((rc = download_soa<float>(ctx, q_trace, b->q_trace, b->n, b->trace_iters * b->na, b->capacity))) return rc;
  if ((rc = download_soa<float>(ctx, lam_trace, b->lam_trace, b->n, b->trace_iters, b->capacity))) return rc;
  return RK_OK;
}

int rk_batch_metrics(rk_batch* b, float pos_thr, float ang_thr, float* pos_err, float* ang_err, int32_t* success) {
  if (!b || !pos_err || !ang_err || !success) return fail(RK_EINVAL, "null argument");
  if (b->th.n_pose == 0) return fail(RK_EINVAL, "metrics need pose targets");
  if (b->traj_T) return fail(RK_EINVAL, "trajectory batches report rk_traj_metrics");
  rk_context* ctx = b->ctx;
  std::lock_guard<std::mutex> lk(ctx->mu);
  HIP_TRY(hipSetDevice(ctx->device));
  const int B = b->n, k = b->th.n_pose;
  float *dpe = nullptr, *dae = nullptr;
  int* dok = nullptr;
  HIP_TRY(hipMalloc(&dpe, sizeof(float) * B * k));
  HIP_TRY(hipMalloc(&dae, sizeof(float) * B * k));
  HIP_TRY(hipMalloc(&dok, sizeof(int) * B));
  hipLaunchKernelGGL(k_ik_metrics, dim3((B + 127) / 128), dim3(128), 0, ctx->stream, make_problems(b), pos_thr,
                     ang_thr, dpe, dae, dok);
  hipMemcpyAsync(pos_err, dpe, sizeof(float) * B * k, hipMemcpyDeviceToHost, ctx->stream);
  hipMemcpyAsync(ang_err, dae, sizeof(float) * B * k, hipMemcpyDeviceToHost, ctx->stream);
  hipMemcpyAsync(success, dok, sizeof(int) * B, hipMemcpyDeviceToHost, ctx->stream);
  hipError_t e = hipStreamSynchronize(ctx->stream);
  hipFree(dpe);
  hipFree(dae);
  hipFree(dok);
  if (e != hipSuccess) return fail(RK_EHIP, std::string("rk_batch_metrics: ") + hipGetErrorString(e));
  return RK_OK;
}

int rk_batch_metrics_noalloc(rk_batch* b, float pos_thr, float ang_thr, float* pos_err, float* ang_err,
                             int32_t* success) {
  if (!b || !pos_err || !ang_err || !success) return fail(RK_EINVAL, "null argument");
  if (b->th.n_pose == 0) return fail(RK_EINVAL, "metrics need pose targets");
  if (b->traj_T) return fail(RK_EINVAL, "trajectory batches report rk_traj_metrics");
  const int B = b->n, k = b->th.n_pose;
  if (B > RK_NOALLOC_MAX_PROBLEMS)
    return fail(RK_EINVAL, "rk_batch_metrics_noalloc: " + std::to_string(B) + " problems exceed the static staging of " +
                               std::to_string(RK_NOALLOC_MAX_PROBLEMS));
  rk_context* ctx = b->ctx;
  std::lock_guard<std::mutex> lk(ctx->mu);
  HIP_TRY(hipSetDevice(ctx->device));
  float* const base = scratch_of(ctx->device);
  if (!base) return fail(RK_EHIP, "rk_batch_metrics_noalloc: no staging address for this device");
  float* const dpe = base;
  float* const dae = base + RK_NOALLOC_PAIRS;
  int* const dok = reinterpret_cast<int*>(base + 2 * RK_NOALLOC_PAIRS);
  hipLaunchKernelGGL(k_ik_metrics, dim3((B + 127) / 128), dim3(128), 0, ctx->stream, make_problems(b), pos_thr,
                     ang_thr, dpe, dae, dok);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipMemcpyAsync(pos_err, dpe, sizeof(float) * B * k, hipMemcpyDeviceToHost, ctx->stream));
  HIP_TRY(hipMemcpyAsync(ang_err, dae, sizeof(float) * B * k, hipMemcpyDeviceToHost, ctx->stream));
  HIP_TRY(hipMemcpyAsync(success, dok, sizeof(int) * B, hipMemcpyDeviceToHost, ctx->stream));
  const hipError_t e = hipStreamSynchronize(ctx->stream);
  if (e != hipSuccess) return fail(RK_EHIP, std::string("rk_batch_metrics_noalloc: ") + hipGetErrorString(e));
  return RK_OK;
}

// ---------------------------------------------------------------------------------------------
// uneven batch (persistent kernel)
// ---------------------------------------------------------------------------------------------
int rk_group_create(rk_context* ctx, int32_t nb, rk_batch* const* batches, rk_group** out) {
  if (!ctx || nb <= 0 || !batches || !out) return fail(RK_EINVAL, "bad argument to rk_group_create");
  for (int i = 0; i < nb; ++i)
    if (!batches[i] || batches[i]->ctx != ctx) return fail(RK_EINVAL, "batch of another context or null");
  if (nb > 32) return fail(RK_EINVAL, "at most 32 batches per group");
  std::lock_guard<std::mutex> lk(ctx->mu);
  HIP_TRY(hipSetDevice(ctx->device));
  rk_group* g = new rk_group();
  g->ctx = ctx;
  g->batches.assign(batches, batches + nb);
  g->tpp.assign(nb, 0);
  for (int i = 0; i < nb; ++i) (batches[i]->traj_T ? g->traj : g->ik).push_back(i);
  const int nik = std::max<int>(1, (int)g->ik.size()), ntj = std::max<int>(1, (int)g->traj.size());
  auto release = [&]() {
    dfree(ctx, g->buckets_dev, sizeof(Bucket) * nik);
    dfree(ctx, g->buckets_graph, sizeof(Bucket) * nik);
    dfree(ctx, g->traj_dev, sizeof(TrajBucket) * ntj);
    dfree(ctx, g->traj_graph, sizeof(TrajBucket) * ntj);
    dfree(ctx, g->counter, sizeof(int) * RK_QSTRIDE * NVARIANT);
    delete g;
  };
  int rc;
  if ((rc = dmalloc(ctx, (void**)&g->buckets_dev, sizeof(Bucket) * nik, &g->bytes)) ||
      (rc = dmalloc(ctx, (void**)&g->buckets_graph, sizeof(Bucket) * nik, &g->bytes)) ||
      (rc = dmalloc(ctx, (void**)&g->traj_dev, sizeof(TrajBucket) * ntj, &g->bytes)) ||
      (rc = dmalloc(ctx, (void**)&g->traj_graph, sizeof(TrajBucket) * ntj, &g->bytes)) ||
      (rc = dmalloc(ctx, (void**)&g->counter, sizeof(int) * RK_QSTRIDE * NVARIANT, &g->bytes))) {
    release();
    return rc;
  }
  // (creation is never captured into a graph)
  if (hipMemsetAsync(g->counter, 0, sizeof(int) * RK_QSTRIDE * NVARIANT, ctx->stream) != hipSuccess ||
      hipStreamSynchronize(ctx->stream) != hipSuccess) {
    release();
    return fail(RK_EHIP, "could not clear the group's queue counters");
  }
  *out = g;
  return RK_OK;
}

int rk_group_destroy(rk_group* g) {
  if (!g) return RK_OK;
  rk_context* ctx = g->ctx;
  hipSetDevice(ctx->device);
  hipStreamSynchronize(ctx->stream);
  if (g->graph) hipGraphExecDestroy(g->graph);
  for (int k = 0; k < NVARIANT + RK_GROUP_TRAJ_MAX; ++k) {
    if (g->aux[k]) hipStreamDestroy(g->aux[k]);
    if (g->ev_join[k]) hipEventDestroy(g->ev_join[k]);
    if (g->ev_k0[k]) hipEventDestroy(g->ev_k0[k]);
    if (g->ev_k1[k]) hipEventDestroy(g->ev_k1[k]);
  }
  if (g->ev_fork) hipEventDestroy(g->ev_fork);
  hipFree(g->buckets_dev);
  hipFree(g->buckets_graph);
  hipFree(g->traj_dev);
  hipFree(g->traj_graph);
  hipFree(g->counter);
  ctx->mem_bytes -= g->bytes;
  delete g;
  return RK_OK;
}

int rk_group_set_tuning(rk_group* g, const int32_t* tpp, int32_t wg_per_cu, int32_t iters_per_launch) {
  if (!g) return fail(RK_EINVAL, "null group");
  if (wg_per_cu < 0 || iters_per_launch < 0) return fail(RK_EINVAL, "negative tuning value");
  if (tpp) {  // every entry is checked before any is applied
    for (size_t i = 0; i < g->batches.size(); ++i) {
      if (g->batches[i]->traj_T) {  // trajectories have one kernel form
        if (tpp[i] != 0 && tpp[i] != RK_TPP_TRAJ) return fail(RK_EINVAL, "a trajectory batch takes tpp 0");
        continue;
      }
      int v;
      int rc = pick_variant(g->batches[i], tpp[i], &v);
      if (rc) return rc;
    }
    for (size_t i = 0; i < g->batches.size(); ++i) g->tpp[i] = tpp[i];
  }
  g->wg_per_cu = wg_per_cu;
  g->iters_per_launch = iters_per_launch;
  if (g->graph) {
    hipGraphExecDestroy(g->graph);
    g->graph = nullptr;
  }
  return RK_OK;
}

}  // extern "C"

struct GroupLaunch {
  const Bucket* table = nullptr;  // device bucket table the launches read
  // per variant present: first bucket, bucket count, tiles, grid, dynamic LDS, batch mask
  std::vector<int> var, first, nb, tiles, grid;
  std::vector<unsigned> mask;
  std::vector<size_t> lds;
  // trajectory classes: dofs, bound-row / geometry kernel, first table entry, entries, workgroups
  const TrajBucket* ttable = nullptr;
  std::vector<int> t_na, t_lim, t_first, t_nb, t_blocks;
  std::vector<unsigned> t_mask;
};

template <int V>
static int variant_grid(rk_context* ctx, int wg_per_cu, size_t lds, int tiles, int* grid) {
  int per_cu = wg_per_cu;
  if (per_cu <= 0) {
    int occ = 0;
    HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, (const void*)k_persistent<V>, WAVE, lds));
    per_cu = std::max(1, occ);
  }
  *grid = std::max(1, std::min(tiles, per_cu * ctx->cu_count));
  return RK_OK;
}

static int variant_grid_any(rk_context* ctx, int v, int wg_per_cu, size_t lds, int tiles, int* grid) {
  switch (v) {
    case V_S4: return variant_grid<V_S4>(ctx, wg_per_cu, lds, tiles, grid);
    case V_S8: return variant_grid<V_S8>(ctx, wg_per_cu, lds, tiles, grid);
    case V_S16: return variant_grid<V_S16>(ctx, wg_per_cu, lds, tiles, grid);
    case V_L16: return variant_grid<V_L16>(ctx, wg_per_cu, lds, tiles, grid);
    case V_L32: return variant_grid<V_L32>(ctx, wg_per_cu, lds, tiles, grid);
    case V_L64: return variant_grid<V_L64>(ctx, wg_per_cu, lds, tiles, grid);
    case V_T5: return variant_grid<V_T5>(ctx, wg_per_cu, lds, tiles, grid);
    case V_T6: return variant_grid<V_T6>(ctx, wg_per_cu, lds, tiles, grid);
    case V_T7: return variant_grid<V_T7>(ctx, wg_per_cu, lds, tiles, grid);
    case V_T8: return variant_grid<V_T8>(ctx, wg_per_cu, lds, tiles, grid);
    case V_W40: return variant_grid<V_W40>(ctx, wg_per_cu, lds, tiles, grid);
    case V_W012: return variant_grid<V_W012>(ctx, wg_per_cu, lds, tiles, grid);
    case V_H40: return variant_grid<V_H40>(ctx, wg_per_cu, lds, tiles, grid);
    case V_HS40: return variant_grid<V_HS40>(ctx, wg_per_cu, lds, tiles, grid);
    case V_Q6: return variant_grid<V_Q6>(ctx, wg_per_cu, lds, tiles, grid);
    case V_Q7: return variant_grid<V_Q7>(ctx, wg_per_cu, lds, tiles, grid);
    case V_HS012: return variant_grid<V_HS012>(ctx, wg_per_cu, lds, tiles, grid);
    case V_P5: return variant_grid<V_P5>(ctx, wg_per_cu, lds, tiles, grid);
    case V_P6: return variant_grid<V_P6>(ctx, wg_per_cu, lds, tiles, grid);
    case V_P7: return variant_grid<V_P7>(ctx, wg_per_cu, lds, tiles, grid);
    case V_P8: return variant_grid<V_P8>(ctx, wg_per_cu, lds, tiles, grid);
    case V_TC6: return variant_grid<V_TC6>(ctx, wg_per_cu, lds, tiles, grid);
    case V_TC7: return variant_grid<V_TC7>(ctx, wg_per_cu, lds, tiles, grid);
    case V_PC6: return variant_grid<V_PC6>(ctx, wg_per_cu, lds, tiles, grid);
    case V_PC7: return variant_grid<V_PC7>(ctx, wg_per_cu, lds, tiles, grid);
    default: return variant_grid<V_H012>(ctx, wg_per_cu, lds, tiles, grid);
  }
}

static void launch_persistent(int v, dim3 grid, size_t lds, hipStream_t st, const Bucket* bk, int nb, int total,
                              int* counter, const LMParams& lp, int it0, int iters) {
  switch (v) {
#define RK_LAUNCH_P(VV) \
  case VV: hipLaunchKernelGGL(k_persistent<VV>, grid, dim3(WAVE), lds, st, bk, nb, total, counter, lp, it0, iters); break;
    RK_LAUNCH_P(V_S4)
    RK_LAUNCH_P(V_S8)
    RK_LAUNCH_P(V_S16)
    RK_LAUNCH_P(V_L16)
    RK_LAUNCH_P(V_L32)
    RK_LAUNCH_P(V_L64)
    RK_LAUNCH_P(V_T5)
    RK_LAUNCH_P(V_T6)
    RK_LAUNCH_P(V_T7)
    RK_LAUNCH_P(V_T8)
    RK_LAUNCH_P(V_W40)
    RK_LAUNCH_P(V_W012)
    RK_LAUNCH_P(V_H40)
    RK_LAUNCH_P(V_H012)
    RK_LAUNCH_P(V_HS40)
    RK_LAUNCH_P(V_HS012)
    RK_LAUNCH_P(V_Q6)
    RK_LAUNCH_P(V_Q7)
    RK_LAUNCH_P(V_P5)
    RK_LAUNCH_P(V_P6)
    RK_LAUNCH_P(V_P7)
    RK_LAUNCH_P(V_P8)
    RK_LAUNCH_P(V_TC6)
    RK_LAUNCH_P(V_TC7)
    RK_LAUNCH_P(V_PC6)
    RK_LAUNCH_P(V_PC7)
#undef RK_LAUNCH_P
    default: break;
  }
}

// upload the bucket table into `table` (outside any capture) and size each variant's persistent
// grid.  Eager solves upload into g->buckets_dev; a capture uploads into g->buckets_graph, which
// only the next capture rewrites, so a later eager solve (other seeds flag, other sizes) cannot
// change what a replay of the captured graph reads.
static int group_prepare(rk_group* g, GroupLaunch* gl, bool from_seed, Bucket* table, TrajBucket* ttable) {
  rk_context* ctx = g->ctx;
  const size_t nik = g->ik.size();
  std::vector<Bucket> bk(nik);
  for (size_t i = 0; i < nik; ++i) {
    rk_batch* b = g->batches[g->ik[i]];
    int v;
    int rc = pick_variant(b, g->tpp[g->ik[i]], &v);
    if (rc) return rc;
    bk[i].variant = v;
    bk[i].P = make_problems(b, from_seed);
    const int ppw = variant_ppw(v);
    bk[i].tiles = (b->n + ppw - 1) / ppw;
  }
  // group by variant; within a variant, big problems first (longest-processing-time order)
  std::vector<int> order(nik);
  for (size_t i = 0; i < order.size(); ++i) order[i] = (int)i;
  auto ikb = [&](int i) { return g->batches[g->ik[i]]; };
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) {
    if (bk[a].variant != bk[b].variant) return bk[a].variant < bk[b].variant;
    return ikb(a)->na * ikb(a)->th.R > ikb(b)->na * ikb(b)->th.R;
  });
  std::vector<Bucket> sorted(nik);
  *gl = GroupLaunch();
  int base = 0;
  for (size_t i = 0; i < order.size(); ++i) {
    sorted[i] = bk[order[i]];
    const int v = sorted[i].variant;
    if (gl->var.empty() || gl->var.back() != v) {
      gl->var.push_back(v);
      gl->first.push_back((int)i);
      gl->nb.push_back(0);
      gl->tiles.push_back(0);
      gl->mask.push_back(0u);
      base = 0;
    }
    gl->mask.back() |= 1u << g->ik[order[i]];
    sorted[i].tile_base = base;
    base += sorted[i].tiles;
    gl->nb.back() += 1;
    gl->tiles.back() = base;
  }
  gl->table = table;
  if (nik)
    HIP_TRY(hipMemcpyAsync(table, sorted.data(), sizeof(Bucket) * sorted.size(), hipMemcpyHostToDevice,
                           ctx->stream));
  // trajectory batches by arm class (6 / 7 dofs; bound rows or geometries), longest horizon first
  // within a class: one table entry per batch, consecutive workgroup ranges
  std::vector<int> tord(g->traj);
  auto tclass = [&](int i) { return (g->batches[i]->traj_na == 7 ? 2 : 0) + (g->batches[i]->traj_lim ? 1 : 0); };
  std::stable_sort(tord.begin(), tord.end(), [&](int a, int b) {
    if (tclass(a) != tclass(b)) return tclass(a) < tclass(b);
    return g->batches[a]->traj_T > g->batches[b]->traj_T;
  });
  std::vector<TrajBucket> tb(tord.size());
  int blocks = 0;
  for (size_t i = 0; i < tord.size(); ++i) {
    rk_batch* b = g->batches[tord[i]];
    if (i == 0 || tclass(tord[i]) != tclass(tord[i - 1])) {
      gl->t_na.push_back(b->traj_na);
      gl->t_lim.push_back(b->traj_lim ? 1 : 0);
      gl->t_first.push_back((int)i);
      gl->t_nb.push_back(0);
      gl->t_blocks.push_back(0);
      gl->t_mask.push_back(0u);
      blocks = 0;
    }
    memset(&tb[i], 0, sizeof(TrajBucket));
    tb[i].P = make_problems(b, from_seed);
    tb[i].tt = b->tt;
    tb[i].block0 = blocks;
    blocks += b->n;
    gl->t_nb.back() += 1;
    gl->t_blocks.back() = blocks;
    gl->t_mask.back() |= 1u << tord[i];
  }
  gl->ttable = ttable;
  if (!tb.empty())
    HIP_TRY(hipMemcpyAsync(ttable, tb.data(), sizeof(TrajBucket) * tb.size(), hipMemcpyHostToDevice, ctx->stream));
  HIP_TRY(hipStreamSynchronize(ctx->stream));
  const size_t nkern = gl->var.size() + gl->t_na.size();
  for (size_t k = 0; k < nkern; ++k) {
    if (!g->aux[k]) HIP_TRY(hipStreamCreateWithFlags(&g->aux[k], hipStreamNonBlocking));
    if (!g->ev_join[k]) HIP_TRY(hipEventCreateWithFlags(&g->ev_join[k], hipEventDisableTiming));
    if (!g->ev_k0[k]) HIP_TRY(hipEventCreate(&g->ev_k0[k]));
    if (!g->ev_k1[k]) HIP_TRY(hipEventCreate(&g->ev_k1[k]));
  }
  if (!g->ev_fork) HIP_TRY(hipEventCreateWithFlags(&g->ev_fork, hipEventDisableTiming));
  for (size_t k = 0; k < gl->var.size(); ++k) {
    const size_t lds = 0;  // the persistent kernels' LDS is static (variant_lds bytes)
    int grid;
    int rc = variant_grid_any(ctx, gl->var[k], g->wg_per_cu, lds, gl->tiles[k], &grid);
    if (rc) return rc;
    gl->lds.push_back(lds);
    gl->grid.push_back(grid);
  }
  return RK_OK;
}

// enqueue the solve on the context stream: fork one stream per variant, each runs its
// persistent kernel over its queue, join; or, for a group of register-resident variants outside
// the timed path, one k_persistent_reg launch.  Capturable: no allocation, copy or memset.
// the trajectory classes of a group solve, on their own streams forked from the context stream
// (parallel branches of a captured graph), joined back to it
static int group_launch_traj(rk_group* g, const GroupLaunch& gl, int iters, const LMParams& lp, bool timed) {
  rk_context* ctx = g->ctx;
  for (size_t c = 0; c < gl.t_na.size(); ++c) {
    const size_t k = gl.var.size() + c;
    hipStream_t st = g->aux[k];
    HIP_TRY(hipStreamWaitEvent(st, g->ev_fork, 0));
    if (timed) HIP_TRY(hipEventRecord(g->ev_k0[k], st));
    const dim3 grid(gl.t_blocks[c]), blk(2 * WAVE);
    const TrajBucket* tb = gl.ttable + gl.t_first[c];
    const int nb = gl.t_nb[c];
    if (gl.t_na[c] == 6 && gl.t_lim[c])
      hipLaunchKernelGGL((k_traj_multi<6, 8, true>), grid, blk, 0, st, tb, nb, lp, 0, iters);
    else if (gl.t_na[c] == 6)
      hipLaunchKernelGGL((k_traj_multi<6, 8>), grid, blk, 0, st, tb, nb, lp, 0, iters);
    else if (gl.t_lim[c])
      hipLaunchKernelGGL((k_traj_multi<7, 9, true>), grid, blk, 0, st, tb, nb, lp, 0, iters);
    else
      hipLaunchKernelGGL((k_traj_multi<7, 9>), grid, blk, 0, st, tb, nb, lp, 0, iters);
    HIP_TRY(hipGetLastError());
    if (timed) HIP_TRY(hipEventRecord(g->ev_k1[k], st));
    HIP_TRY(hipEventRecord(g->ev_join[k], st));
  }
  return RK_OK;
}

static int group_join_traj(rk_group* g, const GroupLaunch& gl) {
  for (size_t c = 0; c < gl.t_na.size(); ++c)
    HIP_TRY(hipStreamWaitEvent(g->ctx->stream, g->ev_join[gl.var.size() + c], 0));
  return RK_OK;
}

static int group_launch(rk_group* g, const GroupLaunch& gl, int iters, const LMParams& lp, bool timed) {
  rk_context* ctx = g->ctx;
  const int chunk = g->iters_per_launch > 0 ? g->iters_per_launch : std::max(1, iters);
  HIP_TRY(hipEventRecord(g->ev_fork, ctx->stream));
  int rc;
  if ((rc = group_launch_traj(g, gl, iters, lp, timed))) return rc;  // the trajectories start first
  // the variant with the most tiles is enqueued first: its workgroups reach the CUs first, and the
  // few wavefronts of the others fill in around them (about 1 % faster on the uneven benchmark
  // than the thread-per-problem kernels first)
  size_t order[NVARIANT];
  for (size_t k = 0; k < gl.var.size(); ++k) order[k] = k;
  std::stable_sort(order, order + gl.var.size(), [&](size_t a, size_t b) { return gl.tiles[a] > gl.tiles[b]; });
  bool mix = !timed && gl.var.size() > 1;
  for (size_t k = 0; k < gl.var.size(); ++k) mix = mix && variant_mixable(gl.var[k]);
  if (mix) {
    // one launch: block ranges in the same order (the variant with the most tiles first)
    MixArgs ma;
    memset(&ma, 0, sizeof(ma));
    ma.nslots = (int)gl.var.size();
    int blocks = 0;
    for (size_t kk = 0; kk < gl.var.size(); ++kk) {
      const size_t k = order[kk];
      ma.variant[kk] = gl.var[k];
      ma.block0[kk] = blocks;
      ma.nblocks[kk] = gl.grid[k];
      ma.first[kk] = gl.first[k];
      ma.nb[kk] = gl.nb[k];
      ma.tiles[kk] = gl.tiles[k];
      ma.ctr[kk] = (int)k;
      blocks += gl.grid[k];
    }
    for (int it0 = 0; it0 < std::max(1, iters); it0 += chunk) {
      const int n_it = std::max(0, std::min(chunk, iters - it0));
      hipLaunchKernelGGL(k_persistent_reg, dim3(blocks), dim3(WAVE), 0, ctx->stream, gl.table, ma,
                         g->counter, lp, it0, n_it);
      if (iters == 0) break;
    }
    HIP_TRY(hipGetLastError());
    return group_join_traj(g, gl);
  }
  for (size_t kk = 0; kk < gl.var.size(); ++kk) {
    const size_t k = order[kk];
    hipStream_t st = gl.var.size() == 1 ? ctx->stream : g->aux[k];
    if (st != ctx->stream) HIP_TRY(hipStreamWaitEvent(st, g->ev_fork, 0));
    if (timed) HIP_TRY(hipEventRecord(g->ev_k0[k], st));
    for (int it0 = 0; it0 < std::max(1, iters); it0 += chunk) {
      const int n_it = std::max(0, std::min(chunk, iters - it0));
      launch_persistent(gl.var[k], dim3(gl.grid[k]), gl.lds[k], st, gl.table + gl.first[k],
                        gl.nb[k], gl.tiles[k], g->counter + RK_QSTRIDE * k, lp, it0, n_it);
      if (iters == 0) break;
    }
    if (timed) HIP_TRY(hipEventRecord(g->ev_k1[k], st));
    if (st != ctx->stream) {
      HIP_TRY(hipEventRecord(g->ev_join[k], st));
      HIP_TRY(hipStreamWaitEvent(ctx->stream, g->ev_join[k], 0));
    }
  }
  HIP_TRY(hipGetLastError());
  return group_join_traj(g, gl);
}

extern "C" {

int rk_group_solve(rk_group* g, int32_t iters, const rk_lm_params* p, int32_t flags, float* kernel_ms) {
  const bool use_graph = (flags & RK_USE_GRAPH) != 0;
  const bool from_seed = (flags & RK_FROM_SEED) != 0;
  if (!g || !p) return fail(RK_EINVAL, "null argument");
  if (iters < 0) return fail(RK_EINVAL, "iters must be >= 0");
  for (rk_batch* b : g->batches) {
    if (b->n <= 0) return fail(RK_EINVAL, "a batch of the group has no problems");
    if (b->trace && iters > b->trace_iters) return fail(RK_EINVAL, "iters exceeds a batch's trace length");
  }
  rk_context* ctx = g->ctx;
  std::lock_guard<std::mutex> lk(ctx->mu);
  HIP_TRY(hipSetDevice(ctx->device));
  const LMParams lp = to_lp(p);
  std::vector<float> key = {(float)iters, lp.lm_lambda, lp.lambda_factor, lp.lambda_min, lp.lambda_max, lp.rho_min,
                            lp.gain_ratio_epsilon};
  for (rk_batch* b : g->batches) key.push_back((float)b->n);
  key.push_back(from_seed ? 1.f : 0.f);
  int rc;
  if (use_graph) {
    if (!g->graph || g->graph_key != key) {
      if (g->graph) {
        hipGraphExecDestroy(g->graph);
        g->graph = nullptr;
      }
      GroupLaunch gl;
      if ((rc = group_prepare(g, &gl, from_seed, g->buckets_graph, g->traj_graph))) return rc;
      hipGraph_t graph = nullptr;
      HIP_TRY(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
      rc = group_launch(g, gl, iters, lp, false);
      const std::string launch_err = g_err;
      hipError_t e = hipStreamEndCapture(ctx->stream, &graph);
      if (rc || e != hipSuccess) {
        if (e == hipSuccess && graph) hipGraphDestroy(graph);
        if (rc) return fail(rc, launch_err);
        return fail(RK_EHIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
      }
      e = hipGraphInstantiate(&g->graph, graph, nullptr, nullptr, 0);
      hipGraphDestroy(graph);
      if (e != hipSuccess) return fail(RK_EHIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
      g->graph_key = key;
    }
    // the start mark of a synchronous solve's kernel time; an enqueued solve (RK_ASYNC) records
    // none: between back-to-back graph launches the event cost 5-6 us per uneven solve (1.3 %)
    if (!(flags & RK_ASYNC)) HIP_TRY(hipEventRecord(ctx->ev0, ctx->stream));
    HIP_TRY(hipGraphLaunch(g->graph, ctx->stream));
    g->last_var.clear();  // per-kernel times are recorded by eager solves only
    g->last_mask.clear();
  } else {
    GroupLaunch gl;
    if ((rc = group_prepare(g, &gl, from_seed, g->buckets_dev, g->traj_dev))) return rc;
    HIP_TRY(hipEventRecord(ctx->ev0, ctx->stream));
    if ((rc = group_launch(g, gl, iters, lp, true))) return rc;
    g->last_var = gl.var;
    g->last_mask = gl.mask;
    for (size_t c = 0; c < gl.t_na.size(); ++c) {
      g->last_var.push_back(RK_VAR_TRAJ + (int)c);
      g->last_mask.push_back(gl.t_mask[c]);
    }
  }
  if (flags & RK_ASYNC) return RK_OK;  // enqueued; the caller synchronises
  HIP_TRY(hipEventRecord(ctx->ev1, ctx->stream));
  HIP_TRY(hipEventSynchronize(ctx->ev1));
  if (kernel_ms) HIP_TRY(hipEventElapsedTime(kernel_ms, ctx->ev0, ctx->ev1));
  return RK_OK;
}

int rk_group_kernel_times(rk_group* g, int32_t cap, int32_t* tpp, float* ms, int32_t* batch_mask, int32_t* n) {
  if (!g || !n) return fail(RK_EINVAL, "null argument");
  const int m = (int)g->last_var.size();
  *n = m;
  for (int k = 0; k < m && k < cap; ++k) {
    const int v = g->last_var[k];
    if (tpp)
      tpp[k] = v >= RK_VAR_TRAJ ? RK_TPP_TRAJ
               : variant_wave(v) ? RK_TPP_WAVE
               : variant_super(v) ? RK_TPP_SUPER
               : variant_quad(v) ? RK_TPP_QUAD
               : variant_half(v) ? RK_TPP_HALF
                                 : variant_G(v);
    if (ms) HIP_TRY(hipEventElapsedTime(&ms[k], g->ev_k0[k], g->ev_k1[k]));
    if (batch_mask) batch_mask[k] = (int32_t)g->last_mask[k];
  }
  return RK_OK;
}

// one batch of a group: time every legal threads-per-problem code on `iters` iterations from the
// batch's current q (restored after each trial and at the end, traces detached meanwhile); on any
// exit the batch's q, trace pointers and the scratch copy are restored / freed
static int autotune_batch(rk_group* g, rk_batch* b, int iters, const LMParams& lp, int* best_out,
                          std::vector<std::pair<float, int>>* ranked) {
  rk_context* ctx = g->ctx;
  const size_t S = b->capacity, qbytes = sizeof(float) * b->na * S;
  float* q_save = nullptr;
  int rc;
  if ((rc = dmalloc(ctx, (void**)&q_save, qbytes, nullptr))) return rc;
  float* const tr = b->trace;
  int* const at = b->acc_trace;
  float* const qt = b->q_trace;
  float* const lt = b->lam_trace;
  auto restore = [&](int code) {
    if (hipMemcpyAsync(b->q, q_save, qbytes, hipMemcpyDeviceToDevice, ctx->stream) != hipSuccess ||
        hipStreamSynchronize(ctx->stream) != hipSuccess) {
      if (code == RK_OK) code = fail(RK_EHIP, "rk_group_autotune: could not restore the batch's q");
    }
    dfree(ctx, q_save, qbytes);
    b->trace = tr;
    b->acc_trace = at;
    b->q_trace = qt;
    b->lam_trace = lt;
    return code;
  };
  if (hipMemcpyAsync(q_save, b->q, qbytes, hipMemcpyDeviceToDevice, ctx->stream) != hipSuccess) {
    dfree(ctx, q_save, qbytes);
    return fail(RK_EHIP, "rk_group_autotune: could not save the batch's q");
  }
  b->trace = nullptr;
  b->acc_trace = nullptr;
  b->q_trace = nullptr;
  b->lam_trace = nullptr;
  std::vector<int> cands;
  legal_tpps(b, cands);
  float best = 1e30f;
  int best_g = cands[0];
  for (int G : cands) {
    int v;
    if (pick_variant(b, G, &v)) continue;
    float ms_sum = 0.f;
    for (int rep = 0; rep < 2; ++rep) {
      hipError_t e = hipMemcpyAsync(b->q, q_save, qbytes, hipMemcpyDeviceToDevice, ctx->stream);
      if (e == hipSuccess) e = hipEventRecord(ctx->ev0, ctx->stream);
      if (e != hipSuccess) return restore(fail(RK_EHIP, std::string("rk_group_autotune: ") + hipGetErrorString(e)));
      if ((rc = launch_variant(ctx->stream, v, make_problems(b), lp, 0, iters))) return restore(rc);
      float ms = 0.f;
      e = hipEventRecord(ctx->ev1, ctx->stream);
      if (e == hipSuccess) e = hipEventSynchronize(ctx->ev1);
      if (e == hipSuccess) e = hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1);
      if (e != hipSuccess) return restore(fail(RK_EHIP, std::string("rk_group_autotune: ") + hipGetErrorString(e)));
      if (rep == 1) ms_sum += ms;
    }
    ranked->push_back({ms_sum, G});
    if (ms_sum < best) {
      best = ms_sum;
      best_g = G;
    }
  }
  *best_out = best_g;
  std::stable_sort(ranked->begin(), ranked->end());
  return restore(RK_OK);
}

// The whole group with the kernel codes tpp, launched as a solve launches it (one k_persistent_reg
// launch when every variant is register-resident): the median HIP-event time of three runs after
// a warm-up, each from the batches' current q (restored before each run), traces detached.  (A
// single short run measured too noisily to separate combinations 15 % apart, and the shortest of
// several favours a combination whose time varies with the workgroups' placement.)  A variant that is fastest
// alone need not be fastest beside the others: the launch shares the chip, and a variant with many
// short wavefronts can leave the resident slots to the long ones of another.
static int autotune_group_time(rk_group* g, const std::vector<int>& tpp, int iters, const LMParams& lp,
                               const std::vector<float*>& q_save, float* ms_out) {
  rk_context* ctx = g->ctx;
  for (size_t i = 0; i < g->batches.size(); ++i) g->tpp[i] = tpp[i];
  GroupLaunch gl;
  int rc;
  if ((rc = group_prepare(g, &gl, false, g->buckets_dev, g->traj_dev))) return rc;
  if (iters <= 0) {  // only the IK workgroups of the launch (for the resident-slot test)
    int blocks = 0;
    for (size_t k = 0; k < gl.var.size(); ++k) blocks += gl.grid[k];
    *ms_out = (float)blocks;
    return RK_OK;
  }
  float t[3] = {0.f, 0.f, 0.f};
  for (int rep = 0; rep < 4; ++rep) {
    for (size_t i = 0; i < g->batches.size(); ++i) {
      rk_batch* b = g->batches[i];
      HIP_TRY(hipMemcpyAsync(b->q, q_save[i], sizeof(float) * b->na * b->capacity, hipMemcpyDeviceToDevice,
                             ctx->stream));
    }
    HIP_TRY(hipEventRecord(ctx->ev0, ctx->stream));
    if ((rc = group_launch(g, gl, iters, lp, false))) return rc;
    HIP_TRY(hipEventRecord(ctx->ev1, ctx->stream));
    HIP_TRY(hipEventSynchronize(ctx->ev1));
    float ms = 0.f;
    HIP_TRY(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
    if (rep > 0) t[rep - 1] = ms;
  }
  std::sort(t, t + 3);
  *ms_out = t[1];
  return RK_OK;
}

// Group pass of the autotuner: the combinations of each IK batch's front-runners alone (codes
// within 2x of its best, at most three per batch, the first four batches) timed as a whole group;
// the fastest combination is kept.  State (q, trace pointers) is restored on every exit.
static int autotune_group(rk_group* g, int iters, const LMParams& lp,
                          const std::vector<std::vector<std::pair<float, int>>>& ranked) {
  rk_context* ctx = g->ctx;
  const size_t nb = g->batches.size();
  std::vector<std::vector<int>> opts(nb);
  size_t combos = 1;
  for (size_t i = 0; i < nb; ++i) {
    opts[i].push_back(g->tpp[i]);
    if (g->batches[i]->traj_T || i >= 4) continue;
    for (const auto& r : ranked[i])
      if (r.second != g->tpp[i] && r.first <= 2.f * ranked[i][0].first && opts[i].size() < 3)
        opts[i].push_back(r.second);
    combos *= opts[i].size();
  }
  if (combos <= 1) return RK_OK;
  std::vector<float*> q_save(nb, nullptr);
  std::vector<float*> tr(nb), qt(nb), lt(nb);
  std::vector<int*> at(nb);
  const std::vector<int> tpp0 = g->tpp;
  auto cleanup = [&](int code) {
    if (hipStreamSynchronize(ctx->stream) != hipSuccess && code == RK_OK)
      code = fail(RK_EHIP, "rk_group_autotune: stream error");
    for (size_t i = 0; i < nb; ++i) {
      rk_batch* b = g->batches[i];
      if (q_save[i]) {
        if (hipMemcpy(b->q, q_save[i], sizeof(float) * b->na * b->capacity, hipMemcpyDeviceToDevice) != hipSuccess &&
            code == RK_OK)
          code = fail(RK_EHIP, "rk_group_autotune: could not restore a batch's q");
        dfree(ctx, q_save[i], sizeof(float) * b->na * b->capacity);
      }
      b->trace = tr[i];
      b->acc_trace = at[i];
      b->q_trace = qt[i];
      b->lam_trace = lt[i];
    }
    if (code != RK_OK) g->tpp = tpp0;
    return code;
  };
  int rc;
  for (size_t i = 0; i < nb; ++i) {
    rk_batch* b = g->batches[i];
    tr[i] = b->trace;
    at[i] = b->acc_trace;
    qt[i] = b->q_trace;
    lt[i] = b->lam_trace;
    b->trace = nullptr;
    b->acc_trace = nullptr;
    b->q_trace = nullptr;
    b->lam_trace = nullptr;
    if ((rc = dmalloc(ctx, (void**)&q_save[i], sizeof(float) * b->na * b->capacity, nullptr))) return cleanup(rc);
    if (hipMemcpyAsync(q_save[i], b->q, sizeof(float) * b->na * b->capacity, hipMemcpyDeviceToDevice, ctx->stream) !=
        hipSuccess)
      return cleanup(fail(RK_EHIP, "rk_group_autotune: could not save a batch's q"));
  }
  // a single launch whose workgroups exceed the resident slots (2 wavefronts per SIMD) runs its
  // last workgroups only as others finish: its time then hangs on the workgroups' placement.  When
  // some combination fits, the ones that do not are not considered.
  auto combo = [&](size_t c, std::vector<int>& cur) {
    size_t k = c;
    for (size_t i = 0; i < nb; ++i) {
      cur[i] = opts[i][k % opts[i].size()];
      k /= opts[i].size();
    }
  };
  const float slots = 2.f * 4.f * (float)ctx->cu_count;
  std::vector<char> fits(combos, 1);
  bool any_fits = false;
  std::vector<int> cur(nb);
  for (size_t c = 0; c < combos; ++c) {
    combo(c, cur);
    float blocks = 0.f;
    if ((rc = autotune_group_time(g, cur, 0, lp, q_save, &blocks))) return cleanup(rc);
    fits[c] = blocks <= slots;
    any_fits = any_fits || fits[c];
  }
  float best = 1e30f;
  std::vector<int> best_tpp = tpp0;
  for (size_t c = 0; c < combos; ++c) {
    if (any_fits && !fits[c]) continue;
    combo(c, cur);
    float ms = 0.f;
    if ((rc = autotune_group_time(g, cur, std::max(iters, 20), lp, q_save, &ms))) return cleanup(rc);
    if (ms < best) {
      best = ms;
      best_tpp = cur;
    }
  }
  g->tpp = best_tpp;
  return cleanup(RK_OK);
}

int rk_group_autotune(rk_group* g, int32_t iters, const rk_lm_params* p, int32_t* chosen) {
  if (!g || !p || iters <= 0) return fail(RK_EINVAL, "bad argument to rk_group_autotune");
  rk_context* ctx = g->ctx;
  const LMParams lp = to_lp(p);
  // each batch alone, each legal G; state is restored after every trial
  std::vector<std::vector<std::pair<float, int>>> ranked(g->batches.size());
  for (size_t i = 0; i < g->batches.size(); ++i) {
    rk_batch* b = g->batches[i];
    if (b->traj_T) continue;  // trajectories have one kernel form
    std::lock_guard<std::mutex> lk(ctx->mu);
    HIP_TRY(hipSetDevice(ctx->device));
    int best_g = 0;
    const int rc = autotune_batch(g, b, iters, lp, &best_g, &ranked[i]);
    if (rc) return rc;
    g->tpp[i] = best_g;
    ctx->tuned[signature(b->th)] = best_g;
  }
  // then the front-runners' combinations as a whole group
  {
    std::lock_guard<std::mutex> lk(ctx->mu);
    HIP_TRY(hipSetDevice(ctx->device));
    const int rc = autotune_group(g, iters, lp, ranked);
    if (rc) return rc;
  }
  for (size_t i = 0; i < g->batches.size(); ++i)
    if (chosen) chosen[i] = g->batches[i]->traj_T ? RK_TPP_TRAJ : g->tpp[i];
  if (g->graph) {
    hipGraphExecDestroy(g->graph);
    g->graph = nullptr;
  }
  return RK_OK;
}

}  // extern "C"

#include "rk_serve.h"

// rk_serve.h -- the request server: continuous batching of LM solves under a device-memory
// budget (included at the end of rk_lm.hip, after the batch and group entry points).
//
// A worker thread owns a device arena of `budget` bytes and S launch streams.  Callers submit
// requests (a number of problems of one registered term structure, their seeds and targets);
// submission copies the inputs and returns a ticket at once.  Whenever a stream is free the
// worker packs the oldest queued request together with every later queued request of the same
// term structure and iteration count that fits (problem cap, arena space) into ONE launch:
// inputs staged through pinned memory, one fused LM kernel over all their problems, the
// converged/failed metrics, results copied back; the launch's arena block is released when its
// event completes.  Requests arriving while launches are on the device therefore join the next
// launch as soon as a stream and memory are free, and launches on different streams overlap on
// the device.  With trajectory term structures registered, stream 0 carries the trajectory
// launches, enqueued a chunk of iterations at a time, and the other streams the IK launches.  A
// problem's result does not depend on what it was packed with: every problem is solved
// independently by the same kernel variant (chosen per term structure).
//
// Warm start: a request may name a session.  The session keeps the last solution of its problems
// resident in the arena; a later request of the session with no seeds starts from it and uses it
// as the smoothness reference (retargeting a stream of frames).  Requests of one session run in
// submission order.

namespace {

double now_s() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

// first-fit allocator over one device block (worker thread only)
struct Arena {
  size_t size = 0;
  std::map<size_t, size_t> free_;  // offset -> length
  size_t used = 0, peak = 0;
  void init(size_t n) {
    size = n;
    free_.clear();
    free_[0] = n;
    used = peak = 0;
  }
  bool alloc(size_t n, size_t* off) {
    n = (n + 255) & ~size_t(255);
    for (auto it = free_.begin(); it != free_.end(); ++it) {
      if (it->second >= n) {
        *off = it->first;
        const size_t rest = it->second - n, o = it->first;
        free_.erase(it);
        if (rest) free_[o + n] = rest;
        used += n;
        peak = std::max(peak, used);
        return true;
      }
    }
    return false;
  }
  void release(size_t off, size_t n) {
    n = (n + 255) & ~size_t(255);
    used -= n;
    auto it = free_.emplace(off, n).first;
    auto nx = std::next(it);
    if (nx != free_.end() && it->first + it->second == nx->first) {
      it->second += nx->second;
      free_.erase(nx);
    }
    if (it != free_.begin()) {
      auto pv = std::prev(it);
      if (pv->first + pv->second == it->first) {
        pv->second += it->second;
        free_.erase(it);
      }
    }
  }
};

}  // namespace

struct SrvTerms {
  Terms th{};
  Terms* td = nullptr;
  int variant = 0, na = 0;  // na: variables per problem (frames * dofs for trajectories)
  int traj_na = 0;          // > 0: trajectory problems of an arm with this many dofs
  bool traj_lim = false;    // they carry bound rows or scene geometries
  TrajTerms* tt = nullptr;
};

// success of a trajectory problem: goal within the thresholds and every sphere clear of the scene
__global__ void k_srv_traj_ok(const float* pe, const float* ae, const float* sd, int* ok, int n, float pos_thr,
                              float ang_thr) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) ok[i] = (pe[i] < pos_thr && ae[i] < ang_thr && sd[i] >= 0.f) ? 1 : 0;
}

struct SrvReq {
  int64_t ticket = 0, session = 0;
  int term = 0, n = 0, iters = 0;
  LMParams lp{};
  bool has_q = false, has_prev = false;
  std::vector<float> q, pose, pos, qprev;  // inputs, row-major per problem
  std::vector<float> q_out, cost, pe, ae;
  std::vector<int> ok;
  double t_submit = 0, t_start = 0, t_done = 0;
  int state = 0;  // 0 queued, 1 on the device, 2 done, 3 failed
  std::string err;
  int first = 0;  // first problem of this request in its launch
};

struct SrvSession {
  size_t off = 0, bytes = 0;  // arena block: last solution, row-major [n][na]
  int n = 0, na = 0;
  bool valid = false;
  int64_t busy = 0;  // ticket of the request of this session on the device (0: none)
};

struct SrvLaunch {
  bool active = false;
  std::vector<SrvReq*> reqs;
  int term = 0, n = 0;
  size_t off = 0, bytes = 0;
  // device slices (SoA, stride n)
  float *q = nullptr, *qs = nullptr, *pose = nullptr, *pos = nullptr, *qp = nullptr, *cost = nullptr,
        *lam = nullptr, *pe = nullptr, *ae = nullptr;
  int *nacc = nullptr, *ok = nullptr;
  // a trajectory launch runs its iterations in chunks, each enqueued when the previous one's event
  // completes: it_done iterations are on the device, iters in all; tail: the metrics and the
  // download are enqueued (the next completion finishes the launch)
  int it_done = 0, iters = 0;
  bool tail = false;
};

// iterations per enqueued trajectory chunk: an IK launch whose stream shares a hardware queue with
// the trajectory stream waits behind at most one chunk (~0.7 ms for 64-frame trajectories), not a
// whole 50-iteration solve
constexpr int SRV_TRAJ_CHUNK = 5;

struct rk_server {
  rk_context* ctx = nullptr;
  char* arena_base = nullptr;
  Arena arena;
  int nstreams = 0, max_problems = 0;
  float pos_thr = 1e-3f, ang_thr = 1e-2f;
  std::vector<hipStream_t> streams;
  std::vector<hipEvent_t> events;
  std::vector<float*> h_in, h_out;  // pinned staging per stream
  size_t stage_floats = 0;
  std::vector<SrvLaunch> launches;  // one slot per stream
  std::vector<SrvTerms> terms;
  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  std::deque<SrvReq*> queue;
  std::map<int64_t, SrvReq*> reqs;
  std::map<int64_t, SrvSession> sessions;
  int64_t next_ticket = 1;
  bool stop = false;
  std::thread worker;
  std::string worker_err;
  // statistics
  int64_t submitted = 0, completed = 0, failed = 0, nlaunch = 0, problems = 0, converged = 0;
  std::vector<double> lat_ms;
  double first_submit = 0, last_done = 0;
};

// row-major [n][na] (session state) <-> SoA slice [k * stride + first + i]
__global__ void k_srv_gather(float* __restrict__ dst, int stride, int first, const float* __restrict__ src, int n,
                             int na) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n * na) return;
  const int p = i / na, k = i % na;
  dst[(size_t)k * stride + first + p] = src[i];
}
__global__ void k_srv_scatter(float* __restrict__ dst, const float* __restrict__ src, int stride, int first, int n,
                              int na) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n * na) return;
  const int p = i / na, k = i % na;
  dst[i] = src[(size_t)k * stride + first + p];
}

// device layout of a launch over n problems (SoA, stride n): the input region (one upload), then
// state and outputs (one download); every slice 256-byte aligned.  Returns the total bytes and,
// with L, sets its slice pointers inside `base`.
static size_t srv_layout(const SrvTerms& t, int n, char* base, SrvLaunch* L, size_t* in_bytes, size_t* out_bytes) {
  const int na = t.na, np = t.th.n_pose, nq = t.th.n_pos;
  size_t o = 0;
  auto take = [&](size_t floats) {
    float* p = reinterpret_cast<float*>(base + o);
    o += ((floats * sizeof(float) + 255) / 256) * 256;
    return p;
  };
  float* qs = take((size_t)na * n);
  float* pose = take((size_t)std::max(1, 7 * np) * n);
  float* pos = take((size_t)std::max(1, 3 * nq) * n);
  float* qp = take((size_t)na * n);
  const size_t ib = o;
  float* q = take((size_t)na * n);
  float* cost = take(n);
  float* pe = take((size_t)std::max(1, np) * n);
  float* ae = take((size_t)std::max(1, np) * n);
  float* ok = take(n);
  const size_t ob = o - ib;
  float* lam = take(n);
  float* nacc = take(n);
  if (L) {
    L->qs = qs, L->pose = pose, L->pos = pos, L->qp = qp, L->q = q, L->cost = cost, L->pe = pe, L->ae = ae;
    L->ok = reinterpret_cast<int*>(ok);
    L->lam = lam;
    L->nacc = reinterpret_cast<int*>(nacc);
  }
  if (in_bytes) *in_bytes = ib;
  if (out_bytes) *out_bytes = ob;
  return o;
}
static size_t srv_launch_bytes(const SrvTerms& t, int n) { return srv_layout(t, n, nullptr, nullptr, nullptr, nullptr); }
// do the staging buffers hold a launch of n problems of t?
static bool srv_fits_stage(const rk_server* sv, const SrvTerms& t, int n) {
  size_t ib = 0, ob = 0;
  srv_layout(t, n, nullptr, nullptr, &ib, &ob);
  return ib <= sizeof(float) * sv->stage_floats && ob <= sizeof(float) * sv->stage_floats;
}

// the problem view of launch slot s (SoA slices of its arena block, stride n)
static Problems srv_problems(rk_server* sv, const SrvLaunch& L) {
  const SrvTerms& T = sv->terms[L.term];
  Problems P;
  memset(&P, 0, sizeof(P));
  P.robot = sv->ctx->robots_dev + T.th.robot;
  P.terms = T.td;
  P.q = L.q;
  P.q_seed = L.qs;
  P.pose_tgt = L.pose;
  P.pos_tgt = L.pos;
  P.q_prev = L.qp;
  P.cost = L.cost;
  P.lam = L.lam;
  P.nacc = L.nacc;
  P.n = L.n;
  P.stride = L.n;
  return P;
}

// the launch's results: converged / failed flags, the sessions' resident state, one download and
// the completion event (worker thread)
static int srv_enqueue_tail(rk_server* sv, int s) {
  SrvLaunch& L = sv->launches[s];
  const SrvTerms& T = sv->terms[L.term];
  const int n = L.n, na = T.na, np = T.th.n_pose;
  hipStream_t st = sv->streams[s];
  const Problems P = srv_problems(sv, L);
  if (T.traj_na) {
    // lambda is not read back: its slice holds the clearance for the success flags
    launch_traj_metrics(st, T.traj_na, T.tt, P, L.pe, L.ae, L.lam);
    hipLaunchKernelGGL(k_srv_traj_ok, dim3((n + 127) / 128), dim3(128), 0, st, L.pe, L.ae, L.lam, L.ok, n,
                       sv->pos_thr, sv->ang_thr);
  } else if (np) {
    hipLaunchKernelGGL(k_ik_metrics, dim3((n + 127) / 128), dim3(128), 0, st, P, sv->pos_thr, sv->ang_thr, L.pe,
                       L.ae, L.ok);
  }
  // session state <- this solution
  for (SrvReq* r : L.reqs) {
    if (!r->session) continue;
    const SrvSession& ss = sv->sessions[r->session];
    float* dst = reinterpret_cast<float*>(sv->arena_base + ss.off);
    const int tot = r->n * na, blk = 256;
    hipLaunchKernelGGL(k_srv_scatter, dim3((tot + blk - 1) / blk), dim3(blk), 0, st, dst, L.q, n, r->first, r->n, na);
  }
  size_t in_bytes = 0, out_bytes = 0;
  srv_layout(T, n, nullptr, nullptr, &in_bytes, &out_bytes);
  HIP_TRY(hipMemcpyAsync(sv->h_out[s], L.q, out_bytes, hipMemcpyDeviceToHost, st));
  HIP_TRY(hipEventRecord(sv->events[s], st));
  HIP_TRY(hipGetLastError());
  L.tail = true;
  return RK_OK;
}

// the next chunk of a trajectory launch and its event, or, after the last one, the tail
static int srv_traj_next(rk_server* sv, int s) {
  SrvLaunch& L = sv->launches[s];
  const SrvTerms& T = sv->terms[L.term];
  if (L.it_done >= L.iters) return srv_enqueue_tail(sv, s);
  const int n_it = std::min(SRV_TRAJ_CHUNK, L.iters - L.it_done);
  int rc = launch_traj(sv->streams[s], T.traj_na, T.tt, srv_problems(sv, L), L.reqs[0]->lp, L.it_done, n_it,
                       T.traj_lim, true);
  if (rc) return rc;
  L.it_done += n_it;
  HIP_TRY(hipEventRecord(sv->events[s], sv->streams[s]));
  return RK_OK;
}

// stage, launch and enqueue the copies of launch slot s (worker thread): an IK launch enqueues
// every step at once; a trajectory launch its first chunk (srv_traj_next enqueues the rest)
static int srv_start(rk_server* sv, int s) {
  SrvLaunch& L = sv->launches[s];
  const SrvTerms& T = sv->terms[L.term];
  const int n = L.n, na = T.na, np = T.th.n_pose, nq = T.th.n_pos;
  hipStream_t st = sv->streams[s];
  char* base = sv->arena_base + L.off;
  size_t in_bytes = 0, out_bytes = 0;
  srv_layout(T, n, base, &L, &in_bytes, &out_bytes);
  if (in_bytes > sizeof(float) * sv->stage_floats || out_bytes > sizeof(float) * sv->stage_floats)
    return fail(RK_EINVAL, "launch exceeds the staging buffers");
  // host staging: SoA of the launch with the same region layout as the device input block
  float* hi = sv->h_in[s];
  auto hslice = [&](float* dev) { return hi + (reinterpret_cast<char*>(dev) - base) / sizeof(float); };
  float *hqs = hslice(L.qs), *hpose = hslice(L.pose), *hpos = hslice(L.pos), *hqp = hslice(L.qp);
  for (SrvReq* r : L.reqs) {
    for (int i = 0; i < r->n; ++i) {
      const int p = r->first + i;
      if (r->has_q)
        for (int k = 0; k < na; ++k) hqs[(size_t)k * n + p] = r->q[(size_t)i * na + k];
      for (int c = 0; c < 7 * np; ++c) hpose[(size_t)c * n + p] = r->pose[(size_t)i * 7 * np + c];
      for (int c = 0; c < 3 * nq; ++c) hpos[(size_t)c * n + p] = r->pos[(size_t)i * 3 * nq + c];
      const std::vector<float>& prev = r->has_prev ? r->qprev : r->q;
      if (r->has_prev || r->has_q)
        for (int k = 0; k < na; ++k) hqp[(size_t)k * n + p] = prev[(size_t)i * na + k];
    }
  }
  HIP_TRY(hipMemcpyAsync(base, hi, in_bytes, hipMemcpyHostToDevice, st));
  // session warm start: seeds (and the smoothness reference) from the resident last solution
  for (SrvReq* r : L.reqs) {
    if (!r->session || r->has_q) continue;
    const SrvSession& ss = sv->sessions[r->session];
    if (!ss.valid) return fail(RK_EINVAL, "warm start: the session holds no solution");
    const float* src = reinterpret_cast<const float*>(sv->arena_base + ss.off);
    const int tot = r->n * na, blk = 256;
    hipLaunchKernelGGL(k_srv_gather, dim3((tot + blk - 1) / blk), dim3(blk), 0, st, L.qs, n, r->first, src, r->n, na);
    if (!r->has_prev)
      hipLaunchKernelGGL(k_srv_gather, dim3((tot + blk - 1) / blk), dim3(blk), 0, st, L.qp, n, r->first, src, r->n,
                         na);
  }
  L.iters = L.reqs[0]->iters;
  L.it_done = 0;
  L.tail = false;
  if (T.traj_na) return srv_traj_next(sv, s);
  int rc = launch_variant(st, T.variant, srv_problems(sv, L), L.reqs[0]->lp, 0, L.iters);
  if (rc) return rc;
  L.it_done = L.iters;
  return srv_enqueue_tail(sv, s);
}

// results of a completed launch back into its requests (worker thread, server lock held)
static void srv_finish(rk_server* sv, int s) {
  SrvLaunch& L = sv->launches[s];
  const SrvTerms& T = sv->terms[L.term];
  const int n = L.n, na = T.na, np = T.th.n_pose;
  const float* ho = sv->h_out[s];
  char* base = reinterpret_cast<char*>(L.q);
  auto hslice = [&](const void* dev) { return ho + (reinterpret_cast<const char*>(dev) - base) / sizeof(float); };
  const float *hq = hslice(L.q), *hc = hslice(L.cost), *hpe = hslice(L.pe), *hae = hslice(L.ae);
  const int* hok = reinterpret_cast<const int*>(hslice(L.ok));
  const double t = now_s();
  for (SrvReq* r : L.reqs) {
    r->q_out.resize((size_t)r->n * na);
    r->cost.resize(r->n);
    r->ok.assign(r->n, 0);
    r->pe.assign((size_t)r->n * std::max(np, 0), 0.f);
    r->ae.assign((size_t)r->n * std::max(np, 0), 0.f);
    for (int i = 0; i < r->n; ++i) {
      const int p = r->first + i;
      for (int k = 0; k < na; ++k) r->q_out[(size_t)i * na + k] = hq[(size_t)k * n + p];
      r->cost[i] = hc[p];
      if (np) {
        r->ok[i] = hok[p];
        for (int k = 0; k < np; ++k) {
          r->pe[(size_t)i * np + k] = hpe[(size_t)p * np + k];
          r->ae[(size_t)i * np + k] = hae[(size_t)p * np + k];
        }
        sv->converged += hok[p];
      }
    }
    r->t_done = t;
    r->state = 2;
    sv->completed += 1;
    sv->problems += r->n;
    sv->lat_ms.push_back((r->t_done - r->t_submit) * 1e3);
    sv->last_done = t;
    if (r->session) {
      SrvSession& ss = sv->sessions[r->session];
      ss.valid = true;
      if (ss.busy == r->ticket) ss.busy = 0;
    }
  }
  sv->arena.release(L.off, L.bytes);
  L.active = false;
  L.reqs.clear();
}

static bool srv_same_key(const SrvReq* a, const SrvReq* b) {
  return a->term == b->term && a->iters == b->iters && memcmp(&a->lp, &b->lp, sizeof(LMParams)) == 0;
}

// can r go on the device now?  (its session has no request in flight, and its session state
// exists or it brings seeds; allocates the session block on first use)
static bool srv_ready(rk_server* sv, SrvReq* r) {
  if (!r->session) return true;
  SrvSession& ss = sv->sessions[r->session];
  if (ss.busy) return false;
  const int na = sv->terms[r->term].na;
  if (ss.bytes == 0) {
    size_t off;
    const size_t bytes = sizeof(float) * (size_t)r->n * na;
    if (!sv->arena.alloc(bytes, &off)) return false;
    ss.off = off;
    ss.bytes = bytes;
    ss.n = r->n;
    ss.na = na;
  }
  return true;
}

static bool srv_busy(const rk_server* sv) {
  for (const auto& L : sv->launches)
    if (L.active) return true;
  return false;
}

static void srv_fail(rk_server* sv, SrvReq* r, const std::string& msg) {
  r->state = 3;
  r->err = msg;
  r->t_done = now_s();
  sv->failed += 1;
  if (r->session) {
    SrvSession& ss = sv->sessions[r->session];
    if (ss.busy == r->ticket) ss.busy = 0;
  }
}

// does the server hold trajectory term structures?  Then launch stream 0 carries trajectory
// launches only and the other streams IK launches only: a trajectory solve runs for milliseconds
// (one wavefront pair per trajectory through ~T * dofs serial band columns), an IK launch for well
// under one, and IK requests must not queue behind it.  With one stream everything shares it.
static bool srv_has_traj(const rk_server* sv) {
  for (const auto& T : sv->terms)
    if (T.traj_na) return true;
  return false;
}

// form one launch on free stream slot s from the queue; false if nothing could start.  The head is
// the oldest ready request of the kinds stream s carries.
static bool srv_form(rk_server* sv, int s) {
  const bool split = sv->nstreams > 1 && srv_has_traj(sv);
  auto carries = [&](const SrvReq* r) { return !split || ((sv->terms[r->term].traj_na != 0) == (s == 0)); };
  // Requests of one session run in submission order: once the scan passes over a request of a
  // session (not ready, other key, no room), every later request of that session waits too.
  std::set<int64_t> passed;
  SrvReq* head = nullptr;
  for (SrvReq* r : sv->queue) {
    if (r->session && passed.count(r->session)) continue;
    if (carries(r) && srv_ready(sv, r)) {
      head = r;
      break;
    }
    if (r->session) passed.insert(r->session);
  }
  if (!head) return false;
  // a warm start needs the session's resident solution: when the session's earlier request
  // failed there is none, and the request fails instead of reading unwritten arena memory
  if (head->session && !head->has_q && !sv->sessions[head->session].valid) {
    srv_fail(sv, head, "warm start: the session holds no solution (its earlier request failed)");
    sv->queue.erase(std::find(sv->queue.begin(), sv->queue.end(), head));
    return true;
  }
  const SrvTerms& T = sv->terms[head->term];
  std::vector<SrvReq*> pick;
  std::set<int64_t> sess;  // sessions with a request picked or passed over by this scan
  int n = 0;
  for (SrvReq* r : sv->queue) {
    if (r->session && sess.count(r->session)) continue;
    const bool take = srv_same_key(r, head) && n + r->n <= sv->max_problems && srv_fits_stage(sv, T, n + r->n) &&
                      (!r->session || (srv_ready(sv, r) && (r->has_q || sv->sessions[r->session].valid)));
    if (r->session) sess.insert(r->session);
    if (!take) continue;
    pick.push_back(r);
    n += r->n;
  }
  size_t off = 0, bytes = 0;
  bool ok = false;
  while (true) {
    bytes = srv_launch_bytes(T, n);
    if (sv->arena.alloc(bytes, &off)) {
      ok = true;
      break;
    }
    if (pick.size() <= 1) break;
    n -= pick.back()->n;
    pick.pop_back();
  }
  if (!ok) {
    if (!srv_busy(sv)) {  // nothing will free memory: the head request can never fit
      srv_fail(sv, head, "request needs " + std::to_string(bytes) + " bytes of device memory; the server budget has " +
                             std::to_string(sv->arena.size - sv->arena.used) + " free");
      sv->queue.erase(std::find(sv->queue.begin(), sv->queue.end(), head));
      return true;
    }
    return false;  // wait for a launch to release memory
  }
  SrvLaunch& L = sv->launches[s];
  L.active = true;
  L.reqs = pick;
  L.term = head->term;
  L.n = n;
  L.off = off;
  L.bytes = bytes;
  const double t = now_s();
  int first = 0;
  for (SrvReq* r : pick) {
    r->first = first;
    first += r->n;
    r->state = 1;
    r->t_start = t;
    if (r->session) sv->sessions[r->session].busy = r->ticket;
    sv->queue.erase(std::find(sv->queue.begin(), sv->queue.end(), r));
  }
  const int rc = srv_start(sv, s);
  sv->nlaunch += 1;
  if (rc) {
    // nothing of this launch may still be pending on the stream before its memory is reused
    hipStreamSynchronize(sv->streams[s]);
    for (SrvReq* r : pick) srv_fail(sv, r, g_err);
    sv->arena.release(L.off, L.bytes);
    L.active = false;
    L.reqs.clear();
  }
  return true;
}

static void srv_loop(rk_server* sv) {
  const bool dev_ok = hipSetDevice(sv->ctx->device) == hipSuccess;
  std::unique_lock<std::mutex> lk(sv->mu);
  if (!dev_ok) sv->worker_err = "hipSetDevice failed in the server thread";
  while (true) {
    bool progress = false;
    for (int s = 0; s < sv->nstreams; ++s) {
      SrvLaunch& L = sv->launches[s];
      if (!L.active) continue;
      const hipError_t e = hipEventQuery(sv->events[s]);
      if (e == hipSuccess && !L.tail) {  // a trajectory chunk ended: the next one (or the tail)
        if (srv_traj_next(sv, s)) {
          hipStreamSynchronize(sv->streams[s]);
          for (SrvReq* r : L.reqs) srv_fail(sv, r, g_err);
          L.active = false;
          L.reqs.clear();
          sv->arena.release(L.off, L.bytes);
        }
        progress = true;
      } else if (e == hipSuccess) {
        srv_finish(sv, s);
        progress = true;
      } else if (e != hipErrorNotReady) {
        for (SrvReq* r : L.reqs) srv_fail(sv, r, std::string("launch failed: ") + hipGetErrorString(e));
        L.active = false;
        L.reqs.clear();
        sv->arena.release(L.off, L.bytes);
        progress = true;
      }
    }
    if (progress) sv->cv_done.notify_all();
    if (sv->stop && sv->queue.empty() && !srv_busy(sv)) break;
    if (dev_ok) {
      for (int s = 0; s < sv->nstreams && !sv->queue.empty(); ++s) {
        if (sv->launches[s].active) continue;
        if (!srv_form(sv, s)) break;
        progress = true;
      }
    } else {
      while (!sv->queue.empty()) {
        srv_fail(sv, sv->queue.front(), sv->worker_err);
        sv->queue.pop_front();
        progress = true;
      }
    }
    if (progress) {
      sv->cv_done.notify_all();
      continue;
    }
    sv->cv_work.wait_for(lk, std::chrono::microseconds(srv_busy(sv) ? 20 : 5000));
  }
  sv->cv_done.notify_all();
}

extern "C" {

int rk_server_create(rk_context* ctx, int64_t mem_budget, int32_t streams, int32_t max_launch_problems,
                     float pos_thr, float ang_thr, rk_server** out) {
  if (!ctx || !out) return fail(RK_EINVAL, "null argument");
  if (mem_budget < (1 << 20)) return fail(RK_EINVAL, "memory budget below 1 MiB");
  if (streams < 1 || streams > 16) return fail(RK_EINVAL, "streams must be 1..16");
  if (max_launch_problems < 1) return fail(RK_EINVAL, "max_launch_problems must be positive");
  std::lock_guard<std::mutex> lk(ctx->mu);
  HIP_TRY(hipSetDevice(ctx->device));
  rk_server* sv = new rk_server();
  sv->ctx = ctx;
  sv->nstreams = streams;
  sv->max_problems = max_launch_problems;
  sv->pos_thr = pos_thr;
  sv->ang_thr = ang_thr;
  int rc = dmalloc(ctx, (void**)&sv->arena_base, (size_t)mem_budget, nullptr);
  if (rc) {
    delete sv;
    return rc;
  }
  sv->arena.init((size_t)mem_budget);
  // pinned staging per stream, sized for the largest launch of the largest supported problem
  sv->stage_floats = (size_t)max_launch_problems * (2 * RK_NA_MAX + 7 * RK_MAXT + 3 * RK_MAXQ + 64) + 16 * 64;
  sv->streams.resize(streams);
  sv->events.resize(streams);
  sv->h_in.resize(streams, nullptr);
  sv->h_out.resize(streams, nullptr);
  sv->launches.resize(streams);
  for (int s = 0; s < streams; ++s) {
    HIP_TRY(hipStreamCreateWithFlags(&sv->streams[s], hipStreamNonBlocking));
    HIP_TRY(hipEventCreateWithFlags(&sv->events[s], hipEventDisableTiming));
    HIP_TRY(hipHostMalloc((void**)&sv->h_in[s], sizeof(float) * sv->stage_floats, hipHostMallocDefault));
    HIP_TRY(hipHostMalloc((void**)&sv->h_out[s], sizeof(float) * sv->stage_floats, hipHostMallocDefault));
  }
  sv->worker = std::thread(srv_loop, sv);
  *out = sv;
  return RK_OK;
}

int rk_server_add_terms(rk_server* sv, const rk_terms_desc* t, int32_t* term_id) {
  if (!sv || !t || !term_id) return fail(RK_EINVAL, "null argument");
  rk_context* ctx = sv->ctx;
  SrvTerms T;
  int rc;
  {
    std::lock_guard<std::mutex> lk(ctx->mu);
    HIP_TRY(hipSetDevice(ctx->device));
    if ((rc = make_terms(ctx, t, &T.th))) return rc;
    if (T.th.n_rot) return fail(RK_EINVAL, "server requests carry pose and position targets (no rotation targets)");
    // launches carry a few hundred problems at most: the default for a small batch
    if ((rc = pick_variant(ctx, T.th, 0, &T.variant, 1))) return rc;
    T.na = ctx->robots[T.th.robot].na;
    if ((rc = dmalloc(ctx, (void**)&T.td, sizeof(Terms), nullptr))) return rc;
    HIP_TRY(hipMemcpyAsync(T.td, &T.th, sizeof(Terms), hipMemcpyHostToDevice, ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));  // robot uploads and this copy precede server launches
  }
  std::lock_guard<std::mutex> lk(sv->mu);
  sv->terms.push_back(T);
  *term_id = (int32_t)sv->terms.size() - 1;
  return RK_OK;
}

int rk_server_add_traj(rk_server* sv, const rk_traj_desc* d, int32_t* term_id) {
  if (!sv || !d || !term_id) return fail(RK_EINVAL, "null argument");
  rk_context* ctx = sv->ctx;
  SrvTerms T;
  int rc;
  {
    std::lock_guard<std::mutex> lk(ctx->mu);
    HIP_TRY(hipSetDevice(ctx->device));
    TrajTerms tt;
    if ((rc = make_traj_terms(ctx, d, &tt))) return rc;
    T.th = traj_goal_terms(d, tt);
    T.traj_na = ctx->robots[d->robot].na;
    T.traj_lim = tt.lim_orders != 0 || tt.ng > 0;
    T.na = d->frames * T.traj_na;
    if ((rc = dmalloc(ctx, (void**)&T.td, sizeof(Terms), nullptr))) return rc;
    if ((rc = dmalloc(ctx, (void**)&T.tt, sizeof(TrajTerms), nullptr))) return rc;
    HIP_TRY(hipMemcpyAsync(T.td, &T.th, sizeof(Terms), hipMemcpyHostToDevice, ctx->stream));
    HIP_TRY(hipMemcpyAsync(T.tt, &tt, sizeof(TrajTerms), hipMemcpyHostToDevice, ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
  }
  std::lock_guard<std::mutex> lk(sv->mu);
  sv->terms.push_back(T);
  *term_id = (int32_t)sv->terms.size() - 1;
  return RK_OK;
}

int rk_server_submit(rk_server* sv, const rk_request* rq, int64_t* ticket) {
  if (!sv || !rq || !ticket) return fail(RK_EINVAL, "null argument");
  std::lock_guard<std::mutex> lk(sv->mu);
  if (sv->stop) return fail(RK_EINVAL, "server is shutting down");
  if (rq->terms < 0 || rq->terms >= (int)sv->terms.size()) return fail(RK_EINVAL, "unknown term structure id");
  const SrvTerms& T = sv->terms[rq->terms];
  if (rq->n <= 0 || rq->n > sv->max_problems)
    return fail(RK_EINVAL, "request size must be 1.." + std::to_string(sv->max_problems));
  if (!srv_fits_stage(sv, T, rq->n)) return fail(RK_EINVAL, "request too large for the server's staging buffers");
  if (rq->iters < 1) return fail(RK_EINVAL, "iters must be >= 1");
  if (T.th.n_pose && !rq->pose_targets) return fail(RK_EINVAL, "pose targets missing");
  if (T.th.n_pos && !rq->pos_targets) return fail(RK_EINVAL, "position targets missing");
  if (!rq->q) {
    if (!rq->session) return fail(RK_EINVAL, "seeds missing (and no session to warm-start from)");
    auto it = sv->sessions.find(rq->session);
    if (it == sv->sessions.end()) return fail(RK_EINVAL, "seeds missing: the session has no earlier request");
  }
  if (rq->session) {
    auto it = sv->sessions.find(rq->session);
    if (it != sv->sessions.end() && (it->second.n != rq->n || it->second.na != T.na))
      return fail(RK_EINVAL, "a session keeps one problem count and robot");
    if (it == sv->sessions.end()) {
      SrvSession ss;
      ss.n = rq->n;
      ss.na = T.na;
      sv->sessions[rq->session] = ss;
    }
  }
  SrvReq* r = new SrvReq();
  r->ticket = sv->next_ticket++;
  r->session = rq->session;
  r->term = rq->terms;
  r->n = rq->n;
  r->iters = rq->iters;
  r->lp = to_lp(&rq->params);
  r->has_q = rq->q != nullptr;
  r->has_prev = rq->q_prev != nullptr;
  const int na = T.na;
  if (rq->q) r->q.assign(rq->q, rq->q + (size_t)rq->n * na);
  if (rq->q_prev) r->qprev.assign(rq->q_prev, rq->q_prev + (size_t)rq->n * na);
  if (T.th.n_pose) r->pose.assign(rq->pose_targets, rq->pose_targets + (size_t)rq->n * 7 * T.th.n_pose);
  if (T.th.n_pos) r->pos.assign(rq->pos_targets, rq->pos_targets + (size_t)rq->n * 3 * T.th.n_pos);
  r->t_submit = now_s();
  if (sv->submitted == 0) sv->first_submit = r->t_submit;
  sv->submitted += 1;
  sv->reqs[r->ticket] = r;
  sv->queue.push_back(r);
  *ticket = r->ticket;
  sv->cv_work.notify_one();
  return RK_OK;
}

int rk_server_poll(rk_server* sv, int64_t ticket, int32_t wait_ms, float* q, float* cost, int32_t* success,
                   float* pos_err, float* ang_err, float* latency_ms) {
  if (!sv) return fail(RK_EINVAL, "null server");
  std::unique_lock<std::mutex> lk(sv->mu);
  auto it = sv->reqs.find(ticket);
  if (it == sv->reqs.end()) return fail(RK_EINVAL, "unknown ticket (already collected?)");
  SrvReq* r = it->second;
  auto done = [&] { return r->state >= 2; };
  if (!done()) {
    if (wait_ms < 0) sv->cv_done.wait(lk, done);
    else if (wait_ms > 0) sv->cv_done.wait_for(lk, std::chrono::milliseconds(wait_ms), done);
  }
  if (!done()) return fail(RK_EBUSY, "request still running");
  const SrvTerms& T = sv->terms[r->term];
  if (r->state == 3) {
    const std::string msg = r->err;
    sv->reqs.erase(it);
    delete r;
    return fail(RK_EHIP, msg);
  }
  if (q) memcpy(q, r->q_out.data(), sizeof(float) * r->q_out.size());
  if (cost) memcpy(cost, r->cost.data(), sizeof(float) * r->n);
  if (success) memcpy(success, r->ok.data(), sizeof(int32_t) * r->n);
  if (pos_err && T.th.n_pose) memcpy(pos_err, r->pe.data(), sizeof(float) * r->pe.size());
  if (ang_err && T.th.n_pose) memcpy(ang_err, r->ae.data(), sizeof(float) * r->ae.size());
  if (latency_ms) *latency_ms = (float)((r->t_done - r->t_submit) * 1e3);
  sv->reqs.erase(it);
  delete r;
  return RK_OK;
}

int rk_server_end_session(rk_server* sv, int64_t session) {
  if (!sv) return fail(RK_EINVAL, "null server");
  std::unique_lock<std::mutex> lk(sv->mu);
  auto it = sv->sessions.find(session);
  if (it == sv->sessions.end()) return fail(RK_EINVAL, "unknown session");
  for (SrvReq* r : sv->queue)
    if (r->session == session) return fail(RK_EBUSY, "the session has queued requests");
  if (it->second.busy) return fail(RK_EBUSY, "the session has a request on the device");
  if (it->second.bytes) sv->arena.release(it->second.off, it->second.bytes);
  sv->sessions.erase(it);
  return RK_OK;
}

int rk_server_stats(rk_server* sv, rk_serve_stats* st) {
  if (!sv || !st) return fail(RK_EINVAL, "null argument");
  std::lock_guard<std::mutex> lk(sv->mu);
  memset(st, 0, sizeof(*st));
  st->submitted = sv->submitted;
  st->completed = sv->completed;
  st->failed = sv->failed;
  st->launches = sv->nlaunch;
  st->problems = sv->problems;
  st->converged = sv->converged;
  st->queued = (int64_t)sv->queue.size();
  std::vector<double> l = sv->lat_ms;
  if (!l.empty()) {
    std::sort(l.begin(), l.end());
    auto pct = [&](double q) {
      const double x = q * (l.size() - 1);
      const size_t i = (size_t)x;
      const double f = x - i;
      return i + 1 < l.size() ? l[i] * (1 - f) + l[i + 1] * f : l[i];
    };
    st->p50_ms = pct(0.50);
    st->p99_ms = pct(0.99);
    double s = 0;
    for (double v : l) s += v;
    st->mean_ms = s / l.size();
    st->max_ms = l.back();
  }
  st->mean_problems_per_launch = sv->nlaunch ? (double)sv->problems / sv->nlaunch : 0.0;
  st->budget_bytes = (int64_t)sv->arena.size;
  st->peak_bytes = (int64_t)sv->arena.peak;
  st->bytes_in_use = (int64_t)sv->arena.used;
  st->span_s = sv->completed ? sv->last_done - sv->first_submit : 0.0;
  return RK_OK;
}

int rk_server_destroy(rk_server* sv) {
  if (!sv) return RK_OK;
  {
    std::lock_guard<std::mutex> lk(sv->mu);
    sv->stop = true;
  }
  sv->cv_work.notify_all();
  if (sv->worker.joinable()) sv->worker.join();
  rk_context* ctx = sv->ctx;
  hipSetDevice(ctx->device);
  for (int s = 0; s < sv->nstreams; ++s) {
    if (sv->streams[s]) hipStreamSynchronize(sv->streams[s]);
    if (sv->streams[s]) hipStreamDestroy(sv->streams[s]);
    if (sv->events[s]) hipEventDestroy(sv->events[s]);
    if (sv->h_in[s]) hipHostFree(sv->h_in[s]);
    if (sv->h_out[s]) hipHostFree(sv->h_out[s]);
  }
  for (auto& T : sv->terms) {
    dfree(ctx, T.td, sizeof(Terms));
    if (T.tt) dfree(ctx, T.tt, sizeof(TrajTerms));
  }
  dfree(ctx, sv->arena_base, sv->arena.size);
  for (auto& kv : sv->reqs) delete kv.second;
  delete sv;
  return RK_OK;
}

}  // extern "C"

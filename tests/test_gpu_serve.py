"""The request server (rk_server): continuous batching under a memory budget, on the GPU.

The bar set for serving: each result is the same alone and in the mix -- here bitwise, since every
problem is solved independently by the kernel variant of its term structure, whatever it is packed
with -- and equal to the batch API's solve of the same problems; admission respects the memory
budget; session warm start equals the device-resident warm start of the batch API.
"""
import numpy as np
import pytest

from robokit_amd import native as N
from robokit_amd import workloads as W

pytestmark = pytest.mark.gpu

P = N.LMParams(lm_lambda=1e-2, lambda_factor=10.0)
ITERS = 50


def _robots():
    h = W.humanoid29(0)
    return {"ur10": (W.ur10(), None), "panda": (W.panda_dh(), None), "humanoid": (h, W.humanoid_ee_links(h))}


def _ik_terms(spec, links):
    return N.TermSpec(pose_links=links, pose_weights=np.ones((len(links), 6), np.float32),
                      limit_weights=np.ones(spec.num_actuated_joints, np.float32))


def _requests(seed, count):
    """a mixed stream of IK requests: (kind, IKBatch) with 1..64 problems each"""
    rng = np.random.default_rng(seed)
    robots = _robots()
    out = []
    for i in range(count):
        kind = ["ur10", "panda", "humanoid"][int(rng.integers(3))]
        spec, links = robots[kind]
        links = links or [spec.num_links - 1]
        n = int(np.exp(rng.uniform(0, np.log(65))))
        out.append((kind, W.ik_batch(spec, links, n, seed=1000 * seed + i)))
    return out


def _server(ctx, budget=256 << 20, streams=4):
    sv = N.Server(ctx, budget, streams=streams, max_launch_problems=4096)
    robots = _robots()
    tids = {}
    for kind, (spec, links) in robots.items():
        tids[kind] = sv.add_terms(spec, _ik_terms(spec, links or [spec.num_links - 1]))
    return sv, tids


def test_alone_equals_mixed_equals_batch(ctx):
    reqs = _requests(1, 24)
    sv, tids = _server(ctx)
    alone = []
    for kind, wl in reqs:  # one at a time: each launch holds a single request
        t = sv.submit(tids[kind], ITERS, P, q=wl.q_seed, pose_targets=wl.targets)
        alone.append(sv.poll(t))
    tickets = [sv.submit(tids[kind], ITERS, P, q=wl.q_seed, pose_targets=wl.targets) for kind, wl in reqs]
    mixed = [sv.poll(t) for t in tickets]
    st = sv.stats()
    print(f"\nserver stats: {st}")
    assert st["failed"] == 0 and st["completed"] == 2 * len(reqs)
    assert st["launches"] < 2 * len(reqs)  # the mixed submissions were packed into shared launches
    for (kind, wl), a, m in zip(reqs, alone, mixed):
        assert np.array_equal(a["q"], m["q"]) and np.array_equal(a["cost"], m["cost"]), kind
        assert np.array_equal(a["success"], m["success"])
    # the batch API on the same problems (same kernel variant) gives the same bits
    for (kind, wl), a in list(zip(reqs, alone))[:6]:
        spec = wl.spec
        b = N.Batch(ctx, spec, _ik_terms(spec, wl.links), wl.size)
        b.set(q=wl.q_seed, pose_targets=wl.targets)
        b.solve(ITERS, P)
        ref = b.get()
        assert np.array_equal(ref["q"], a["q"]), kind
        assert np.array_equal(b.metrics()[2], a["success"])
    sv.close()


def test_memory_budget_admission(ctx):
    """a budget that holds only a few launches: requests wait for memory, all complete, the peak
    stays within the budget; a request larger than the whole budget fails cleanly"""
    reqs = _requests(2, 40)
    sv, tids = _server(ctx, budget=1 << 20, streams=4)
    tickets = [sv.submit(tids[kind], 20, P, q=wl.q_seed, pose_targets=wl.targets) for kind, wl in reqs]
    res = [sv.poll(t) for t in tickets]
    st = sv.stats()
    assert st["completed"] == len(reqs) and st["failed"] == 0
    assert 0 < st["peak_bytes"] <= st["budget_bytes"]
    # too big for the budget: fails at poll with a message, the server stays usable
    spec, links = _robots()["humanoid"]
    big = W.ik_batch(spec, links, 4000, seed=9)
    t = sv.submit(tids["humanoid"], 5, P, q=big.q_seed, pose_targets=big.targets)
    with pytest.raises(N.RKError, match="budget"):
        sv.poll(t)
    kind, wl = reqs[0]
    again = sv.poll(sv.submit(tids[kind], 20, P, q=wl.q_seed, pose_targets=wl.targets))
    assert np.array_equal(again["q"], res[0]["q"])
    sv.close()


def test_session_warm_start_equals_batch_advance(ctx):
    clip = W.config2(seed=6, frames=7)
    spec, links = clip.spec, clip.links
    na = spec.num_actuated_joints
    S = 16
    ts = N.TermSpec(pos_links=links, pos_weights=np.ones((len(links), 3), np.float32),
                    limit_weights=np.ones(na, np.float32), smooth_weights=np.full(na, 0.1, np.float32))
    q0 = np.repeat(clip.q_traj[:1], S, 0) + np.random.default_rng(0).normal(0, 0.02, (S, na)).astype(np.float32)
    sv = N.Server(ctx, 64 << 20, streams=2)
    tid = sv.add_terms(spec, ts)
    b = N.Batch(ctx, spec, ts, S)
    for t in range(1, 7):
        tg = np.repeat(clip.targets[t][None], S, 0)
        if t == 1:
            tk = sv.submit(tid, 10, P, q=q0, q_prev=q0, pos_targets=tg, session=42)
            b.set(q=q0, q_prev=q0, pos_targets=tg)
        else:
            tk = sv.submit(tid, 10, P, pos_targets=tg, session=42)  # seeds + reference from the session
            b.advance()
            b.set(pos_targets=tg)
        r = sv.poll(tk)
        b.solve(10, P, from_seed=True)
        assert np.array_equal(r["q"], b.get()["q"]), t
    sv.end_session(42)
    sv.close()


def test_streaming_arrivals_overlap(ctx):
    """requests submitted while others run join later launches; p50/p99 are reported"""
    reqs = _requests(3, 60)
    sv, tids = _server(ctx)
    import time

    tickets = []
    for kind, wl in reqs:
        tickets.append(sv.submit(tids[kind], ITERS, P, q=wl.q_seed, pose_targets=wl.targets))
        time.sleep(0.0005)
    res = [sv.poll(t) for t in tickets]
    st = sv.stats()
    print(f"\nstreaming: {st}")
    assert st["completed"] == len(reqs) and st["p99_ms"] >= st["p50_ms"] > 0
    assert all(r["latency_ms"] > 0 for r in res)


def test_trajectory_requests_equal_batch(ctx):
    """trajectory requests (T = 32) mixed with IK requests: each equals the trajectory batch's solve"""
    tb = W.config3(seed=8, batch=24, T=32)
    sv, tids = _server(ctx)
    ttid = sv.add_traj(tb)
    ik = _requests(4, 8)
    tickets = []
    cuts = [0, 5, 6, 13, 24]
    for i in range(4):
        lo, hi = cuts[i], cuts[i + 1]
        tickets.append(sv.submit(ttid, 10, P, q=tb.q_seed[lo:hi], pose_targets=tb.goals[lo:hi, None]))
        kind, wl = ik[i]
        sv.submit(tids[kind], ITERS, P, q=wl.q_seed, pose_targets=wl.targets)
    res = [sv.poll(t) for t in tickets]
    b = N.TrajBatch(ctx, tb, 24)
    b.set(q=tb.q_seed, goals=tb.goals)
    b.solve(10, P)
    ref = b.get()
    pe, ae, sd = b.metrics()
    ok = (pe < 1e-3) & (ae < 1e-2) & (sd >= 0)
    for i in range(4):
        lo, hi = cuts[i], cuts[i + 1]
        assert np.array_equal(res[i]["q"].reshape(hi - lo, 32, -1), ref["q"][lo:hi]), i
        assert np.array_equal(res[i]["cost"], ref["cost"][lo:hi])
        assert np.array_equal(res[i]["success"], ok[lo:hi])
    sv.close()

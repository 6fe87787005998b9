"""The allocation-free metrics entry point (rk_batch_metrics_noalloc) against the allocating one.

rk_batch_metrics_noalloc stages its device results in the library's static per-GPU array
instead of per-call allocations: its outputs must equal rk_batch_metrics' bit for bit, on calls
that follow each other with different sizes (the array keeps the previous call's contents), and a
batch beyond the array's size is refused with a clear error.  The definition of the metrics is
_ik_metrics_kernel (helpers/ik/ik.py:59-82).
"""
import numpy as np
import pytest

from robokit_amd import native as N
from robokit_amd import workloads as W

pytestmark = pytest.mark.gpu

P = N.LMParams(lm_lambda=1e-2, lambda_factor=10.0)


def _solved(ctx, spec, links, n, seed, iters=15):
    ts = N.TermSpec(pose_links=links, pose_weights=np.ones((len(links), 6), np.float32),
                    limit_weights=np.ones(spec.num_actuated_joints, np.float32))
    wl = W.ik_batch(spec, links, n, seed=seed)
    b = N.Batch(ctx, spec, ts, n)
    b.set(q=wl.q_seed, pose_targets=wl.targets)
    b.solve(iters, P)
    return b


def test_noalloc_metrics_equal_the_allocating_entry(ctx):
    h = W.humanoid29(0)
    big = _solved(ctx, h, W.humanoid_ee_links(h), 700, seed=31)  # 4 targets per problem
    p = W.panda_dh()
    small = _solved(ctx, p, [p.num_links - 1], 37, seed=32)  # then a smaller call, 1 target
    for b in (big, small, big):
        for thr in ((1e-3, 1e-2), (5e-3, 5e-2)):
            ref = b.metrics(*thr)
            got = b.metrics(*thr, noalloc=True)
            for r, g in zip(ref, got):
                assert r.shape == g.shape and np.array_equal(r, g)
    assert big.metrics(1e9, 1e9, noalloc=True)[2].all()  # thresholds reach the kernel


def test_noalloc_metrics_refuses_beyond_its_staging(ctx):
    p = W.panda_dh()
    n = 65536 + 1
    ts = N.TermSpec(pose_links=[p.num_links - 1], pose_weights=np.ones((1, 6), np.float32))
    b = N.Batch(ctx, p, ts, n)
    q = np.zeros((n, 7), np.float32)
    b.set(q=q, pose_targets=np.tile(np.array([0, 0, 0.5, 1, 0, 0, 0], np.float32), (n, 1, 1)))
    with pytest.raises(ValueError, match="exceed"):
        b.metrics(noalloc=True)
    b.set(q=q[:65536], pose_targets=np.tile(np.array([0, 0, 0.5, 1, 0, 0, 0], np.float32), (65536, 1, 1)))
    ref, got = b.metrics(), b.metrics(noalloc=True)  # the largest accepted size
    assert all(np.array_equal(r, g) for r, g in zip(ref, got))

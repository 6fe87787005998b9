"""GPU parity of the collision-aware trajectory LM (banded kernel) against the trajectory oracle.

Same bar as the IK path: every GPU step re-executed by the FP32 oracle and its FP64 twin from
the GPU's own iterate (residual norms within 1e-4 relative where FP32 itself is; proposal-norm
error distribution and accept flips no worse than the FP32 oracle's), the cost at the seeds to
1e-5 relative, and the goal / clearance metrics of the solutions to FP32 rounding.
"""
import numpy as np
import pytest

from oracle import oracle as O
from robokit_amd import native as N
from robokit_amd import workloads as W
from tests.parity import shadow_check

pytestmark = pytest.mark.gpu

P = N.LMParams(lm_lambda=1e-2, lambda_factor=10.0)


# T = 64 / 16: the band solve split over the two wavefronts (nested dissection, equal halves);
# T = 7: unequal halves (2 frames before the separator, 3 after); T = 5: too short to split, one
# wavefront solves it whole.  The UR10 cases run the 6-DOF kernel; 13 spheres + 3 boxes leave
# the padded obstacle groups partly filled; the last case fills every table to its maximum
# (64 robot spheres, 64 obstacle spheres, 16 boxes, 64 frames).
@pytest.mark.parametrize("T,B,iters,arm,ns,nb,spl", [(64, 12, 12, "panda", 32, 8, 6), (16, 24, 20, "panda", 32, 8, 6),
                                                     (7, 16, 12, "panda", 32, 8, 6), (5, 16, 12, "panda", 13, 3, 6),
                                                     (16, 16, 12, "ur10", 13, 3, 6), (64, 6, 8, "ur10", 64, 16, 8)])
def test_traj_parity(ctx, T, B, iters, arm, ns, nb, spl):
    tb = W.config3(seed=5, batch=B, T=T, n_obs_spheres=ns, n_obs_boxes=nb, arm=arm, spheres_per_link=spl)
    b = N.TrajBatch(ctx, tb, B, trace_iters=iters)
    b.set(q=tb.q_seed, goals=tb.goals)
    b.solve(iters, P)
    gpu = b.get(traces=True)
    q_it, lam_it = b.iterates()
    ot = O.OracleTraj(tb)
    rows = np.repeat(np.arange(B), iters)
    qi = q_it.reshape(B * iters, T, -1)
    p = O.LMParams(**P.__dict__)
    s32 = O.traj_lm_step(ot, qi, lam_it.reshape(-1), p, rows=rows)
    s64 = O.traj_lm_step(ot, qi, lam_it.reshape(-1), p, dtype="f64", rows=rows)
    rep = shadow_check(gpu["cost_trace"], gpu["accept"], s32, s64)
    print(f"\ntrajectory {arm} T={T} B={B} scene {ns}+{nb}: {rep}")
    assert rep.ok(), rep
    # the seed cost (first trace entry) against the FP64 oracle
    r64 = np.array([O.traj_residual_jacobian(ot, i, tb.q_seed[i], dtype="f64", jac=False)[4].sum() for i in range(B)])
    np.testing.assert_allclose(gpu["cost_trace"][:, 0], r64, rtol=1e-5)
    # metrics of the GPU solution
    pe, ae, sd = b.metrics()
    pe_o, ae_o, sd_o = O.traj_metrics(ot, gpu["q"], dtype="f64")
    np.testing.assert_allclose(pe, pe_o, atol=2e-5)
    np.testing.assert_allclose(ae, ae_o, atol=1e-4)
    np.testing.assert_allclose(sd, sd_o, atol=2e-5)


def test_traj_cost_only_and_continue(ctx):
    """iters = 0 scores the seeds; two continued halves equal one launch bitwise"""
    tb = W.config3(seed=6, batch=8, T=32)
    b = N.TrajBatch(ctx, tb, 8)
    b.set(q=tb.q_seed, goals=tb.goals)
    b.solve(0, P)
    c0 = b.get()["cost"]
    ot = O.OracleTraj(tb)
    ref = np.array([O.traj_residual_jacobian(ot, i, tb.q_seed[i], dtype="f64", jac=False)[4].sum() for i in range(8)])
    np.testing.assert_allclose(c0, ref, rtol=1e-5)
    b.solve(10, P, from_seed=True)
    one = b.get()
    b.set(q=tb.q_seed, goals=tb.goals)
    b.solve(5, P)
    b.solve(5, P, cont=True)
    two = b.get()
    assert np.array_equal(one["q"], two["q"]) and np.array_equal(one["cost"], two["cost"])


def test_traj_lm_improves_clearance(ctx):
    """the solve drives the trajectories out of collision and towards the goal (behavioural check
    on the full workload shape)"""
    tb = W.config3(seed=7, batch=64, T=64)
    b = N.TrajBatch(ctx, tb, 64)
    b.set(q=tb.q_seed, goals=tb.goals)
    pe0, ae0, sd0 = b.metrics()
    b.solve(50, P)
    pe, ae, sd = b.metrics()
    print(f"\nmin clearance median {np.median(sd0):.3f} -> {np.median(sd):.3f} m, goal error median "
          f"{np.median(pe0):.3f} -> {np.median(pe):.4f} m")
    assert np.median(sd) > np.median(sd0) and np.median(pe) < np.median(pe0)

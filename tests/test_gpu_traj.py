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


def _bounded(tb, orders=0b111, w=(3.0, 2.0, 1.5), half=(1.2, 0.15, 0.05)):
    """TrajectoryLimitTask rows (forward stencils) of every order, tight enough to be active"""
    import dataclasses

    na = tb.spec.num_actuated_joints
    lo = np.stack([np.full(na, -half[0]), np.full(na, -half[1]), np.full(na, -half[2])]).astype(np.float32)
    wts = np.stack([np.full(na, wi) for wi in w]).astype(np.float32)
    return dataclasses.replace(tb, lim_orders=orders, lim_w=wts, lim_lo=lo, lim_hi=-lo)


# bound rows of orders 0, 1, 2 together and alone; the 6-dof arm; a 64-frame problem whose band
# solve is split over both wavefronts
@pytest.mark.parametrize("T,B,iters,arm,orders,half", [(16, 16, 15, "panda", 0b111, (1.2, 0.15, 0.05)),
                                                       (16, 12, 12, "panda", 0b001, (1.2, 0.15, 0.05)),
                                                       (7, 12, 12, "panda", 0b100, (1.2, 0.15, 0.001)),
                                                       (16, 12, 12, "ur10", 0b110, (1.2, 0.02, 0.001)),
                                                       (64, 6, 10, "panda", 0b111, (1.2, 0.15, 0.05))])
def test_traj_bound_rows_parity(ctx, T, B, iters, arm, orders, half):
    """TrajectoryLimitTask bound rows in the trajectory kernel (its LIM instantiation): every GPU step
    shadow-checked against the trajectory oracle's rows (trajectory_limit_task.py:436-498), the seed
    cost to 1e-5 relative of FP64; continued halves equal one launch bitwise"""
    tb = _bounded(W.config3(seed=9, batch=B, T=T, arm=arm), orders=orders, half=half)
    b = N.TrajBatch(ctx, tb, B, trace_iters=iters)
    b.set(q=tb.q_seed, goals=tb.goals)
    b.solve(iters, P)
    gpu = b.get(traces=True)
    q_it, lam_it = b.iterates()
    ot = O.OracleTraj(tb)
    rows = np.repeat(np.arange(B), iters)
    qi = q_it.reshape(B * iters, T, -1)
    p = O.LMParams(**P.__dict__)
    rep = shadow_check(gpu["cost_trace"], gpu["accept"], O.traj_lm_step(ot, qi, lam_it.reshape(-1), p, rows=rows),
                       O.traj_lm_step(ot, qi, lam_it.reshape(-1), p, dtype="f64", rows=rows))
    ct = np.array([O.traj_residual_jacobian(ot, i, tb.q_seed[i], dtype="f64", jac=False)[4] for i in range(B)])
    cf = np.array([O.traj_residual_jacobian(ot, i, gpu["q"][i], dtype="f64", jac=False)[4] for i in range(B)])
    share = max(ct[:, 4:].sum() / ct.sum(), cf[:, 4:].sum() / cf.sum())
    print(f"\nbounds {arm} T={T} orders {orders:03b}: {rep}; bound cost share at the seeds "
          f"{ct[:, 4:].sum() / ct.sum():.3f}, at the solutions {cf[:, 4:].sum() / cf.sum():.3f}")
    assert rep.ok(), rep
    assert share > 0.005  # the bound rows carry weight on the path: the check is not vacuous
    np.testing.assert_allclose(gpu["cost_trace"][:, 0], ct.sum(1), rtol=1e-5)
    half = iters // 2
    b.set(q=tb.q_seed, goals=tb.goals)
    b.solve(half, P)
    b.solve(iters - half, P, cont=True)
    assert np.array_equal(b.get()["q"], gpu["q"])


def test_traj_bounds_keep_trajectories_inside(ctx):
    """position bounds (a hinge of weight 50) on the benchmark's trajectory shape: the unbounded
    solve leaves |q| <= 1 far behind while every frame of every bounded solution stays within the
    bound up to the soft hinge's balance against the goal rows (5 %)"""
    tb = W.config3(seed=10, batch=32, T=64)
    na = tb.spec.num_actuated_joints
    bnd = _bounded(tb, orders=0b001, w=(50.0, 0.0, 0.0), half=(1.0, 0.0, 0.0))
    q0 = np.clip(tb.q_seed, -1.0, 1.0)
    free = N.TrajBatch(ctx, tb, 32)
    free.set(q=q0, goals=tb.goals)
    free.solve(50, P)
    bb = N.TrajBatch(ctx, bnd, 32)
    bb.set(q=q0, goals=tb.goals)
    bb.solve(50, P)
    qf, qb = free.get()["q"], bb.get()["q"]
    print(f"\nmax |q| unbounded {np.abs(qf).max():.3f}, bounded {np.abs(qb).max():.3f} (bound 1.0, {na} dofs)")
    assert np.abs(qf).max() > 1.5
    assert np.abs(qb).max() < 1.05


# conjugate-gradient steps (LMOptimizerConfig(linear_solver="cg")): the reference's default budget
# max(4, min(64, D // 4)) on 16 and 64 frames, a short problem, the 6-dof arm, bound rows.  FP32 CG
# loses conjugacy within a few iterations, so accept decisions near rho_min flip against FP64 about
# as often for the oracle's FP32 CG as for the GPU's; the short problem takes 48 trajectories so
# the flip count is a statistic rather than one or two events
@pytest.mark.parametrize("T,B,iters,arm,bounds,cg", [(16, 16, 15, "panda", False, None), (64, 8, 10, "panda", False, None),
                                                    (5, 48, 12, "panda", False, None), (16, 12, 12, "ur10", True, None),
                                                    (16, 12, 10, "panda", True, 6)])
def test_traj_cg_parity(ctx, T, B, iters, arm, bounds, cg):
    """the register-resident CG step of the trajectory kernel: every GPU step shadow-checked against
    the oracle's matrix-free CG over the sparse rows (opt/lm_optimizer.py:781-832, same budget), the
    seed cost to 1e-5 relative of FP64; continued halves equal one launch bitwise"""
    import dataclasses

    tb = W.config3(seed=21, batch=B, T=T, arm=arm)
    if bounds:
        tb = _bounded(tb)
    tb = dataclasses.replace(tb, cg_iters=cg or tb.cg_default_iters)
    b = N.TrajBatch(ctx, tb, B, trace_iters=iters)
    b.set(q=tb.q_seed, goals=tb.goals)
    b.solve(iters, P)
    gpu = b.get(traces=True)
    q_it, lam_it = b.iterates()
    ot = O.OracleTraj(tb)
    rows = np.repeat(np.arange(B), iters)
    qi = q_it.reshape(B * iters, T, -1)
    p = O.LMParams(**P.__dict__)
    rep = shadow_check(gpu["cost_trace"], gpu["accept"], O.traj_lm_step(ot, qi, lam_it.reshape(-1), p, rows=rows),
                       O.traj_lm_step(ot, qi, lam_it.reshape(-1), p, dtype="f64", rows=rows))
    print(f"\nCG {arm} T={T} budget {tb.cg_iters} bounds {bounds}: {rep}")
    assert rep.ok(), rep
    r64 = np.array([O.traj_residual_jacobian(ot, i, tb.q_seed[i], dtype="f64", jac=False)[4].sum() for i in range(B)])
    np.testing.assert_allclose(gpu["cost_trace"][:, 0], r64, rtol=1e-5)
    assert gpu["accept"].mean() > 0.3  # the steps are real descent steps, not rejected noise
    half = iters // 2
    b.set(q=tb.q_seed, goals=tb.goals)
    b.solve(half, P)
    b.solve(iters - half, P, cont=True)
    assert np.array_equal(b.get()["q"], gpu["q"])


def test_traj_cg_solves_the_workload(ctx):
    """the benchmark's trajectory shape (64 frames, 7 dofs, the scene) on CG steps with the default
    budget: clearance and goal error improve as with the band Cholesky"""
    import dataclasses

    tb = W.config3(seed=7, batch=64, T=64)
    tb = dataclasses.replace(tb, cg_iters=tb.cg_default_iters)
    b = N.TrajBatch(ctx, tb, 64)
    b.set(q=tb.q_seed, goals=tb.goals)
    pe0, ae0, sd0 = b.metrics()
    b.solve(50, P)
    pe, ae, sd = b.metrics()
    print(f"\nCG: min clearance median {np.median(sd0):.3f} -> {np.median(sd):.3f} m, goal error median "
          f"{np.median(pe0):.3f} -> {np.median(pe):.4f} m")
    assert np.median(sd) > np.median(sd0) and np.median(pe) < np.median(pe0)


def test_group_packs_uneven_horizons(ctx):
    """an uneven group of IK batches (6 / 7 / 29 dofs) and trajectory batches of horizons 16, 32 and
    64 (6- and 7-dof arms, with and without bound rows): every batch's result equals its solve
    alone, bitwise, eagerly and from two replays of the captured graph; the kernel times list one
    trajectory launch per arm class with its batches"""
    from robokit_amd.native import TPP_TRAJ

    iters = 15
    mb = W.config1(seed=3, total=300)
    ik, ik_single = [], []
    for bk in mb.buckets:
        ts = N.TermSpec(pose_links=bk.links, pose_weights=np.ones((len(bk.links), 6), np.float32),
                        limit_weights=np.ones(bk.spec.num_actuated_joints, np.float32))
        b = N.Batch(ctx, bk.spec, ts, bk.size, trace_iters=iters)
        b.set(q=bk.q_seed, pose_targets=bk.targets)
        b.solve(iters, P)
        ik_single.append(b.get(traces=True))
        b.set(q=bk.q_seed, pose_targets=bk.targets)
        ik.append(b)
    tbs = [W.config3(seed=31, batch=8, T=16), W.config3(seed=32, batch=4, T=64),
           W.config3(seed=33, batch=6, T=32, arm="ur10"), _bounded(W.config3(seed=34, batch=5, T=16))]
    tj, tj_single = [], []
    for tb in tbs:
        b = N.TrajBatch(ctx, tb, tb.size, trace_iters=iters)
        b.set(q=tb.q_seed, goals=tb.goals)
        b.solve(iters, P)
        tj_single.append(b.get(traces=True))
        b.set(q=tb.q_seed, goals=tb.goals)
        tj.append(b)
    g = N.Group(ctx, ik + tj)
    print(f"\n[group] {len(ik)} IK batches + trajectories T = {[tb.T for tb in tbs]}", flush=True)

    def check():
        for b, ref in zip(ik + tj, ik_single + tj_single):
            out = b.get(traces=True)
            assert np.array_equal(out["q"], ref["q"])
            assert np.array_equal(out["cost_trace"], ref["cost_trace"])

    g.solve(iters, P, use_graph=False, from_seed=True)
    check()
    kt = g.kernel_times()
    traj = [k for k in kt if k[0] == TPP_TRAJ]
    # classes: 6 dofs (the UR10), 7 dofs, 7 dofs with bound rows
    assert sorted(sorted(k[2]) for k in traj) == [[3, 4], [5], [6]], kt
    assert all(k[1] > 0 for k in kt)
    for _ in range(2):
        g.solve(iters, P, use_graph=True, from_seed=True)
        check()

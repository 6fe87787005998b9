"""The reference-shaped solver API (robokit_amd.lm): construction and validation on the CPU,
solves on the GPU through the C-ABI.

Mirrors the reference's tests of the LM lifecycle: config defaults
(tests/opt/test_optimizer_acceptance.py:66-67), capacity errors (opt/optimizer.py:321-322),
unsupported residual modes raise ValueError (frame_task.py:87-88, position_limit.py:225-226),
iteration callbacks see every step (opt/optimizer.py:226-236), early stopping at cost_tol
(opt/optimizer.py:237-239).
"""
import numpy as np
import pytest

from oracle import oracle as O
from robokit_amd import lm as L
from robokit_amd import native as N
from robokit_amd import workloads as W


def _ik_problem(B=64, seed=0, spread=0.4):
    spec = W.panda_dh()
    robot = L.Robot(spec)
    wl = W.ik_batch(spec, [spec.num_links - 1], B, seed=seed)
    lim = spec.actuated_joint_limits_f32
    q0 = np.clip(wl.q_target + np.random.default_rng(seed).uniform(-spread, spread, wl.q_target.shape),
                 lim[:, 0], lim[:, 1]).astype(np.float32)
    return robot, wl, q0


# ----------------------------------------------------------------------------- CPU
def test_config_defaults_match_reference():
    c = L.LMOptimizerConfig()
    assert (c.max_iter, c.early_stopping_interval, c.cost_tol, c.cuda_graph_mode) == (100, 1, 1e-9, "none")
    assert (c.lm_lambda, c.lambda_factor, c.lambda_min, c.lambda_max) == (1e-3, 2.0, 1e-5, 1e10)
    assert (c.rho_min, c.gain_ratio_epsilon, c.linear_solver) == (1e-3, 1e-8, "cholesky")


def test_term_validation():
    robot = L.Robot(W.ur10())
    T = np.array([0.5, 0.1, 0.4, 1, 0, 0, 0], np.float32)
    with pytest.raises(ValueError):
        L.FrameTask(robot, 7, T, residual_mode="bogus")
    with pytest.raises(NotImplementedError):
        L.FrameTask(robot, 7, T, residual_mode="distance")
    with pytest.raises(ValueError):
        L.PositionLimit(robot, residual_mode="bogus")
    with pytest.raises(ValueError):
        L.FrameTask(robot, 99, T)
    with pytest.raises(ValueError):
        L.FrameTask(robot, [6, 7], T[None])  # one target for two frames
    ft = L.FrameTask(robot, 7, T, position_weight=2.0, orientation_weight=0.5)
    np.testing.assert_array_equal(ft.weights, [[2, 2, 2, 0.5, 0.5, 0.5]])
    assert ft.T_world_target.shape == (1, 1, 7) and ft.residual_dim == 6
    pt = L.PositionTask(robot, [5, 7], np.zeros((2, 3)), weight=[1, 2, 3])
    assert pt.weights.shape == (2, 3) and pt.T_world_target.shape == (1, 2, 3)
    assert L.PositionTask(robot, 7, T).weights[0, 0] == 20.0  # reference default weight
    np.testing.assert_array_equal(L.PositionLimit(robot).weights, np.ones(6))
    with pytest.raises(ValueError):
        L.RobotState(robot, np.zeros((3, 5)))


# ----------------------------------------------------------------------------- GPU
@pytest.mark.gpu
def test_optimizer_matches_batch_api(ctx):
    robot, wl, q0 = _ik_problem()
    cfg = L.LMOptimizerConfig(max_iter=30, early_stopping_interval=0, lm_lambda=1e-2, lambda_factor=10.0)
    terms = [L.FrameTask(robot, wl.links[0], wl.targets[:, 0]), L.PositionLimit(robot, weight=1.0)]
    var = L.VarValues({L.VarKey.ROBOT: robot.state(q0)})
    var, costs = cfg.build(terms).solve(var)
    b = N.Batch(ctx, robot.spec, N.TermSpec(pose_links=wl.links, pose_weights=np.ones((1, 6), np.float32),
                                            limit_weights=np.ones(7, np.float32)), 64)
    b.set(q=q0, pose_targets=wl.targets)
    b.solve(30, cfg.params())
    ref = b.get()
    np.testing.assert_array_equal(var.get(L.VarKey.ROBOT).q, ref["q"])
    np.testing.assert_array_equal(costs, ref["cost"])


@pytest.mark.gpu
def test_steps_continue_the_solve(ctx):
    """early_stopping_interval 5 runs the solve as continued chunks; with no early stop it equals
    one 30-iteration launch bitwise (damping and iterate carry over on the device)."""
    robot, wl, q0 = _ik_problem(seed=1, spread=2.0)
    terms = [L.FrameTask(robot, wl.links[0], wl.targets[:, 0]), L.PositionLimit(robot)]
    base = dict(max_iter=30, lm_lambda=1e-2, lambda_factor=10.0, cost_tol=0.0)
    v1 = L.VarValues({L.VarKey.ROBOT: robot.state(q0)})
    v2 = L.VarValues({L.VarKey.ROBOT: robot.state(q0)})
    _, c1 = L.LMOptimizerConfig(early_stopping_interval=0, **base).build(terms).solve(v1)
    _, c2 = L.LMOptimizerConfig(early_stopping_interval=5, **base).build(terms).solve(v2)
    np.testing.assert_array_equal(v1.get(L.VarKey.ROBOT).q, v2.get(L.VarKey.ROBOT).q)
    np.testing.assert_array_equal(c1, c2)


@pytest.mark.gpu
def test_iteration_callback_and_cost_trace(ctx):
    robot, wl, q0 = _ik_problem(seed=2)
    terms = [L.FrameTask(robot, wl.links[0], wl.targets[:, 0]), L.PositionLimit(robot)]
    cfg = L.LMOptimizerConfig(max_iter=12, early_stopping_interval=0, lm_lambda=1e-2, lambda_factor=10.0)
    seen = []
    var = L.VarValues({L.VarKey.ROBOT: robot.state(q0)})
    cfg.build(terms).solve(var, iteration_callback=lambda i, v, c: seen.append((i, c.copy())))
    assert [i for i, _ in seen] == list(range(13))
    costs = np.stack([c for _, c in seen])
    assert np.all(np.diff(costs, axis=0) <= 0)  # LM never accepts an uphill step
    # step 0 is the cost at the seeds: the oracle's seed cost
    ora = O.lm_solve(robot.spec, O.ik_terms(wl.links, wl.targets, 7), q0, 0, O.LMParams())
    np.testing.assert_allclose(costs[0], ora["cost"], rtol=1e-5, atol=1e-9)
    with pytest.raises(ValueError):
        L.LMOptimizerConfig(cuda_graph_mode="full").build(terms).solve(var, iteration_callback=lambda *a: None)


@pytest.mark.gpu
def test_early_stop_and_capacity(ctx):
    robot, wl, q0 = _ik_problem(seed=3, spread=0.05)
    terms = [L.FrameTask(robot, wl.links[0], wl.targets[:, 0]), L.PositionLimit(robot)]
    calls = []
    cfg = L.LMOptimizerConfig(max_iter=100, early_stopping_interval=1, cost_tol=1e-6, lm_lambda=1e-2,
                              lambda_factor=10.0)
    opt = cfg.build(terms)
    var = L.VarValues({L.VarKey.ROBOT: robot.state(q0)})
    _, costs = opt.solve(var, iteration_callback=lambda i, v, c: calls.append(i))
    assert np.all(costs <= 1e-6) and len(calls) < 101  # stopped early once every cost met cost_tol
    # a smaller batch is a prefix; a larger one is refused
    small = L.VarValues({L.VarKey.ROBOT: robot.state(q0[:10])})
    terms[0].set_target(wl.targets[:10, 0])
    opt.solve(small)
    big = L.VarValues({L.VarKey.ROBOT: robot.state(np.concatenate([q0, q0]))})
    terms[0].set_target(np.concatenate([wl.targets[:, 0]] * 2))
    with pytest.raises(ValueError, match="exceeds"):
        opt.solve(big)


@pytest.mark.gpu
def test_ik_metrics_api(ctx):
    robot, wl, q0 = _ik_problem(seed=4)
    pe, ae, ok = L.ik_metrics(robot, wl.q_target, wl.links, wl.targets[:, 0])
    assert ok.all() and pe.max() < 1e-5
    pe2, ae2, ok2 = O.ik_metrics(robot.spec, q0, wl.links, wl.targets)
    pe3, ae3, ok3 = L.ik_metrics(robot, q0, wl.links, wl.targets[:, 0])
    np.testing.assert_allclose(pe3, pe2, atol=2e-6)
    np.testing.assert_array_equal(ok3, ok2)


@pytest.mark.gpu
def test_rotation_position_tasks_solve(ctx):
    """the presets.basic pair (PositionTask + RotationTask) through the solver API reaches the targets"""
    robot, wl, q0 = _ik_problem(seed=5, spread=0.3)
    link = wl.links[0]
    T = wl.targets[:, 0]
    terms = [L.PositionTask(robot, link, T, weight=20.0), L.RotationTask(robot, link, T, weight=10.0),
             L.PositionLimit(robot)]
    cfg = L.LMOptimizerConfig(max_iter=50, early_stopping_interval=0, lm_lambda=1e-2, lambda_factor=10.0)
    var = L.VarValues({L.VarKey.ROBOT: robot.state(q0)})
    var, costs = cfg.build(terms).solve(var)
    pe, ae, ok = L.ik_metrics(robot, var.get(L.VarKey.ROBOT).q, wl.links, T)
    assert ok.mean() > 0.8, (pe.max(), ae.max())


def _collision_robot():
    tb = W.config3(seed=12, batch=2, T=5)
    sc = tb.scene
    robot = L.Robot(tb.spec, collision_spheres=(sc.sphere_links, sc.sphere_centers, sc.sphere_radii))
    scene = L.PrimitiveScene(sc.obstacle_spheres, sc.obstacle_boxes)
    return robot, scene


def test_scene_collision_task_validation():
    """SceneCollisionTask (terms/dense/scene_collision_task.py:281-356): the penalty is checked, a
    robot without collision spheres is refused (RuntimeError), a robot-free construction is a
    template, sphere subsets and per-sphere weights map onto the term's rows"""
    robot, scene = _collision_robot()
    with pytest.raises(ValueError):
        L.SceneCollisionTask(robot, scene, penalty="bogus")
    with pytest.raises(RuntimeError):
        L.SceneCollisionTask(L.Robot(W.panda_dh()), scene)
    assert L.SceneCollisionTask(scene=scene, weight=2.0).is_template
    t = L.SceneCollisionTask(robot, scene, weight=np.arange(1, 4, dtype=np.float32), sphere_indices=[0, 5, 7],
                             margin=0.05)
    assert t.residual_dim == 3 and np.array_equal(t.weights, [1, 2, 3])
    assert list(t.sphere_links) == list(robot.collision_spheres[0][[0, 5, 7]])
    with pytest.raises(ValueError):
        L.SceneCollisionTask(robot, scene, weight=[1.0, 2.0], sphere_indices=[0, 5, 7])
    with pytest.raises(NotImplementedError):
        L.SceneCollisionTask(robot, scene, residual_mode="sqrt_abs")
    assert scene.num_geoms == len(scene.spheres) + len(scene.boxes)


@pytest.mark.gpu
def test_scene_collision_through_the_optimizer(ctx):
    """FrameTask + PositionLimit + SceneCollisionTask through LMOptimizer equals the batch API's
    solve of the same term structure (the register kernel with collision rows), bitwise"""
    robot, scene = _collision_robot()
    spec = robot.spec
    wl = W.ik_batch(spec, [spec.num_links - 1], 48, seed=21)
    terms = [L.FrameTask(robot, wl.links[0], wl.targets[:, 0]), L.PositionLimit(robot),
             L.SceneCollisionTask(robot, scene, weight=2.0, margin=0.05)]
    cfg = L.LMOptimizerConfig(max_iter=20, early_stopping_interval=0, lm_lambda=1e-2, lambda_factor=10.0)
    var = L.VarValues({L.VarKey.ROBOT: robot.state(wl.q_seed)})
    var, costs = cfg.build(terms).solve(var)
    links, centers, radii = robot.collision_spheres
    ts = N.TermSpec(pose_links=wl.links, pose_weights=np.ones((1, 6), np.float32), limit_weights=np.ones(7, np.float32),
                    coll_links=list(links), coll_centers=centers, coll_radii=radii,
                    coll_weights=np.full(len(links), 2.0, np.float32), coll_margin=0.05,
                    obstacle_spheres=scene.spheres, obstacle_boxes=scene.boxes)
    b = N.Batch(ctx, spec, ts, 48)
    b.set(q=wl.q_seed, pose_targets=wl.targets)
    b.solve(20, cfg.params())
    ref = b.get()
    np.testing.assert_array_equal(var.get(L.VarKey.ROBOT).q, ref["q"])
    np.testing.assert_array_equal(costs, ref["cost"])


# ----------------------------------------------------------------------------- trajectory terms, variables
def _traj_terms(tb, **kw):
    """the reference-shaped terms of a workloads.TrajBatch problem set"""
    sc = tb.scene
    robot = L.Robot(tb.spec, collision_spheres=(sc.sphere_links, sc.sphere_centers, sc.sphere_radii))
    scene = L.PrimitiveScene(sc.obstacle_spheres, sc.obstacle_boxes)
    goal = L.TrajectoryTask(L.FrameTask(robot, tb.goal_link, tb.goals[:, None]), tb.T, [-1])
    terms = [goal, L.TrajectorySmoothnessTask(robot, tb.T, tb.wv, order=1, dt=tb.dt),
             L.TrajectorySmoothnessTask(robot, tb.T, tb.wa, order=2, dt=tb.dt),
             L.TrajectoryCollisionTask(robot, scene, tb.T, weight=tb.coll_w, margin=tb.margin)]
    if tb.lim_orders:
        orders = [o for o in range(3) if tb.lim_orders >> o & 1]
        terms.append(L.TrajectoryLimitTask(robot, tb.T, orders, lower=tb.lim_lo[orders], upper=tb.lim_hi[orders],
                                           weight=tb.lim_w[orders], dt=tb.dt, stencil="forward"))
    return robot, terms


def test_trajectory_terms_build_the_kernel_problem_set():
    """the optimizer's trajectory description equals the workload's: goal link and weights, the
    difference weights, the collision model, the CG budget of the reference's default formula"""
    tb = W.config3(seed=3, batch=4, T=16)
    robot, terms = _traj_terms(tb)
    opt = L.LMOptimizerConfig(linear_solver="cg").build(terms)
    d = opt._traj
    assert (d.T, d.goal_link, d.wv, d.wa, d.dt, d.coll_w, d.margin) == (16, tb.goal_link, tb.wv, tb.wa, 1.0, 20.0, 0.02)
    np.testing.assert_array_equal(d.scene.sphere_links, tb.scene.sphere_links)
    np.testing.assert_array_equal(d.scene.obstacle_boxes, tb.scene.obstacle_boxes)
    assert d.cg_iters == max(4, min(64, 16 * 7 // 4)) == 28
    np.testing.assert_array_equal(opt._goal_weights, np.ones(6))
    assert L.LMOptimizerConfig(linear_solver="cg", cg_max_iter=9).build(terms)._traj.cg_iters == 9
    assert L.LMOptimizer(terms)._traj.cg_iters == 0
    assert sum(t.residual_dim for t in terms) == 6 + 15 * 7 + 14 * 7 + 16 * 48


def test_trajectory_term_validation():
    tb = W.config3(seed=3, batch=2, T=8)
    robot, terms = _traj_terms(tb)
    with pytest.raises(NotImplementedError):
        L.TrajectoryTask(L.FrameTask(robot, tb.goal_link, tb.goals[:1]), 8, [3])  # goal on a middle frame
    with pytest.raises(ValueError):
        L.TrajectorySmoothnessTask(robot, 8, order=8)
    with pytest.raises(NotImplementedError):
        L.TrajectorySmoothnessTask(robot, 8, order=3)
    with pytest.raises(NotImplementedError):
        L.TrajectoryCollisionTask(robot, L.PrimitiveScene(), 8, penalty="smooth")
    with pytest.raises(NotImplementedError):
        L.TrajectoryLimitTask(robot, 8, order=1, upper=1.0)  # central stencil by default
    with pytest.raises(ValueError):
        L.TrajectoryLimitTask(robot, 8, order=2, stencil="forward")  # no default acceleration bound
    lt = L.TrajectoryLimitTask(robot, 8, order=[0, 1], upper=(robot.spec.actuated_joint_limits[:, 1], 2.0),
                               stencil="forward")
    np.testing.assert_array_equal(lt.lower[1], -2.0 * np.ones(7))  # derivatives: -upper
    np.testing.assert_array_equal(lt.lower[0], robot.spec.actuated_joint_limits[:, 0].astype(np.float32))
    with pytest.raises(ValueError):  # horizons disagree
        L.LMOptimizer(terms + [L.TrajectorySmoothnessTask(robot, 9, order=1)])
    with pytest.raises(NotImplementedError):  # mixed dense and trajectory terms
        L.LMOptimizer(terms + [L.PositionLimit(robot)])
    with pytest.raises(NotImplementedError):
        L.LMOptimizer([L.PositionLimit(robot)], L.LMOptimizerConfig(linear_solver="cg"))
    with pytest.raises(NotImplementedError):
        L.LMOptimizer(terms, L.LMOptimizerConfig(linear_solver="torch_cholesky"))
    tpl = L.TrajectorySmoothnessTask(num_frames=8, order=2)
    assert tpl.is_template and tpl.resolve(robot, {}).order == 2


def test_var_values_column_blocks():
    """VarValues (opt/var_values.py:88-248): column offsets in insertion order, integrate splits the
    velocity by block, view shares storage, gather / accept per key; VectorVar's unit and max_step
    follow the reference doctest (opt/var_values.py:255-266)"""
    robot = L.Robot(W.panda_dh())
    st = robot.state(np.zeros((3, 7)))
    vv = L.VarValues({L.VarKey.ROBOT: st, L.VarKey.TIMING: L.VectorVar(np.ones((3, 2)))})
    assert (vv.tangent_dim, vv.tangent_offset(L.VarKey.ROBOT), vv.tangent_offset(L.VarKey.TIMING)) == (9, 0, 7)
    v = np.arange(27, dtype=np.float32).reshape(3, 9)
    out = vv.integrate(v)
    np.testing.assert_array_equal(out[L.VarKey.ROBOT].q, v[:, :7])
    np.testing.assert_array_equal(out[L.VarKey.TIMING].x, 1 + v[:, 7:])
    np.testing.assert_array_equal(vv.difference(out), v)
    view = vv.view(2, 1)
    view[L.VarKey.TIMING].x[0, 0] = 5.0
    assert vv[L.VarKey.TIMING].x[1, 0] == 5.0
    g = vv.gather([2, 0])
    assert g.batch_size == 2 and g[L.VarKey.TIMING].x[1, 0] == 1.0
    vv.accept(np.array([0, 1, 0]), out)
    np.testing.assert_array_equal(vv[L.VarKey.ROBOT].q[1], v[1, :7])
    np.testing.assert_array_equal(vv[L.VarKey.ROBOT].q[0], np.zeros(7))
    one = L.VarValues.from_var(L.VectorVar(np.array([[1.0, 2.0]])), L.VarKey.ACTION)
    np.testing.assert_array_equal(one.integrate(np.array([[0.5, -1.0]], np.float32))[L.VarKey.ACTION].x, [[1.5, 1.0]])
    capped = L.VectorVar(np.zeros((1, 2)), unit=10.0, max_step=1.0)
    np.testing.assert_array_equal(capped.integrate(np.array([[0.0625, 1.0]], np.float32)).x, [[0.625, 1.0]])
    with pytest.raises(ValueError):
        L.VarValues({L.VarKey.ROBOT: st, L.VarKey.ACTION: L.VectorVar(np.ones((2, 2)))})  # batch sizes differ
    with pytest.raises(ValueError):
        L.VarValues({})


@pytest.mark.gpu
@pytest.mark.parametrize("solver,bounds", [("cholesky", False), ("cg", False), ("cholesky", True)])
def test_trajectory_optimizer_equals_the_batch(ctx, solver, bounds):
    """LMOptimizer over the trajectory terms runs the trajectory kernel on the same problem set as
    the native batch of the workload: results bitwise equal, a second variable carried unchanged"""
    import dataclasses

    tb = W.config3(seed=5, batch=16, T=32)
    if bounds:
        na = 7
        lo = np.stack([-np.full(na, 1.5), -np.full(na, 0.2), -np.full(na, 0.1)]).astype(np.float32)
        tb = dataclasses.replace(tb, lim_orders=7, lim_w=np.full((3, na), 5.0, np.float32), lim_lo=lo, lim_hi=-lo)
    cfg = L.LMOptimizerConfig(max_iter=20, early_stopping_interval=0, lm_lambda=1e-2, lambda_factor=10.0,
                              linear_solver=solver)
    if solver == "cg":
        tb = dataclasses.replace(tb, cg_iters=tb.cg_default_iters)
    robot, terms = _traj_terms(tb)
    opt = cfg.build(terms)
    extra = L.VectorVar(np.full((tb.size, 3), 7.0))
    var = L.VarValues({L.VarKey.ROBOT: robot.state(tb.q_seed), L.VarKey.TIMING: extra})
    var, costs = opt.solve(var)
    b = N.TrajBatch(ctx, tb, tb.size)
    b.set(q=tb.q_seed, goals=tb.goals)
    b.solve(20, cfg.params())
    ref = b.get()
    np.testing.assert_array_equal(var[L.VarKey.ROBOT].q, ref["q"])
    np.testing.assert_array_equal(costs, ref["cost"])
    np.testing.assert_array_equal(var[L.VarKey.TIMING].x, np.full((tb.size, 3), 7.0, np.float32))
    assert np.all(np.isfinite(costs))


@pytest.mark.gpu
def test_extra_variable_leaves_the_ik_solve_unchanged(ctx):
    """a VectorVar no term reads has zero Jacobian columns: the solve of the robot key is the
    ROBOT-only solve, bit for bit, and the vector is unchanged"""
    robot, wl, q0 = _ik_problem(B=32, seed=4)
    terms = [L.FrameTask(robot, wl.links[0], wl.targets), L.PositionLimit(robot)]
    cfg = L.LMOptimizerConfig(max_iter=15, early_stopping_interval=0, lm_lambda=1e-2, lambda_factor=10.0)
    a = L.VarValues({L.VarKey.ROBOT: robot.state(q0)})
    b = L.VarValues({L.VarKey.ROBOT: robot.state(q0), L.VarKey.ACTION: L.VectorVar(np.ones((32, 4)))})
    qa, ca = cfg.build(terms).solve(a)
    qb, cb = cfg.build(terms).solve(b)
    np.testing.assert_array_equal(qa[L.VarKey.ROBOT].q, qb[L.VarKey.ROBOT].q)
    np.testing.assert_array_equal(ca, cb)
    np.testing.assert_array_equal(qb[L.VarKey.ACTION].x, np.ones((32, 4), np.float32))

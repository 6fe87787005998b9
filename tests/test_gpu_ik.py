"""GPU parity of the IKSolver front end (robokit_amd.ik) against the oracle population.

Bars: the winners classify like the FP32 oracle population (no more disagreements than FP32 vs
FP64); compute_metrics_numpy equals the oracle's metrics at the reference's default thresholds;
the previous state is kept exactly where its cost is no worse than the winner's; candidates
around an initial state put candidate 0 on it.
"""
import numpy as np
import pytest

from oracle import oracle as O
from oracle import population as OP
from robokit_amd import ik as IK
from robokit_amd import population as P
from robokit_amd import workloads as W

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def panda():
    spec = W.panda_dh()
    return spec, W.ik_batch(spec, [spec.num_links - 1], 32, seed=41)


def _stages(cfg):
    return [(s.iters, w, O.LMParams(**p.__dict__)) for s, w, p in
            zip(cfg.stages, cfg.num_candidates, cfg.optimizer)]


def test_solver_matches_oracle_population(ctx, panda):
    spec, wl = panda
    solver = IK.IKSolver(IK.IKSolverConfig(), spec, wl.links, ctx=ctx)
    print(f"\n[ik] panda {wl.size} targets, default schedule", flush=True)
    q = solver.solve_numpy(wl.targets)
    q32, c32, _ = OP.population_ik(spec, wl.links, wl.targets, solver._pop.seeds, _stages(solver.config.solver))
    q64, _, _ = OP.population_ik(spec, wl.links, wl.targets, solver._pop.seeds, _stages(solver.config.solver),
                                 dtype="f64")
    ok = O.ik_metrics(spec, q, wl.links, wl.targets)[2]
    ok32 = O.ik_metrics(spec, q32, wl.links, wl.targets)[2]
    ok64 = O.ik_metrics(spec, q64, wl.links, wl.targets, dtype="f64")[2]
    mism, n_ref = int(np.sum(ok != ok32)), int(np.sum(ok64 != ok32))
    print(f"  converged gpu {ok.mean():.3f} f32 {ok32.mean():.3f}; differ gpu/f32 {mism}, f32/f64 {n_ref}")
    assert mism <= max(n_ref, 1)
    # metrics at the reference's defaults (5 mm, 0.05 rad)
    pe, ae, succ = solver.compute_metrics_numpy(q, wl.targets)
    pe_o, ae_o, succ_o = O.ik_metrics(spec, q, wl.links, wl.targets, pos_thr=0.005, ang_thr=0.05)
    np.testing.assert_allclose(pe, pe_o, rtol=1e-4, atol=1e-5)
    np.testing.assert_allclose(ae, ae_o, rtol=1e-4, atol=1e-5)
    assert np.sum(succ != succ_o) <= 1
    with pytest.raises(ValueError, match="mismatch"):
        solver.solve_numpy(wl.targets[:5])
    with pytest.raises(ValueError, match="shape"):
        solver.solve_numpy(np.zeros((32, 6), np.float32))


@pytest.mark.parametrize("rank_final", [False, True])
def test_several_solutions_per_target(ctx, panda, rank_final):
    """A final stage of 4 candidates (helpers/ik/ik.py:175-178, 233-304): solve_numpy returns 4
    solutions per target, instance-major, the final population of the oracle's flow
    (PopulationSolver.final_var, population_solver.py:384-404); with rank_final best first, the
    first being the winner of the default schedule (the same stages and a final select of one).
    The solutions classify like the FP32 oracle population's (per target, the converged counts
    differ by no more in total than FP32 vs FP64), and a previous state needs one solution."""
    spec, wl = panda
    from robokit_amd import native as N
    cfg = IK.IKSolverConfig(solver=P.PopulationSolverConfig(
        stages=[P.Stage(6, 64, N.LMParams(lm_lambda=10.0)), P.Stage(10, 4, N.LMParams(lm_lambda=1.0))],
        rank_final=rank_final))
    solver = IK.IKSolver(cfg, spec, wl.links, ctx=ctx)
    assert solver.num_solutions == 4
    print(f"\n[ik] panda {wl.size} targets x 4 solutions, rank_final={rank_final}", flush=True)
    q = solver.solve_numpy(wl.targets)
    assert q.shape == (wl.size * 4, 7)
    stages = _stages(cfg.solver)
    q32, c32, _ = OP.population_ik(spec, wl.links, wl.targets, solver._pop.seeds, stages, keep_final=True,
                                   rank_final=rank_final)
    q64, _, _ = OP.population_ik(spec, wl.links, wl.targets, solver._pop.seeds, stages, dtype="f64",
                                 keep_final=True, rank_final=rank_final)
    rep = lambda a: np.repeat(a, 4, axis=0)  # noqa: E731
    ok = O.ik_metrics(spec, q, wl.links, rep(wl.targets))[2].reshape(-1, 4).sum(1)
    ok32 = O.ik_metrics(spec, q32, wl.links, rep(wl.targets))[2].reshape(-1, 4).sum(1)
    ok64 = O.ik_metrics(spec, q64, wl.links, rep(wl.targets), dtype="f64")[2].reshape(-1, 4).sum(1)
    mism, n_ref = int(np.abs(ok - ok32).sum()), int(np.abs(ok64 - ok32).sum())
    print(f"  converged solutions gpu {ok.sum()} f32 {ok32.sum()} f64 {ok64.sum()}; differ gpu/f32 {mism}, "
          f"f32/f64 {n_ref}")
    assert mism <= max(n_ref, 1)
    if rank_final:
        assert np.all(np.diff(solver.last_cost.reshape(-1, 4), axis=1) >= 0)
        win = IK.IKSolver(IK.IKSolverConfig(), spec, wl.links, ctx=ctx).solve_numpy(wl.targets)
        np.testing.assert_array_equal(q.reshape(-1, 4, 7)[:, 0], win)
    with pytest.raises(ValueError, match="num_candidates=1"):
        solver.solve_numpy(wl.targets, prev_q=q[::4])


def test_prev_state_kept_on_ties(ctx, panda):
    """select_better_score_kernel (helpers/ik/ik.py:43-57): prev_q wins where its cost <= the
    winner's; the solve itself is reproducible, so the other rows are the plain solve's winners"""
    spec, wl = panda
    solver = IK.IKSolver(IK.IKSolverConfig(), spec, wl.links, ctx=ctx)
    q1 = solver.solve_numpy(wl.targets)
    q2 = solver.solve_numpy(wl.targets)
    np.testing.assert_array_equal(q1, q2)  # same seeds, same bits
    # a previous state that is the winner itself ties and is kept; a perturbed one loses
    prev = q1.copy()
    prev[::2] += 0.05
    prev = np.clip(prev, spec.actuated_joint_limits_f32[:, 0], spec.actuated_joint_limits_f32[:, 1])
    q3 = solver.solve_numpy(wl.targets, prev_q=prev)
    keep = solver.last_prev_cost <= solver.last_cost
    np.testing.assert_array_equal(q3[keep], prev[keep])
    np.testing.assert_array_equal(q3[~keep], q1[~keep])
    # the unperturbed rows score the winner itself: its cost again, to rounding
    np.testing.assert_allclose(solver.last_prev_cost[1::2], solver.last_cost[1::2], rtol=1e-4, atol=1e-9)
    # the previous state's score is the terms' cost at prev_q
    c_o = O.lm_solve(spec, O.ik_terms(wl.links, wl.targets, 7), prev, 0, O.LMParams(), traces=False)["cost"]
    np.testing.assert_allclose(solver.last_prev_cost, c_o, rtol=1e-4, atol=1e-7)


def test_candidates_around_an_initial_state(ctx, panda):
    """zero-iteration schedule: the select keeps the best seed, and candidate 0 sits on init_q, the
    exact joint solution, so it wins"""
    spec, wl = panda
    cfg = IK.IKSolverConfig(solver=P.PopulationSolverConfig([P.Stage(0, 64), P.Stage(0, 1)]))
    solver = IK.IKSolver(cfg, spec, wl.links, ctx=ctx)
    init = np.clip(wl.q_target, spec.actuated_joint_limits_f32[:, 0], spec.actuated_joint_limits_f32[:, 1])
    q = solver.solve_numpy(wl.targets, init_q=init, init_sample_range=0.1)
    np.testing.assert_array_equal(q, init)
    # full schedule from the initial states converges at least as often as from the mid-range
    s_mid = IK.IKSolver(IK.IKSolverConfig(), spec, wl.links, ctx=ctx)
    ok_mid = s_mid.compute_metrics_numpy(s_mid.solve_numpy(wl.targets), wl.targets)[2]
    noisy = np.clip(wl.q_target + np.random.default_rng(2).normal(0, 0.05, wl.q_target.shape).astype(np.float32),
                    spec.actuated_joint_limits_f32[:, 0], spec.actuated_joint_limits_f32[:, 1])
    ok_init = s_mid.compute_metrics_numpy(s_mid.solve_numpy(wl.targets, init_q=noisy), wl.targets)[2]
    assert ok_init.sum() >= ok_mid.sum()


def test_solver_humanoid_four_targets(ctx):
    """a tree with four pose targets (the stages run the half-wave register kernels), candidates
    around noisy initial states: the winners classify like the oracle population"""
    spec = W.humanoid29(0)
    links = W.humanoid_ee_links(spec)
    wl = W.ik_batch(spec, links, 12, seed=43)
    lim32 = spec.actuated_joint_limits_f32
    init = np.clip(wl.q_target + np.random.default_rng(4).normal(0, 0.15, wl.q_target.shape).astype(np.float32),
                   lim32[:, 0], lim32[:, 1])
    solver = IK.IKSolver(IK.IKSolverConfig(), spec, links, ctx=ctx)
    print(f"\n[ik] humanoid {wl.size} x {len(links)} targets", flush=True)
    q = solver.solve_numpy(wl.targets, init_q=init, init_sample_range=0.1)
    seeds = solver._pop.candidate_seeds(0.1, center=init)
    stages = _stages(solver.config.solver)
    q32, _, _ = OP.population_ik(spec, links, wl.targets, seeds, stages)
    q64, _, _ = OP.population_ik(spec, links, wl.targets, seeds, stages, dtype="f64")
    ok = O.ik_metrics(spec, q, links, wl.targets, pos_thr=0.005, ang_thr=0.05)[2]
    ok32 = O.ik_metrics(spec, q32, links, wl.targets, pos_thr=0.005, ang_thr=0.05)[2]
    ok64 = O.ik_metrics(spec, q64, links, wl.targets, pos_thr=0.005, ang_thr=0.05, dtype="f64")[2]
    mism, n_ref = int(np.sum(ok != ok32)), int(np.sum(ok64 != ok32))
    print(f"  converged (5 mm) gpu {ok.mean():.3f} f32 {ok32.mean():.3f}; differ gpu/f32 {mism}, f32/f64 {n_ref}")
    assert mism <= max(n_ref, 1)
    assert ok32.sum() > 0  # the comparison is not vacuous
    assert np.sum(solver.compute_metrics_numpy(q, wl.targets)[2] != ok) <= 1


def _preset_terms(links, targets, na):
    """the oracle terms of presets.basic: PositionTask w 20 + RotationTask w 10 on every target link,
    PositionLimit w 50 (helpers/ik/presets.py:13-18)"""
    k = len(links)
    return [O.Term(O.POSITION, links, targets, np.full(3 * k, 20.0, np.float32)),
            O.Term(O.ROTATION, links, targets, np.full(3 * k, 10.0, np.float32)),
            O.Term(O.LIMIT, (), None, np.full(na, 50.0, np.float32))]


def test_presets_basic_matches_oracle_population(ctx):
    """IKSolver(presets.basic) on the Panda with its TCP link (the reference benchmark's robot and
    target link): 64 Roberts candidates (keep_first, scale 1) x 6 iterations at lambda 10, best 4 x 10
    at lambda 1, best 1; the stages run the one-lane-per-problem register kernel's position +
    rotation rows.  The winners classify like the FP32 oracle population at the reference
    benchmark's 5 mm / 0.05 rad, and the winners' costs agree with the oracle's to its FP32 spread."""
    spec = W.panda_tcp()
    tcp = spec.link_names.index("panda_hand_tcp")
    wl = W.ik_batch(spec, [tcp], 96, seed=45)
    solver = IK.IKSolver(IK.presets.basic, spec, tcp, ctx=ctx)
    print(f"\n[ik] presets.basic panda_tcp {wl.size} targets", flush=True)
    q = solver.solve_numpy(wl.targets)
    stages = _stages(solver.config.solver)
    fn = lambda inst: _preset_terms([tcp], wl.targets[inst], 7)  # noqa: E731
    q32, c32, _ = OP.population_ik(spec, [tcp], wl.targets, solver._pop.seeds, stages, terms_fn=fn, threads=8)
    q64, c64, _ = OP.population_ik(spec, [tcp], wl.targets, solver._pop.seeds, stages, dtype="f64", terms_fn=fn,
                                   threads=8)
    thr = dict(pos_thr=0.005, ang_thr=0.05)
    ok = O.ik_metrics(spec, q, [tcp], wl.targets, **thr)[2]
    ok32 = O.ik_metrics(spec, q32, [tcp], wl.targets, **thr)[2]
    ok64 = O.ik_metrics(spec, q64, [tcp], wl.targets, dtype="f64", **thr)[2]
    mism, n_ref = int(np.sum(ok != ok32)), int(np.sum(ok64 != ok32))
    print(f"  converged (5 mm) gpu {ok.mean():.3f} f32 {ok32.mean():.3f} f64 {ok64.mean():.3f}; "
          f"differ gpu/f32 {mism}, f32/f64 {n_ref}")
    assert mism <= max(n_ref, 1)
    assert ok32.sum() > 0.5 * wl.size  # the preset solves most targets: the comparison is not vacuous
    # compute_metrics_numpy of the preset solver (its own pose-target metrics batch) = the oracle's
    pe, ae, succ = solver.compute_metrics_numpy(q, wl.targets)
    pe_o, ae_o, _ = O.ik_metrics(spec, q, [tcp], wl.targets, **thr)
    np.testing.assert_allclose(pe, pe_o, rtol=1e-4, atol=1e-5)
    np.testing.assert_allclose(ae, ae_o, rtol=1e-4, atol=1e-5)
    # a previous state is scored with the preset's terms: the oracle's cost at prev_q
    prev = np.clip(q + 0.01, spec.actuated_joint_limits_f32[:, 0], spec.actuated_joint_limits_f32[:, 1])
    solver.solve_numpy(wl.targets, prev_q=prev)
    c_o = O.lm_solve(spec, _preset_terms([tcp], wl.targets, 7), prev, 0, O.LMParams(), traces=False)["cost"]
    np.testing.assert_allclose(solver.last_prev_cost, c_o, rtol=1e-4, atol=1e-6)


def test_presets_smooth_needs_and_uses_previous_state(ctx):
    """presets.smooth adds SmoothnessTask(prev_var=PREVIOUS_STATE, w 0.1): a solve without prev_q is
    refused; with prev_q the winners' costs include the smoothness rows (the oracle's cost of the
    returned joints with prev_q as reference agrees with the solver's)"""
    spec = W.panda_tcp()
    tcp = spec.link_names.index("panda_hand_tcp")
    wl = W.ik_batch(spec, [tcp], 16, seed=46)
    solver = IK.IKSolver(IK.presets.smooth, spec, tcp, ctx=ctx)
    with pytest.raises(ValueError, match="prev_q"):
        solver.solve_numpy(wl.targets)
    prev = wl.q_target.copy()
    q = solver.solve_numpy(wl.targets, prev_q=prev)
    terms = _preset_terms([tcp], wl.targets, 7) + [O.Term(O.SMOOTH, (), None, np.full(7, 0.1, np.float32), prev)]
    c_o = O.lm_solve(spec, terms, q, 0, O.LMParams(), traces=False)["cost"]
    best = np.minimum(solver.last_cost, solver.last_prev_cost)
    np.testing.assert_allclose(best, c_o, rtol=1e-3, atol=1e-6)


def test_solver_scene_and_rest_state_match_oracle_population(ctx):
    """IKSolver with the basic preset's terms plus a SceneCollisionTask on IKContext.SCENE (primitive
    obstacles and a box mesh) and a SmoothnessTask pinning the rest configuration IKContext.REST_Q
    (solve_numpy(rest_q=...)): the winners classify like the FP32 oracle population with the same
    terms; the rest reference defaults to the zero configuration"""
    from robokit_amd import geom as G
    from robokit_amd import lm as L

    tb = W.config3(seed=12, batch=4, T=5)
    spec, sc = tb.spec, tb.scene
    ee = spec.num_links - 1
    wl = W.ik_batch(spec, [ee], 64, seed=47)
    mesh = G.box_mesh([0.45, 0.0, 0.45], [0.08, 0.25, 0.05])
    scene = L.PrimitiveScene(sc.obstacle_spheres, sc.obstacle_boxes, geoms=[mesh])
    robot = L.Robot(spec, collision_spheres=(sc.sphere_links, sc.sphere_centers, sc.sphere_radii))
    cfg = (IK.IKSolverConfig(sampler=IK.RobertsSamplerConfig(keep_first=True, scale=1.0))
           .add(L.PositionTask(frame_index=IK.IKContext.TARGET_LINKS, T_world_target=IK.IKContext.TARGET, weight=20.0))
           .add(L.RotationTask(frame_index=IK.IKContext.TARGET_LINKS, T_world_target=IK.IKContext.TARGET, weight=10.0))
           .add(L.PositionLimit(weight=50.0))
           .add(L.SmoothnessTask(prev_var=IK.IKContext.REST_Q, weight=0.05))
           .add(L.SceneCollisionTask(scene=IK.IKContext.SCENE, margin=0.05)))
    with pytest.raises(ValueError):
        IK.IKSolver(cfg, robot, ee, ctx=ctx)  # IKContext.SCENE without a scene
    solver = IK.IKSolver(cfg, robot, ee, ctx=ctx, scene=scene)
    rest = np.clip(wl.q_target + 0.3, spec.actuated_joint_limits_f32[:, 0], spec.actuated_joint_limits_f32[:, 1])
    print(f"\n[ik] scene + rest state, panda {wl.size} targets", flush=True)
    q = solver.solve_numpy(wl.targets, rest_q=rest)
    na = spec.num_actuated_joints
    coll = O.CollisionScene(sc.sphere_links, sc.sphere_centers, sc.sphere_radii, np.ones(len(sc.sphere_links)),
                            margin=0.05, penalty="smooth", obstacle_spheres=sc.obstacle_spheres,
                            obstacle_boxes=sc.obstacle_boxes, geoms=[mesh])

    def fn(inst):
        return _preset_terms([ee], wl.targets[inst], na) + [
            O.Term(O.SMOOTH, (), None, np.full(na, 0.05, np.float32), rest[inst]),
            O.Term(O.COLLISION, (), None, None, None, coll)]

    stages = _stages(solver.config.solver)
    q32, _, _ = OP.population_ik(spec, [ee], wl.targets, solver._pop.seeds, stages, terms_fn=fn, threads=8)
    q64, _, _ = OP.population_ik(spec, [ee], wl.targets, solver._pop.seeds, stages, dtype="f64", terms_fn=fn,
                                 threads=8)
    thr = dict(pos_thr=0.005, ang_thr=0.05)
    ok = O.ik_metrics(spec, q, [ee], wl.targets, **thr)[2]
    ok32 = O.ik_metrics(spec, q32, [ee], wl.targets, **thr)[2]
    ok64 = O.ik_metrics(spec, q64, [ee], wl.targets, dtype="f64", **thr)[2]
    mism, n_ref = int(np.sum(ok != ok32)), int(np.sum(ok64 != ok32))
    print(f"  converged (5 mm) gpu {ok.mean():.3f} f32 {ok32.mean():.3f} f64 {ok64.mean():.3f}; "
          f"differ gpu/f32 {mism}, f32/f64 {n_ref}")
    assert mism <= max(n_ref, 2)
    assert ok32.sum() > 0.3 * wl.size
    # the default rest reference is the zero configuration
    q0 = solver.solve_numpy(wl.targets)
    q0b = solver.solve_numpy(wl.targets, rest_q=np.zeros(na, np.float32))
    np.testing.assert_array_equal(q0, q0b)


def test_score_terms_rank_the_candidates(ctx):
    """IKSolverConfig.add_score: the basic preset optimises, a FrameTask (orientation weighted 0.1)
    ranks the candidates at every select and the previous state; the winners classify like the
    oracle population ranked by the same score terms, and the reported costs are the scores"""
    from robokit_amd import lm as L

    spec = W.panda_tcp()
    tcp = spec.link_names.index("panda_hand_tcp")
    wl = W.ik_batch(spec, [tcp], 64, seed=53)
    cfg = IK.IKSolverConfig(sampler=IK.RobertsSamplerConfig(keep_first=True, scale=1.0))
    for t in IK.presets.basic.terms:
        cfg.add(t)
    cfg.add_score(L.FrameTask(frame_index=IK.IKContext.TARGET_LINKS, T_world_target=IK.IKContext.TARGET,
                              position_weight=1.0, orientation_weight=0.1), name="score")
    solver = IK.IKSolver(cfg, spec, tcp, ctx=ctx)
    q = solver.solve_numpy(wl.targets)
    stages = _stages(solver.config.solver)
    fn = lambda inst: _preset_terms([tcp], wl.targets[inst], 7)  # noqa: E731
    w6 = np.array([1, 1, 1, 0.1, 0.1, 0.1], np.float32)
    sfn = lambda inst: [O.Term(O.FRAME, [tcp], wl.targets[inst], w6)]  # noqa: E731
    q32, c32, _ = OP.population_ik(spec, [tcp], wl.targets, solver._pop.seeds, stages, terms_fn=fn, threads=8,
                                   score_fn=sfn)
    q64, _, _ = OP.population_ik(spec, [tcp], wl.targets, solver._pop.seeds, stages, dtype="f64", terms_fn=fn,
                                 threads=8, score_fn=sfn)
    thr = dict(pos_thr=0.005, ang_thr=0.05)
    ok = O.ik_metrics(spec, q, [tcp], wl.targets, **thr)[2]
    ok32 = O.ik_metrics(spec, q32, [tcp], wl.targets, **thr)[2]
    ok64 = O.ik_metrics(spec, q64, [tcp], wl.targets, dtype="f64", **thr)[2]
    mism, n_ref = int(np.sum(ok != ok32)), int(np.sum(ok64 != ok32))
    print(f"\n[ik] score terms: converged gpu {ok.mean():.3f} f32 {ok32.mean():.3f}; differ gpu/f32 {mism}, "
          f"f32/f64 {n_ref}")
    assert mism <= max(n_ref, 1)
    # the winners' reported costs are their scores (the oracle's score at the GPU's winners)
    s_o = O.lm_solve(spec, sfn(np.arange(wl.size)), q, 0, O.LMParams(), traces=False)["cost"]
    np.testing.assert_allclose(solver.last_cost, s_o, rtol=1e-4, atol=1e-7)
    # the previous state is ranked by the score terms too
    solver.solve_numpy(wl.targets, prev_q=q)
    np.testing.assert_allclose(solver.last_prev_cost, s_o, rtol=1e-4, atol=1e-7)


def test_solve_torch_adapter(ctx, panda):
    """solve_torch takes 4 x 4 target matrices as torch tensors and returns torch q equal to
    solve_numpy's on the same poses"""
    import torch

    spec, wl = panda
    solver = IK.IKSolver(IK.IKSolverConfig(), spec, wl.links, ctx=ctx)
    x, y, z = wl.targets[..., 4], wl.targets[..., 5], wl.targets[..., 6]
    w = wl.targets[..., 3]
    T = np.zeros(wl.targets.shape[:-1] + (4, 4))
    T[..., 0, :3] = np.stack([1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)], -1)
    T[..., 1, :3] = np.stack([2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)], -1)
    T[..., 2, :3] = np.stack([2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)], -1)
    T[..., :3, 3] = wl.targets[..., :3]
    T[..., 3, 3] = 1
    res = solver.solve_torch(torch.from_numpy(T.astype(np.float32)))
    assert torch.is_tensor(res.q) and res.T_world_base is None
    poses = IK.matrix_to_pose(T.astype(np.float32))
    q_np = solver.solve_numpy(poses)
    np.testing.assert_array_equal(res.q.numpy(), q_np)

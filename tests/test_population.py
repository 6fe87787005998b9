"""Population IK (CPU): the oracle's R_d sequence and best-k ranking pinned against closed forms,
the reference's own test cases and the reference's plain selection loop; the host-side schedule
and seeding of robokit_amd.population checked against them."""
import math

import numpy as np
import pytest

from oracle import population as OP
from robokit_amd import native as N
from robokit_amd import population as P
from robokit_amd import workloads as W


# ----------------------------------------------------------------------------- R_d sequence
def test_roberts_closed_forms():
    """d = 1: x^2 = x + 1 has the golden ratio as its root, so the step is 1 - 1/phi = 2 - phi;
    d = 2: x^3 = x + 1 has the plastic number (Cardano's closed form)."""
    phi = (1 + math.sqrt(5)) / 2
    s1 = OP.roberts_samples(4, 1)
    np.testing.assert_allclose(s1[:, 0], [(n * (2 - phi)) % 1.0 for n in range(4)], rtol=0, atol=1e-12)
    rho = (((9 + math.sqrt(69)) / 18) ** (1 / 3) + ((9 - math.sqrt(69)) / 18) ** (1 / 3))
    assert abs(rho ** 3 - rho - 1) < 1e-12
    s2 = OP.roberts_samples(5, 2, offset=3)
    alpha = np.array([1 - 1 / rho, 1 - 1 / rho ** 2])
    np.testing.assert_allclose(s2, np.modf(np.outer(np.arange(3, 8), alpha))[0], rtol=0, atol=1e-12)


def test_roberts_reference_properties():
    """tests/opt/test_sampler.py:57-61 and the docstring examples of opt/sampler.py:249-254"""
    assert OP.roberts_samples(5, 3).shape == (5, 3)
    assert OP.roberts_samples(64, 2).max() < 1.0
    cells = lambda pts: len({(int(x * 8), int(y * 8)) for x, y in pts})  # noqa: E731
    assert cells(OP.roberts_samples(64, 2)) > cells(np.random.default_rng(0).random((64, 2)))


def test_product_roberts_and_offsets_match_oracle():
    for dim in (1, 2, 6, 7, 29):
        np.testing.assert_array_equal(P.roberts_samples(64, dim), OP.roberts_samples(64, dim))
    off = P.candidate_offsets(64, 7)
    assert off.dtype == np.float32 and off.shape == (64, 7)
    np.testing.assert_array_equal(off[0], 0.0)  # keep_first: candidate 0 on the centre
    np.testing.assert_array_equal(off[1], -0.5)  # then point 0 of the sequence (0 - 1/2)
    np.testing.assert_array_equal(off[2:], (OP.roberts_samples(63, 7)[1:] - 0.5).astype(np.float32))
    off_s = P.candidate_offsets(8, 3, seed=10017, keep_first=False)  # seed % 10000 offsets the sequence
    np.testing.assert_array_equal(off_s, (OP.roberts_samples(8, 3, offset=17) - 0.5).astype(np.float32))


# ----------------------------------------------------------------------------- best-k ranking
def _reference_loop(costs, width, k):
    """the plain selection loop of opt/population_solver.py:24-44, problem by problem"""
    c = np.asarray(costs, np.float32).reshape(-1, width)
    out = np.zeros((c.shape[0], k), np.int64)
    for b in range(c.shape[0]):
        prev_c, prev_j = -np.inf, -1
        for i in range(k):
            best_c, best_j = np.inf, -1
            for j in range(width):
                v = np.inf if np.isnan(c[b, j]) else c[b, j]
                after = v > prev_c or (v == prev_c and j > prev_j)
                if after and (best_j < 0 or v < best_c or (v == best_c and j < best_j)):
                    best_c, best_j = v, j
            out[b, i] = b * width + best_j
            prev_c, prev_j = best_c, best_j
    return out


def test_select_golden_case():
    """tests/helpers/motion_plan/test_motion_plan.py:2871-2889: infinite costs tie in candidate
    order after the finite ones; equal finite costs keep candidate order"""
    keys = np.array([1e-3, 0, np.inf, np.inf, 0.25, 0.25, 0.25, 0.25, 0, np.inf, 1e-3, np.inf], np.float32)
    costs = 0.5 * keys * keys
    assert OP.select_best_k(costs, 4, 4).tolist() == [[1, 0, 2, 3], [4, 5, 6, 7], [8, 10, 9, 11]]


@pytest.mark.parametrize("width,k", [(64, 4), (4, 1), (64, 64), (7, 3), (1, 1)])
def test_select_matches_reference_loop(width, k):
    rng = np.random.default_rng(width * 100 + k)
    c = rng.choice(np.array([0.0, 0.5, 1.0, 2.0, np.inf, np.nan, 1e-7], np.float32), size=(9, width))
    c[::2] = rng.random((5, width)).astype(np.float32)  # some rows all distinct
    np.testing.assert_array_equal(OP.select_best_k(c.reshape(-1), width, k), _reference_loop(c, width, k))


# ----------------------------------------------------------------------------- schedule and seeds
def test_default_ik_schedule():
    """helpers/ik/config.py:37-45 and the PopulationSolverConfig properties
    (opt/population_solver.py:94-121)"""
    cfg = P.default_ik_config()
    assert cfg.num_iters == 16 and cfg.select_at == [6, 16] and cfg.num_candidates == [64, 4, 1]
    lams = [o.lm_lambda for o in cfg.optimizer]
    assert lams == [10.0, 1.0, 1.0]  # an unset stage keeps the previous stage's optimizer
    assert cfg.optimizer[0].lambda_factor == 2.0  # LMOptimizerConfig defaults otherwise
    assert P.PopulationSolverConfig([P.Stage(3)]).num_candidates == [1]
    with pytest.raises(ValueError, match="selects only narrow"):
        P.PopulationSolverConfig([P.Stage(2, 4), P.Stage(2, 8)])
    with pytest.raises(ValueError, match="non-negative"):
        P.PopulationSolverConfig([P.Stage(-1, 4)])
    with pytest.raises(ValueError, match="num_candidates"):
        P.PopulationSolverConfig([P.Stage(2, 65)])


class _NoDevice(P.PopulationIK):
    """the seeding of PopulationIK without its device batches"""

    def __init__(self, spec, config, seed=None, keep_first=True):
        self.spec, self.config, self.widths = spec, config, config.num_candidates
        self.offsets = P.candidate_offsets(self.widths[0], spec.num_actuated_joints, seed, keep_first)


@pytest.mark.parametrize("robot", ["panda_dh", "ur10", "humanoid"])
def test_candidate_seeds(robot):
    """helpers/ik/ik.py:509-532: mid-range + offset x (float32 upper - lower), float32 arithmetic,
    clamped to the float32 limits; candidate 0 is the mid-range"""
    spec = {"panda_dh": W.panda_dh, "ur10": W.ur10, "humanoid": lambda: W.humanoid29(0)}[robot]()
    seeds = _NoDevice(spec, P.default_ik_config()).candidate_seeds(1.0)
    lim = np.asarray(spec.actuated_joint_limits)
    lim32 = spec.actuated_joint_limits_f32
    assert seeds.dtype == np.float32 and seeds.shape == (64, spec.num_actuated_joints)
    np.testing.assert_array_equal(seeds[0], np.clip((0.5 * (lim[:, 0] + lim[:, 1])).astype(np.float32),
                                                    lim32[:, 0], lim32[:, 1]))
    assert np.all(seeds >= lim32[:, 0]) and np.all(seeds <= lim32[:, 1])
    # candidate 1 sits at offset -1/2: the lower limit, up to float32 rounding of mid - range/2
    np.testing.assert_allclose(seeds[1], lim32[:, 0], atol=1e-6)
    # the spread covers the joint box: every joint sees both halves of its range
    mid = 0.5 * (lim[:, 0] + lim[:, 1])
    assert np.all((seeds < mid).sum(0) > 16) and np.all((seeds > mid).sum(0) > 16)


def test_population_oracle_select_consistency():
    """the oracle's population flow on a tiny Panda case: every select keeps rows of the problem
    they belong to, and the winner's cost is the lowest of its final candidates"""
    spec = W.panda_dh()
    wl = W.ik_batch(spec, [spec.num_links - 1], 3, seed=5)
    seeds = _NoDevice(spec, P.default_ik_config()).candidate_seeds(1.0)[:8]
    stages = [(3, 8, OP.O.LMParams(lm_lambda=10.0)), (4, 2, OP.O.LMParams(lm_lambda=1.0)), (0, 1, None)]
    q, cost, sel = OP.population_ik(spec, wl.links, wl.targets, seeds, stages)
    assert q.shape == (3, 7) and cost.shape == (3,)
    assert [s.tolist() for s in sel[0].reshape(3, 2) // 8] == [[0, 0], [1, 1], [2, 2]]
    assert (sel[1] // 2).tolist() == [0, 1, 2]
    assert np.all(np.isfinite(cost))


def test_population_oracle_final_population():
    """keep_final returns the final stage's population (rows i * width + j of problem i); with
    rank_final it is ordered best first, its first row is the winner of the plain flow, and the
    unranked population holds the same rows"""
    spec = W.panda_dh()
    wl = W.ik_batch(spec, [spec.num_links - 1], 3, seed=5)
    seeds = _NoDevice(spec, P.default_ik_config()).candidate_seeds(1.0)[:8]
    stages = [(3, 8, OP.O.LMParams(lm_lambda=10.0)), (4, 4, OP.O.LMParams(lm_lambda=1.0))]
    q1, c1, _ = OP.population_ik(spec, wl.links, wl.targets, seeds, stages)
    qa, ca, _ = OP.population_ik(spec, wl.links, wl.targets, seeds, stages, keep_final=True)
    qr, cr, sel = OP.population_ik(spec, wl.links, wl.targets, seeds, stages, keep_final=True, rank_final=True)
    assert qa.shape == (12, 7) and qr.shape == (12, 7) and cr.shape == (12,)
    assert np.all(np.diff(cr.reshape(3, 4), axis=1) >= 0)
    np.testing.assert_array_equal(qr.reshape(3, 4, 7)[:, 0], q1)
    np.testing.assert_array_equal(cr.reshape(3, 4)[:, 0], c1)
    assert [sorted(r) for r in (sel[-1].reshape(3, 4) // 4).tolist()] == [[0] * 4, [1] * 4, [2] * 4]
    np.testing.assert_array_equal(qr, qa[sel[-1]])


def test_select_abi_declared():
    assert "rk_batch_select" in N.EXPORTS


def test_candidate_seeds_around_an_initial_state():
    """helpers/ik/ik.py:513-532: with init_state, problem i's candidates are init_q[i] + offset x
    (scale x joint range), clamped; candidate 0 stays on init_q (keep_first)"""
    spec = W.panda_dh()
    nd = _NoDevice(spec, P.default_ik_config())
    lim32 = spec.actuated_joint_limits_f32
    rng = np.random.default_rng(3)
    init = (lim32[:, 0] + (lim32[:, 1] - lim32[:, 0]) * rng.random((5, 7))).astype(np.float32)
    s = nd.candidate_seeds(0.2, center=init)
    assert s.shape == (5 * 64, 7)
    np.testing.assert_array_equal(s[::64], np.clip(init, lim32[:, 0], lim32[:, 1]))
    lim = np.asarray(spec.actuated_joint_limits)
    half = 0.1 * (lim[:, 1] - lim[:, 0])
    d = s.reshape(5, 64, 7) - init[:, None, :]
    assert np.all(np.abs(d) <= half + 1e-6)
    assert np.all(s >= lim32[:, 0]) and np.all(s <= lim32[:, 1])


def test_ik_solver_rejects_bad_arguments_before_touching_the_device():
    from robokit_amd import ik as IK
    spec = W.panda_dh()
    with pytest.raises(ValueError, match="out of range"):
        IK.IKSolver(IK.IKSolverConfig(), spec, [spec.num_links + 3])
    with pytest.raises(ValueError, match="at least one"):
        IK.IKSolver(IK.IKSolverConfig(), spec, [])
    with pytest.raises(NotImplementedError):
        IK.IKSolverConfig(enable_T_world_base=True)


def test_matrix_to_pose_round_trips_rotations():
    """the torch adapter's target conversion: Shepperd's quaternion of a rotation matrix, compared
    with the matrix rebuilt from it (every branch: rotations near 0 and pi about each axis and
    random ones), the scalar part non-negative"""
    from robokit_amd.ik import matrix_to_pose

    rng = np.random.default_rng(0)

    def rot(axis, ang):
        a = np.asarray(axis, float) / np.linalg.norm(axis)
        K = np.array([[0, -a[2], a[1]], [a[2], 0, -a[0]], [-a[1], a[0], 0]])
        return np.eye(3) + np.sin(ang) * K + (1 - np.cos(ang)) * K @ K

    Rs = [rot(ax, ang) for ax in np.eye(3) for ang in (0.0, 0.3, np.pi - 1e-3, np.pi)]
    Rs += [rot(rng.normal(size=3), rng.uniform(0, np.pi)) for _ in range(50)]
    T = np.zeros((len(Rs), 4, 4))
    T[:, :3, :3] = Rs
    T[:, :3, 3] = rng.normal(size=(len(Rs), 3))
    T[:, 3, 3] = 1
    p = matrix_to_pose(T).astype(np.float64)
    w, x, y, z = p[:, 3], p[:, 4], p[:, 5], p[:, 6]
    Rq = np.stack([np.stack([1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)], -1),
                   np.stack([2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)], -1),
                   np.stack([2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)], -1)], 1)
    np.testing.assert_allclose(Rq, np.asarray(Rs), atol=2e-6)
    np.testing.assert_allclose(p[:, :3], T[:, :3, 3], atol=1e-6)
    assert np.all(w >= 0)

"""CPU restatement of the reference's population IK path (numpy over the C LM oracle).

TEST INFRASTRUCTURE ONLY: imported by tests/ and bench.py's baseline leg to check results; the
product package (robokit_amd) never imports it.

  roberts_samples      opt/sampler.py:246-264
  select_best_k        opt/population_solver.py:24-44 (_select_best_k_kernel: cost ascending,
                       NaN as +inf, equal costs in candidate order), restated as a stable sort
  population_ik        opt/population_solver.py:291-395 for the IK schedule of
                       helpers/ik/config.py:37-45: each segment a fresh LM (oracle.lm_solve) on
                       the kept candidates, then a select where the width narrows (an equal
                       width keeps the population, :436-443); the last select keeps one winner
                       (keep_final: the final population, best first with rank_final, :384-404)

Pins (tests/test_population.py): the R_d closed forms (golden ratio for d = 1, plastic number for
d = 2), the reference test's coverage property (tests/opt/test_sampler.py:57-61), the ranking
golden case of tests/helpers/motion_plan/test_motion_plan.py:2871-2889, and the reference's
plain selection loop on random costs with ties, infinities and NaNs.
"""
from __future__ import annotations

from typing import List, Sequence, Tuple

import numpy as np

from . import oracle as O


def roberts_samples(num_points: int, dim: int, offset: int = 0) -> np.ndarray:
    """frac(n * (1 - phi^-(k+1))), phi the root > 1 of x^(d+1) = x + 1 (Newton from 1)"""
    x = 1.0
    for _ in range(10000):
        f = x ** (dim + 1) - x - 1.0
        x = x - f / ((dim + 1) * x ** dim - 1.0)
        if abs(f) < 1e-10:
            break
    alpha = 1.0 - 1.0 / x ** (1.0 + np.arange(dim))
    n = (np.arange(num_points) + offset).astype(np.float64)
    return np.modf(np.outer(n, alpha))[0]


def select_best_k(costs: np.ndarray, width_in: int, width_out: int) -> np.ndarray:
    """[n_inst, width_out] flat source rows, best first (stable sort on the NaN-as-inf cost)"""
    c = np.asarray(costs).reshape(-1, width_in)
    key = np.where(np.isnan(c), np.inf, c)
    order = np.argsort(key, axis=1, kind="stable")[:, :width_out]
    return order + (np.arange(c.shape[0]) * width_in)[:, None]


def population_ik(spec, links: Sequence[int], targets: np.ndarray, seeds: np.ndarray,
                  stages: List[Tuple[int, int, O.LMParams]], dtype: str = "f32", terms_fn=None, threads: int = 1,
                  score_fn=None, keep_final: bool = False, rank_final: bool = False):
    """targets [n_inst, n_pose, 7]; seeds [width0, na] shared by every problem, or
    [n_inst * width0, na] per problem; stages [(iters, width, params)].  terms_fn(inst) gives the
    oracle terms of the candidate rows whose problems are ``inst`` (default: FrameTask on every
    link + PositionLimit, oracle.ik_terms).  score_fn(inst): score terms whose cost at the
    candidates ranks every select instead of the optimization cost (population_solver.py:415-434);
    the returned costs are then the winners' scores.  Returns (q [n_inst, na], cost [n_inst], rows of
    every select).  keep_final: return the final stage's population instead of its winners
    (PopulationSolver.final_var, population_solver.py:384-385), q [n_inst * width, na] in rows
    i * width + j; rank_final: rank only that population by score_fn (earlier selects keep the
    optimization cost, population_solver.py:185-186) and order it best first (ranked_indices,
    :232-236, 392-393)."""
    if terms_fn is None:
        terms_fn = lambda inst: O.ik_terms(links, targets[inst], seeds.shape[1])  # noqa: E731
    n_inst, na = targets.shape[0], seeds.shape[1]
    w = stages[0][1]
    inst = np.repeat(np.arange(n_inst), w)
    rt = np.float64 if dtype == "f64" else np.float32
    q = (np.tile(seeds, (n_inst, 1)) if seeds.shape[0] == w else np.asarray(seeds)).astype(rt)
    cost = None
    selects = []
    def rank(q, cost, inst, final=True):
        if score_fn is None or (rank_final and not final):
            return cost
        return O.lm_solve(spec, score_fn(inst), q, 0, O.LMParams(), dtype=dtype, traces=False,
                          threads=threads)["cost"].astype(rt)

    for k, (iters, width, params) in enumerate(stages):
        if k > 0 and width < w:  # an equal width keeps the population (population_solver.py:442-443)
            sc = rank(q, cost, inst, final=False)
            rows = select_best_k(sc, w, width).reshape(-1)
            selects.append(rows)
            q, cost, inst, w = q[rows], sc[rows], inst[rows], width
        if iters > 0 or k == 0:
            res = O.lm_solve(spec, terms_fn(inst), q, iters, params, dtype=dtype, traces=False, threads=threads)
            q, cost = res["q"].astype(rt), res["cost"].astype(rt)
    if keep_final:
        if rank_final and w > 1:
            sc = rank(q, cost, inst)
            rows = select_best_k(sc, w, w).reshape(-1)
            selects.append(rows)
            q, cost, inst = q[rows], sc[rows], inst[rows]
        elif score_fn is not None:
            cost = rank(q, cost, inst)
        return q, cost, selects
    if w > 1:
        sc = rank(q, cost, inst)
        rows = select_best_k(sc, w, 1).reshape(-1)
        selects.append(rows)
        q, cost, inst = q[rows], sc[rows], inst[rows]
    elif score_fn is not None:
        cost = rank(q, cost, inst)
    return q, cost, selects

"""Multi-candidate IK around the LM solve: Roberts seeds, staged population widths, best-k select.

The reference's population path for IK, on the MI355X library:

  roberts_samples              opt/sampler.py:246-264 (R_d low-discrepancy points in [0, 1))
  candidate_offsets            opt/sampler.py:338-346 (keep_first, seed offset), :31-39 (offset x scale)
  Stage, PopulationSolverConfig  opt/population_solver.py:52-156 (stages, select_at, num_candidates)
  PopulationIK.seeds           helpers/ik/ik.py:487-537 (joint mid-range + offset x joint range, clamp)
  PopulationIK.solve           opt/population_solver.py:291-395 (optimize -> score -> select per
                               segment, each segment a fresh LM; the winner of each problem last)
  default_ik_config            helpers/ik/config.py:37-45 (64 x 6 iterations at lambda 10, best 4 x 10
                               at lambda 1, best 1)

Every LM iteration and every select runs on the GPU (librkhip.so: rk_batch_solve,
rk_batch_select); the candidates never leave the device between stages.  There is no CPU
fallback.  The samplers that move candidates during a solve (MPPI / CEM / Gaussian draws),
score terms other than the optimisation cost, early stopping and the floating base are not
implemented: the configuration rejects them.
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import Dict, List, Optional

import numpy as np

from . import native as N

MAX_WIDTH = 64  # candidates per problem a select handles (one lane each)


def roberts_samples(num_points: int, dim: int, offset: int = 0) -> np.ndarray:
    """Points n = offset .. offset + num_points - 1 of the R_d sequence: frac(n * (1 - phi_d^-(k+1)))
    per column k, phi_d the real root > 1 of x^(d+1) = x + 1 (Newton from 1)."""
    root = 1.0
    for _ in range(10000):
        f = root ** (dim + 1) - root - 1.0
        root -= f / ((dim + 1) * root ** dim - 1.0)
        if abs(f) < 1e-10:
            break
    basis = 1.0 - 1.0 / root ** (1.0 + np.arange(dim))  # this rounding order reproduces the published points
    n = np.arange(num_points, dtype=np.float64) + offset
    return np.modf(n[:, None] * basis[None, :])[0]


def candidate_offsets(num_candidates: int, dim: int, seed: Optional[int] = None, keep_first: bool = True) -> np.ndarray:
    """float32 [num_candidates, dim] offsets in [-0.5, 0.5): Roberts points (from seed % 10000) minus
    1/2; keep_first puts candidate 0 on the centre and shifts the rest by one"""
    off = roberts_samples(num_candidates, dim, (seed % 10000) if seed is not None else 0) - 0.5
    if keep_first:
        off = np.vstack([np.zeros_like(off[:1]), off[:-1]])
    return off.astype(np.float32)


@dataclass
class Stage:
    """``iters`` LM iterations at ``num_candidates`` per problem (None: the previous stage's, or 1);
    ``optimizer``: its damping parameters (None: the previous stage's, or the LM defaults).
    Entering a stage keeps each problem's best ``num_candidates``."""

    iters: int = 0
    num_candidates: Optional[int] = None
    optimizer: Optional[N.LMParams] = None


@dataclass
class PopulationSolverConfig:
    stages: List[Stage] = field(default_factory=lambda: [Stage(iters=10)])
    # score only the final population (with score terms; earlier selects keep the optimization
    # cost) and return it best first (population_solver.py:88-89, 384-404)
    rank_final: bool = False

    def __post_init__(self):
        self.stages = list(self.stages)
        if not self.stages or any(s.iters < 0 for s in self.stages):
            raise ValueError("stages must be a non-empty list with non-negative iters")
        w = self.num_candidates
        if any(b > a for a, b in zip(w, w[1:])):
            k = next(i for i, (a, b) in enumerate(zip(w, w[1:])) if b > a)
            raise ValueError(f"select {k} grows {w[k]} -> {w[k + 1]}; selects only narrow")
        if w[0] > MAX_WIDTH or w[0] < 1:
            raise ValueError(f"num_candidates must be in 1..{MAX_WIDTH}, got {w[0]}")

    @property
    def num_iters(self) -> int:
        return sum(s.iters for s in self.stages)

    @property
    def select_at(self) -> List[int]:
        """iterations (after the update) where a stage ends and the next one's width is kept"""
        return [int(i) for i in np.cumsum([s.iters for s in self.stages])[:-1]]

    @property
    def num_candidates(self) -> List[int]:
        widths = [1]
        for s in self.stages:
            widths.append(widths[-1] if s.num_candidates is None else int(s.num_candidates))
        return widths[1:]

    @property
    def optimizer(self) -> List[N.LMParams]:
        out: List[Optional[N.LMParams]] = [None]
        for s in self.stages:
            out.append(s.optimizer or out[-1])
        return [o if o is not None else N.LMParams() for o in out[1:]]


def default_ik_config() -> PopulationSolverConfig:
    return PopulationSolverConfig(stages=[
        Stage(6, 64, N.LMParams(lm_lambda=10.0)),
        Stage(10, 4, N.LMParams(lm_lambda=1.0)),
        Stage(0, 1),
    ])


class PopulationIK:
    """``batch_size`` IK problems of one robot and term structure, each solved as a population of
    candidates (one device batch per population width; they stay resident across solves)."""

    def __init__(self, ctx: N.Context, spec, terms: N.TermSpec, batch_size: int,
                 config: Optional[PopulationSolverConfig] = None, seed: Optional[int] = None,
                 keep_first: bool = True, scale: float = 1.0, score_terms: Optional[N.TermSpec] = None):
        """``score_terms``: a term structure whose cost ranks the candidates at every select instead
        of the optimization cost (PopulationSolver score_terms, opt/population_solver.py:163-186,
        415-434); the winners' costs are then their scores"""
        self.ctx, self.spec, self.terms = ctx, spec, terms
        self.score_terms = score_terms
        self._score_targets: dict = {}
        self._score_set: set = set()  # widths whose score batch holds the current targets
        self.config = config or default_ik_config()
        self.batch_size = int(batch_size)
        self.widths = self.config.num_candidates
        self._batches: Dict[int, N.Batch] = {}
        self._ranked: Optional[N.Batch] = None  # the final population ranked (rank_final)
        self._score: Dict[int, N.Batch] = {}  # per width: the score terms' batch (score_terms only)
        for w in dict.fromkeys(self.widths + [1]):
            self._batches[w] = N.Batch(ctx, spec, terms, self.batch_size * w)
        self.offsets = candidate_offsets(self.widths[0], spec.num_actuated_joints,
                                         None if seed is None else int(seed), keep_first)
        self.seeds = self.candidate_seeds(scale)
        self._global_loaded = False  # the tiled global seeds are resident in the widest batch
        self._ready = False

    def candidate_seeds(self, scale: float = 1.0, center: Optional[np.ndarray] = None) -> np.ndarray:
        """Joint seeds: centre + offset x (scale x joint range), clamped to the float32 limits.
        center None: the joint mid-range, [width0, na] shared by every problem; else
        center [batch, na]: [batch * width0, na], problem i's candidates in rows i * width0 + j."""
        lim = np.asarray(self.spec.actuated_joint_limits)
        rng = (lim.astype(np.float32)[:, 1] - lim.astype(np.float32)[:, 0]) * np.float32(scale)
        step = self.offsets * rng[None, :]
        if center is None:
            q = (0.5 * (lim[:, 0] + lim[:, 1])).astype(np.float32)[None, :] + step
        else:
            c = np.asarray(center, np.float32).reshape(-1, lim.shape[0])
            q = (c[:, None, :] + step[None, :, :]).reshape(-1, lim.shape[0])
        lo, hi = self.spec.actuated_joint_limits_f32[:, 0], self.spec.actuated_joint_limits_f32[:, 1]
        return np.clip(q, lo, hi).astype(np.float32)

    def set_targets(self, pose_targets: Optional[np.ndarray] = None, pos_targets: Optional[np.ndarray] = None,
                    q_prev: Optional[np.ndarray] = None, init_q: Optional[np.ndarray] = None,
                    init_scale: float = 0.2, rot_targets: Optional[np.ndarray] = None,
                    score_targets: Optional[dict] = None) -> None:
        """each problem's targets [batch, ...], broadcast over its candidates (rows i * width + j).
        Seeds: init_q None: the global mid-range candidates (uploaded once, then kept resident);
        else init_q [batch, na] spread by init_scale x joint range (helpers/ik/ik.py:509-532).
        score_targets: the score terms' per-problem arrays (Batch.set keywords), with score_terms"""
        if self.score_terms is not None:
            self._score_targets = dict(score_targets or {})
            self._score_set.clear()
        w0 = self.widths[0]
        rep = lambda a: None if a is None else np.repeat(np.asarray(a, np.float32), w0, axis=0)  # noqa: E731
        b0 = self._batches[w0]
        q = None
        if init_q is not None:
            if w0 == 1:  # one candidate sits on its centre
                q = np.clip(np.asarray(init_q, np.float32), self.spec.actuated_joint_limits_f32[:, 0],
                            self.spec.actuated_joint_limits_f32[:, 1])
            else:
                q = self.candidate_seeds(init_scale, center=init_q)
            self._global_loaded = False
        elif not self._global_loaded:
            q = np.tile(self.seeds, (self.batch_size, 1))
            self._global_loaded = True
        b0.set(q=q, pose_targets=rep(pose_targets), pos_targets=rep(pos_targets), q_prev=rep(q_prev),
               n=self.batch_size * w0, rot_targets=rep(rot_targets))
        self._ready = True

    def solve(self, tpp=0, final_population: bool = False) -> N.Batch:
        """Run the schedule from the seeds; returns the width-1 batch of winners (get / metrics), or
        with ``final_population`` the final stage's batch (width num_candidates[-1], rows
        i * width + j; best first with config.rank_final -- PopulationSolver.final_var,
        population_solver.py:384-404).  ``tpp``: the threads-per-problem code of every stage, or a
        sequence of one per stage.  The launches and selects are only enqueued on the context's
        stream: reading the winners (get / metrics) waits for them."""
        if not self._ready:
            raise RuntimeError("set_targets first")
        cfg, widths = self.config, self.widths
        params = cfg.optimizer
        cur, w = self._batches[widths[0]], widths[0]
        for k, stage in enumerate(cfg.stages):
            # a stage boundary that narrows the population selects; at an equal width the population
            # is kept as it is (PopulationSolver._select returns it, population_solver.py:436-443);
            # either way the next segment restarts LM
            if k > 0 and widths[k] < w:
                nxt = self._batches[widths[k]]
                self._select(cur, w, nxt, widths[k], scored=not cfg.rank_final)
                cur, w = nxt, widths[k]
            if stage.iters > 0 or k == 0:  # a zero-iteration first stage scores the seeds
                tk = tpp[k] if isinstance(tpp, (list, tuple)) else tpp
                cur.solve(stage.iters, params[k], tpp=tk, from_seed=(k == 0), wait=False)
        if final_population:
            if cfg.rank_final and w > 1:  # the final population ranked best first (stable)
                if self._ranked is None:
                    self._ranked = N.Batch(self.ctx, self.spec, self.terms, self.batch_size * w)
                self._select(cur, w, self._ranked, w)
                cur = self._ranked
            return cur
        if w > 1:  # one winner per problem
            win = self._batches[1]
            self._select(cur, w, win, 1)
            cur = win
        return cur

    def score_batch(self, w: int) -> N.Batch:
        """the score terms' batch of width ``w`` (its targets uploaded once per set_targets)"""
        if w not in self._score:
            self._score[w] = N.Batch(self.ctx, self.spec, self.score_terms, self.batch_size * w)
        sb = self._score[w]
        if w not in self._score_set:
            rep = {k: np.repeat(np.asarray(v, np.float32), w, axis=0) for k, v in self._score_targets.items()
                   if v is not None}
            sb.set(n=self.batch_size * w, **rep)
            self._score_set.add(w)
        return sb

    def _select(self, cur: N.Batch, w: int, dst: N.Batch, w_out: int, scored: bool = True) -> None:
        """keep each problem's best w_out of its w candidates, ranked by cost or, with score terms
        (and scored), by the score terms' cost at the candidates' iterates"""
        if self.score_terms is None or not scored:
            cur.select(w, dst, w_out)
            return
        sb = self.score_batch(w)
        cur.copy_q_to(sb)
        sb.solve(0, N.LMParams(), wait=False)
        cur.select(w, dst, w_out, scores=sb)

    def close(self) -> None:
        for b in list(self._batches.values()) + list(self._score.values()) + [self._ranked]:
            if b is not None:
                b.close()

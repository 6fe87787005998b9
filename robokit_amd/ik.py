"""IKSolver front end on the MI355X population path.

Mirrors the reference's IK helper (helpers/ik/ik.py, helpers/ik/config.py) for its default use:

  IKSolverConfig      helpers/ik/config.py:49-66 (solver schedule, Roberts sampler settings)
  IKSolver.solve_numpy  helpers/ik/ik.py:223-304, 349-372: candidates around the joint
                      mid-range (drawn once, kept resident) or around init_q (init_sample_range
                      x joint range), the population schedule, and prev_q kept when its cost is no
                      worse than the winner's (select_better_score_kernel, :43-57)
  compute_metrics_numpy  helpers/ik/ik.py:387-470 (_ik_metrics_kernel, :60-82): per-target
                      position / rotation error and success at 5 mm / 0.05 rad by default

The optimisation terms are FrameTask(log) on every target link (position / orientation weights)
plus PositionLimit: the terms the reference's IK presets start from.  Every solve, score and
select runs on the GPU through librkhip.so.  Score terms, rest_state pinning, scenes, the
floating base and the torch adapter are not implemented (ValueError / NotImplementedError).
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import Optional, Sequence, Tuple, Union

import numpy as np

from . import native as N
from .population import PopulationIK, PopulationSolverConfig, default_ik_config


@dataclass(frozen=True)
class RobertsSamplerConfig:
    """opt/sampler.py:291-306: ``scale`` is the spread around init_q as a fraction of each joint's
    range (the mid-range candidates always span the whole range)"""

    scale: float = 0.2
    seed: Optional[int] = None
    keep_first: bool = False


@dataclass
class IKSolverConfig:
    solver: PopulationSolverConfig = field(default_factory=default_ik_config)
    sampler: RobertsSamplerConfig = field(default_factory=lambda: RobertsSamplerConfig(keep_first=True))
    position_weight: float = 1.0
    orientation_weight: float = 1.0
    limit_weight: float = 1.0
    enable_T_world_base: bool = False

    def __post_init__(self):
        if self.enable_T_world_base:
            raise NotImplementedError("the floating base is not implemented on this path")


class IKSolver:
    """IK for one robot and a fixed list of target links; batches of any size up to the first
    solve's (buffers are sized at the first solve, as the reference's)."""

    def __init__(self, config: IKSolverConfig, spec, target_links: Union[int, Sequence[int]],
                 ctx: Optional[N.Context] = None):
        self.config = config
        self.spec = spec
        self.target_links = [int(target_links)] if np.isscalar(target_links) else [int(x) for x in target_links]
        if not self.target_links:
            raise ValueError("at least one target link")
        if any(not 0 <= l < spec.num_links for l in self.target_links):
            raise ValueError(f"target link out of range (robot has {spec.num_links} links)")
        self.ctx = ctx or N.Context(0)
        na = spec.num_actuated_joints
        k = len(self.target_links)
        w = np.tile(np.array([config.position_weight] * 3 + [config.orientation_weight] * 3, np.float32), (k, 1))
        self.terms = N.TermSpec(pose_links=self.target_links, pose_weights=w,
                                limit_weights=np.full(na, config.limit_weight, np.float32))
        if config.solver.num_candidates[-1] != 1:
            raise NotImplementedError("a final stage of more than one solution per target is not implemented")
        self._pop: Optional[PopulationIK] = None
        self._score: Optional[N.Batch] = None
        self._metrics: Optional[N.Batch] = None
        self.last_cost: Optional[np.ndarray] = None
        self.last_prev_cost: Optional[np.ndarray] = None

    @property
    def num_targets(self) -> int:
        return len(self.target_links)

    def _targets(self, T_world_target) -> np.ndarray:
        t = np.asarray(T_world_target, np.float32)
        if not 1 <= t.ndim <= 3 or t.shape[-1] != 7:
            raise ValueError(f"T_world_target must have shape (..., 7), got {t.shape}")
        return np.ascontiguousarray(t.reshape(-1, self.num_targets, 7))

    def _build(self, batch_size: int) -> None:
        if self._pop is None:
            s = self.config.sampler
            self._pop = PopulationIK(self.ctx, self.spec, self.terms, batch_size, self.config.solver, seed=s.seed,
                                     keep_first=s.keep_first)
            self._score = N.Batch(self.ctx, self.spec, self.terms, batch_size)
        elif batch_size != self._pop.batch_size:
            raise ValueError(f"Batch size mismatch: expected {self._pop.batch_size}, got {batch_size}")

    def solve_numpy(self, T_world_target, prev_q: Optional[np.ndarray] = None, init_q: Optional[np.ndarray] = None,
                    init_sample_range: Optional[float] = None) -> np.ndarray:
        """q [batch, na] of each target's winner"""
        tg = self._targets(T_world_target)
        B = tg.shape[0]
        self._build(B)
        na = self.spec.num_actuated_joints
        if init_q is not None and np.asarray(init_q).reshape(-1, na).shape[0] != B:
            raise ValueError(f"init_q batch size mismatch: expected {B}")
        if prev_q is not None and np.asarray(prev_q).reshape(-1, na).shape[0] != B:
            raise ValueError(f"prev_state batch size mismatch: expected {B}, got {np.asarray(prev_q).reshape(-1, na).shape[0]}")
        scale = init_sample_range if init_sample_range is not None else self.config.sampler.scale
        self._pop.set_targets(pose_targets=tg, init_q=init_q, init_scale=scale)
        win = self._pop.solve()
        out = win.get()
        q, cost = out["q"], out["cost"]
        self.last_cost, self.last_prev_cost = cost, None
        if prev_q is not None:  # score prev with the solver's ranking rule; keep it on a tie
            pq = np.asarray(prev_q, np.float32).reshape(B, na)
            self._score.set(q=pq, pose_targets=tg)
            self._score.solve(0, self.config.solver.optimizer[-1], from_seed=True)
            pc = self._score.get()["cost"]
            keep = pc <= cost
            q = np.where(keep[:, None], pq, q)
            self.last_prev_cost = pc
        return q

    def compute_metrics_numpy(self, q: np.ndarray, T_world_target, position_threshold_m: float = 0.005,
                              orientation_threshold_rad: float = 0.05) -> Tuple[np.ndarray, np.ndarray, np.ndarray]:
        """(position error [batch, targets] m, rotation error [batch, targets] rad, success [batch])"""
        tg = self._targets(T_world_target)
        na = self.spec.num_actuated_joints
        q = np.asarray(q, np.float32).reshape(-1, na)
        if q.shape[0] != tg.shape[0]:
            raise ValueError(f"Target batch size ({tg.shape[0]}) must match state batch size ({q.shape[0]})")
        if self._metrics is None or self._metrics.capacity < q.shape[0]:
            self._metrics = N.Batch(self.ctx, self.spec, self.terms, q.shape[0])
        self._metrics.set(q=q, pose_targets=tg)
        pe, ae, ok = self._metrics.metrics(position_threshold_m, orientation_threshold_rad)
        return pe, ae, ok.astype(bool)

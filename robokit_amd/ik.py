"""IKSolver front end on the MI355X population path.

Mirrors the reference's IK helper (helpers/ik/ik.py, helpers/ik/config.py, helpers/ik/presets.py):

  IKContext           helpers/ik/config.py:12-33 (the values an IKSolver gives its term templates)
  IKSolverConfig      helpers/ik/config.py:49-92 (term templates added with ``add``, the solver
                      schedule, the Roberts sampler settings)
  presets             helpers/ik/presets.py:13-36 (basic, smooth, single_candidate)
  IKSolver.solve_numpy  helpers/ik/ik.py:223-304, 349-372: candidates around the joint
                      mid-range (drawn once, kept resident) or around init_q (init_sample_range
                      x joint range), the population schedule, and prev_q kept when its cost is no
                      worse than the winner's (select_better_score_kernel, :43-57)
  IKSolver.solve_torch  helpers/ik/ik.py:307-347: 4 x 4 target matrices and torch tensors in and
                      out (torch imported only there; matrix_to_pose by Shepperd's method)
  compute_metrics_numpy  helpers/ik/ik.py:387-470 (_ik_metrics_kernel, :60-82): per-target
                      position / rotation error and success at 5 mm / 0.05 rad by default

Term templates (robokit_amd.lm FrameTask / PositionTask / RotationTask / PositionLimit /
SmoothnessTask / SceneCollisionTask built without a robot) take IKContext.TARGET_LINKS /
TARGET_LINK as frames, IKContext.TARGET as targets, IKContext.PREVIOUS_STATE or IKContext.REST_Q
(the rest configuration: zero_q until solve_numpy(rest_q=...)) as the smoothness reference, and
IKContext.SCENE (IKSolver(scene=...)) as the collision scene.  A config without terms solves
FrameTask(log) on every target link (``position_weight`` / ``orientation_weight``) plus
PositionLimit(``limit_weight``).  Every solve, score and select runs on the GPU through
librkhip.so.  Score terms (IKSolverConfig.add_score) rank the candidates at every select and
the previous state by their own cost.  A final stage of several candidates returns all of them
per target, instance-major (solver.rank_final: best first).  The floating base is not implemented
(NotImplementedError).
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import List, NamedTuple, Optional, Sequence, Tuple, Union

import numpy as np

from . import lm as L
from . import native as N
from .population import PopulationIK, PopulationSolverConfig, Stage, default_ik_config


class IKContext(L.Context):
    """Values an IKSolver gives its term templates (helpers/ik/config.py:12-33)."""

    TARGET_LINKS = 0
    TARGET_LINK = 1
    TARGET = 2
    PREVIOUS_STATE = 3
    REST_Q = 4
    REST_BASE = 5
    NUM_SOLUTIONS = 6
    OPTIMIZATION_TERMS = 7
    SCENE = 8
    SCENE_INDICES = 9


@dataclass(frozen=True)
class RobertsSamplerConfig:
    """opt/sampler.py:291-306: ``scale`` is the spread around init_q as a fraction of each joint's
    range (the mid-range candidates always span the whole range)"""

    scale: float = 0.2
    seed: Optional[int] = None
    keep_first: bool = False


@dataclass
class IKSolverConfig:
    terms: List[L.Task] = field(default_factory=list)
    solver: PopulationSolverConfig = field(default_factory=default_ik_config)
    sampler: RobertsSamplerConfig = field(default_factory=lambda: RobertsSamplerConfig(keep_first=True))
    position_weight: float = 1.0
    orientation_weight: float = 1.0
    limit_weight: float = 1.0
    enable_T_world_base: bool = False
    score_terms: List[L.Task] = field(default_factory=list)

    def __post_init__(self):
        if self.enable_T_world_base:
            raise NotImplementedError("the floating base is not implemented on this path")

    def add(self, term: L.Task, *, name: Optional[str] = None) -> "IKSolverConfig":
        """append a term template (helpers/ik/config.py:73-76)"""
        if name is not None:
            term.name = name
            if term._template is not None:
                term._template["name"] = name
        self.terms.append(term)
        return self

    def add_score(self, term: L.Task, *, name: Optional[str] = None) -> "IKSolverConfig":
        """append a score term template: the score terms' cost ranks the candidates at every select
        instead of the optimization cost (helpers/ik/config.py:57-59, 78-81)"""
        if name is not None:
            term.name = name
            if term._template is not None:
                term._template["name"] = name
        self.score_terms.append(term)
        return self

    def task(self, name: str) -> L.Task:
        for t in self.terms + self.score_terms:
            if t.name == name:
                return t
        raise KeyError(name)


class presets:
    """The reference's IK presets (helpers/ik/presets.py:13-36)."""

    basic = (IKSolverConfig(sampler=RobertsSamplerConfig(keep_first=True, scale=1.0))
             .add(L.PositionTask(frame_index=IKContext.TARGET_LINKS, T_world_target=IKContext.TARGET, weight=20.0))
             .add(L.RotationTask(frame_index=IKContext.TARGET_LINKS, T_world_target=IKContext.TARGET, weight=10.0))
             .add(L.PositionLimit(weight=50.0)))
    smooth = (IKSolverConfig(sampler=RobertsSamplerConfig(keep_first=True, scale=1.0))
              .add(L.PositionTask(frame_index=IKContext.TARGET_LINKS, T_world_target=IKContext.TARGET, weight=20.0))
              .add(L.RotationTask(frame_index=IKContext.TARGET_LINKS, T_world_target=IKContext.TARGET, weight=10.0))
              .add(L.PositionLimit(weight=50.0))
              .add(L.SmoothnessTask(prev_var=IKContext.PREVIOUS_STATE, weight=0.1)))
    single_candidate = (
        IKSolverConfig(solver=PopulationSolverConfig(stages=[Stage(30, 1, N.LMParams(lm_lambda=1.0))]))
        .add(L.PositionTask(frame_index=IKContext.TARGET_LINKS, T_world_target=IKContext.TARGET, weight=20.0))
        .add(L.RotationTask(frame_index=IKContext.TARGET_LINKS, T_world_target=IKContext.TARGET, weight=10.0))
        .add(L.PositionLimit(weight=50.0))
        .add(L.SmoothnessTask(prev_var=IKContext.PREVIOUS_STATE, weight=0.1)))


class IKResultTorch(NamedTuple):
    """helpers/ik/ik.py:93-95"""

    q: "object"
    T_world_base: Optional["object"] = None


def matrix_to_pose(T: np.ndarray) -> np.ndarray:
    """homogeneous transforms [..., 4, 4] -> poses [..., 7] (x, y, z, qw, qx, qy, qz): the rotation's
    unit quaternion by Shepperd's method (the largest of 4 qw^2, 4 qx^2, 4 qy^2, 4 qz^2 picks the
    well-conditioned formula), scalar part made non-negative"""
    T = np.asarray(T, np.float64)
    R = T[..., :3, :3]
    t = T[..., :3, 3]
    tr = R[..., 0, 0] + R[..., 1, 1] + R[..., 2, 2]
    cand = np.stack([tr, R[..., 0, 0], R[..., 1, 1], R[..., 2, 2]], -1)
    k = np.argmax(cand, -1)
    q = np.zeros(T.shape[:-2] + (4,))
    m = k == 0
    s0 = np.sqrt(np.maximum(1.0 + tr[m], 0.0)) * 2
    q[m] = np.stack([0.25 * s0, (R[m][:, 2, 1] - R[m][:, 1, 2]) / s0, (R[m][:, 0, 2] - R[m][:, 2, 0]) / s0,
                     (R[m][:, 1, 0] - R[m][:, 0, 1]) / s0], -1)
    for i in range(3):
        m = k == i + 1
        if not np.any(m):
            continue
        j, l = (i + 1) % 3, (i + 2) % 3
        Rm = R[m]
        si = np.sqrt(np.maximum(1.0 + Rm[:, i, i] - Rm[:, j, j] - Rm[:, l, l], 0.0)) * 2
        qq = np.zeros((Rm.shape[0], 4))
        qq[:, 0] = (Rm[:, l, j] - Rm[:, j, l]) / si
        qq[:, 1 + i] = 0.25 * si
        qq[:, 1 + j] = (Rm[:, j, i] + Rm[:, i, j]) / si
        qq[:, 1 + l] = (Rm[:, l, i] + Rm[:, i, l]) / si
        q[m] = qq
    q *= np.where(q[..., :1] < 0, -1.0, 1.0)
    q /= np.linalg.norm(q, axis=-1, keepdims=True)
    return np.concatenate([t, q], -1).astype(np.float32)


class IKSolver:
    """IK for one robot and a fixed list of target links; the batch size is fixed by the first
    solve (buffers are sized then, as the reference's)."""

    def __init__(self, config: IKSolverConfig, spec, target_links: Union[int, Sequence[int]],
                 ctx: Optional[N.Context] = None, scene: Optional[L.PrimitiveScene] = None):
        """``spec``: a RobotSpec, or an lm.Robot (whose collision spheres SceneCollisionTask terms
        use); ``scene``: the collision scene IKContext.SCENE stands for (IKSolver(scene=...),
        helpers/ik/ik.py:101-111)"""
        self.config = config
        self.robot = spec if isinstance(spec, L.Robot) else L.Robot(spec)
        spec = self.robot.spec
        self.spec = spec
        self.scene = scene
        self.target_links = [int(target_links)] if np.isscalar(target_links) else [int(x) for x in target_links]
        if not self.target_links:
            raise ValueError("at least one target link")
        if any(not 0 <= l < spec.num_links for l in self.target_links):
            raise ValueError(f"target link out of range (robot has {spec.num_links} links)")
        self.ctx = ctx or N.Context(0)
        self._plan_terms()
        self._pop: Optional[PopulationIK] = None
        self._score: Optional[N.Batch] = None
        self._metrics: Optional[N.Batch] = None
        self.last_cost: Optional[np.ndarray] = None
        self.last_prev_cost: Optional[np.ndarray] = None

    def _plan_terms(self) -> None:
        """resolve the config's term (and score term) templates against the robot and the target
        links: each term's frames become indices into the target list, whose poses feed its
        targets at solve time"""
        k = len(self.target_links)
        templates = self.config.terms or [
            L.FrameTask(frame_index=IKContext.TARGET_LINKS, T_world_target=IKContext.TARGET,
                        position_weight=self.config.position_weight, orientation_weight=self.config.orientation_weight),
            L.PositionLimit(weight=self.config.limit_weight)]
        self.terms, self._sel, self.smooth_ref = self._resolve(templates)
        self.score_terms, self._score_sel, self._score_ref = (self._resolve(self.config.score_terms)
                                                              if self.config.score_terms else (None, None, None))
        refs = {r for r in (self.smooth_ref, self._score_ref) if r}
        if len(refs) > 1:
            raise NotImplementedError("terms and score terms must share one smoothness reference on this path")
        self.needs_prev = "prev" in refs
        self._metric_terms = N.TermSpec(pose_links=self.target_links, pose_weights=np.ones((k, 6), np.float32))

    def _resolve(self, templates):
        """(TermSpec, target-list index of each row block, smoothness reference "prev" / "rest" /
        None) of a list of term templates"""
        k = len(self.target_links)
        robot = self.robot
        ctxv = {IKContext.TARGET_LINKS: list(self.target_links), IKContext.TARGET_LINK: self.target_links[0],
                IKContext.TARGET: np.tile(np.array([0, 0, 0, 1, 0, 0, 0], np.float32), (1, k, 1)),
                IKContext.PREVIOUS_STATE: L.RobotState(robot, self.spec.zero_q),
                IKContext.REST_Q: L.RobotState(robot, self.spec.zero_q), IKContext.SCENE: self.scene}
        sel = {"pose": [], "pos": [], "rot": []}
        w = {"pose": [], "pos": [], "rot": []}
        limit = smooth = coll = None
        smooth_ref = None
        for t in templates:
            tmpl = t._template or {}
            if isinstance(t, (L.FrameTask, L.PositionTask, L.RotationTask)) and tmpl.get("T_world_target") is not IKContext.TARGET:
                raise NotImplementedError("IK term targets must be IKContext.TARGET on this path")
            if isinstance(t, L.SmoothnessTask):
                ref = tmpl.get("prev_var")
                if ref not in (IKContext.PREVIOUS_STATE, IKContext.REST_Q):
                    raise NotImplementedError("the IK smoothness reference must be IKContext.PREVIOUS_STATE or REST_Q")
                if smooth_ref is not None:
                    raise NotImplementedError("one smoothness reference per IK configuration on this path")
                smooth_ref = "prev" if ref is IKContext.PREVIOUS_STATE else "rest"
            if isinstance(t, L.SceneCollisionTask) and tmpl.get("scene") is IKContext.SCENE and self.scene is None:
                raise ValueError("a SceneCollisionTask on IKContext.SCENE needs IKSolver(scene=...)")
            r = t.resolve(robot, ctxv)
            key = {L.FrameTask: "pose", L.PositionTask: "pos", L.RotationTask: "rot"}.get(type(r))
            if key:
                for f in r.frame_indices:
                    if f not in self.target_links:
                        raise ValueError(f"frame {f} is not a target link")
                    sel[key].append(self.target_links.index(f))
                w[key].append(r.weights)
            elif isinstance(r, L.PositionLimit):
                limit = r.weights
            elif isinstance(r, L.SmoothnessTask):
                smooth = r.weights
            elif isinstance(r, L.SceneCollisionTask):
                if coll is not None:
                    raise NotImplementedError("one SceneCollisionTask per IK configuration on this path")
                coll = r
            else:
                raise NotImplementedError(f"IK term {type(r).__name__} is not on this path")
        cat = lambda key: np.concatenate(w[key]) if w[key] else None  # noqa: E731
        ts = N.TermSpec(pose_links=[self.target_links[i] for i in sel["pose"]], pose_weights=cat("pose"),
                        pos_links=[self.target_links[i] for i in sel["pos"]], pos_weights=cat("pos"),
                        rot_links=[self.target_links[i] for i in sel["rot"]], rot_weights=cat("rot"),
                        limit_weights=limit, smooth_weights=smooth)
        if coll is not None:  # SceneCollisionTask rows (the robot's spheres against the scene)
            ts.coll_links, ts.coll_centers, ts.coll_radii = list(coll.sphere_links), coll.sphere_centers, coll.sphere_radii
            ts.coll_weights, ts.coll_margin, ts.coll_penalty = coll.weights, coll.margin, coll.penalty
            ts.obstacle_spheres, ts.obstacle_boxes, ts.geoms = coll.scene.spheres, coll.scene.boxes, list(coll.scene.geoms)
        return ts, sel, smooth_ref

    def term_targets(self, tg: np.ndarray, sel: Optional[dict] = None) -> dict:
        """the batch's target arrays (pose [B, n, 7], position [B, n, 3], rotation [B, n, 4]) from
        the target poses [B, k, 7] (of the optimization terms, or of the row blocks ``sel``)"""
        sel = self._sel if sel is None else sel
        out = {}
        for key, sl in (("pose_targets", "pose"), ("pos_targets", "pos"), ("rot_targets", "rot")):
            idx = sel[sl]
            if idx:
                a = tg[:, idx]
                out[key] = np.ascontiguousarray(a[..., :3] if sl == "pos" else a[..., 3:] if sl == "rot" else a)
        return out

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
                                     keep_first=s.keep_first, score_terms=self.score_terms)
            # the previous state is ranked like the candidates: by the score terms, else the terms
            self._score = N.Batch(self.ctx, self.spec, self.score_terms or self.terms, batch_size)
        elif batch_size != self._pop.batch_size:
            raise ValueError(f"Batch size mismatch: expected {self._pop.batch_size}, got {batch_size}")

    @property
    def num_solutions(self) -> int:
        """solutions returned per target: the final stage's candidates (helpers/ik/ik.py:175-178)"""
        return self.config.solver.num_candidates[-1]

    def solve_numpy(self, T_world_target, prev_q: Optional[np.ndarray] = None, init_q: Optional[np.ndarray] = None,
                    init_sample_range: Optional[float] = None, rest_q: Optional[np.ndarray] = None) -> np.ndarray:
        """q [batch * num_solutions, na]: each target's final solutions, instance-major (rows
        i * num_solutions + j; with solver.rank_final best first), the winner when num_solutions
        is 1 (helpers/ik/ik.py:233-304).  rest_q [batch, na] or [na]: the rest configuration
        IKContext.REST_Q stands for (default: the robot's zero configuration, helpers/ik/ik.py:576)"""
        tg = self._targets(T_world_target)
        B = tg.shape[0]
        self._build(B)
        na = self.spec.num_actuated_joints
        ns = self.num_solutions
        if prev_q is not None and ns != 1:
            raise ValueError("prev_state requires an IK final stage with num_candidates=1.")
        if init_q is not None and np.asarray(init_q).reshape(-1, na).shape[0] != B:
            raise ValueError(f"init_q batch size mismatch: expected {B}")
        if prev_q is not None and np.asarray(prev_q).reshape(-1, na).shape[0] != B:
            raise ValueError(f"prev_state batch size mismatch: expected {B}, got {np.asarray(prev_q).reshape(-1, na).shape[0]}")
        if self.needs_prev and prev_q is None:
            raise ValueError("this configuration's SmoothnessTask needs prev_q (IKContext.PREVIOUS_STATE)")
        pq = None if prev_q is None else np.asarray(prev_q, np.float32).reshape(B, na)
        scale = init_sample_range if init_sample_range is not None else self.config.sampler.scale
        ref = None  # the smoothness terms' reference rows
        kind = self.smooth_ref or self._score_ref
        if kind == "prev":
            ref = pq
        elif kind == "rest":
            rq = self.spec.zero_q if rest_q is None else rest_q
            ref = np.ascontiguousarray(np.broadcast_to(np.asarray(rq, np.float32).reshape(-1, na), (B, na)))
        score_tg = None
        if self.score_terms is not None:
            score_tg = dict(self.term_targets(tg, self._score_sel), q_prev=ref if self._score_ref else None)
        self._pop.set_targets(**self.term_targets(tg), q_prev=ref if self.smooth_ref else None, init_q=init_q,
                              init_scale=scale, score_targets=score_tg)
        win = self._pop.solve(final_population=ns > 1)
        out = win.get()
        q, cost = out["q"], out["cost"]
        if self.score_terms is not None:  # the solutions' scores (also without a final select)
            sb = self._pop.score_batch(ns)
            win.copy_q_to(sb)
            sb.solve(0, N.LMParams())
            cost = sb.get()["cost"]
        self.last_cost, self.last_prev_cost = cost, None
        if pq is not None:  # score prev with the solver's ranking rule; keep it on a tie
            if self.score_terms is not None:
                self._score.set(q=pq, q_prev=ref if self._score_ref else None, **self.term_targets(tg, self._score_sel))
            else:
                self._score.set(q=pq, q_prev=ref if self.smooth_ref else None, **self.term_targets(tg))
            self._score.solve(0, self.config.solver.optimizer[-1], from_seed=True)
            pc = self._score.get()["cost"]
            keep = pc <= cost
            q = np.where(keep[:, None], pq, q)
            self.last_prev_cost = pc
        return q

    def solve_torch(self, T_world_target, prev_state=None, init_q=None, init_T_world_base=None, rest_q=None,
                    rest_T_world_base=None, init_sample_range: Optional[float] = None) -> IKResultTorch:
        """The torch adapter (helpers/ik/ik.py:307-347): T_world_target [..., 4, 4] homogeneous
        matrices, optional init_q / rest_q tensors and a previous state (tensor or array of q);
        returns IKResultTorch(q) on the targets' device.  torch is imported only by this call."""
        import torch

        if init_T_world_base is not None or rest_T_world_base is not None:
            raise NotImplementedError("the floating base is not implemented on this path")
        T = T_world_target
        if not 2 <= T.ndim <= 4 or tuple(T.shape[-2:]) != (4, 4):
            raise ValueError(f"T_world_target must have shape (..., 4, 4), got {tuple(T.shape)}")
        as_np = lambda x: None if x is None else (x.detach().cpu().numpy() if torch.is_tensor(x)  # noqa: E731
                                                   else np.asarray(getattr(x, "q", x)))
        poses = matrix_to_pose(as_np(T).reshape(-1, self.num_targets, 4, 4))
        q = self.solve_numpy(poses, prev_q=as_np(getattr(prev_state, "q", prev_state)), init_q=as_np(init_q),
                             init_sample_range=init_sample_range, rest_q=as_np(rest_q))
        return IKResultTorch(q=torch.from_numpy(np.ascontiguousarray(q)).to(T.device))

    def compute_metrics_numpy(self, q: np.ndarray, T_world_target, position_threshold_m: float = 0.005,
                              orientation_threshold_rad: float = 0.05) -> Tuple[np.ndarray, np.ndarray, np.ndarray]:
        """(position error [batch, targets] m, rotation error [batch, targets] rad, success [batch])"""
        tg = self._targets(T_world_target)
        na = self.spec.num_actuated_joints
        q = np.asarray(q, np.float32).reshape(-1, na)
        if q.shape[0] != tg.shape[0]:
            raise ValueError(f"Target batch size ({tg.shape[0]}) must match state batch size ({q.shape[0]})")
        if self._metrics is None or self._metrics.capacity < q.shape[0]:
            self._metrics = N.Batch(self.ctx, self.spec, self._metric_terms, q.shape[0])
        self._metrics.set(q=q, pose_targets=tg)
        pe, ae, ok = self._metrics.metrics(position_threshold_m, orientation_threshold_rad)
        return pe, ae, ok.astype(bool)

"""RoboKit's LM solver and residual-term API, on the MI355X library.

The classes mirror the reference's names, arguments, defaults and error behaviour so that a
caller of the reference's Levenberg-Marquardt path can switch without edits:

  VarKey, VarValues            opt/var_values.py:24-342 (one variable: the robot configuration)
  Robot, RobotState            robo/robot.py:105-130, robo/robot_state.py:41-476 (q, batch)
  FrameTask                    terms/dense/frame_task.py:72-82 (residual_mode "log")
  PositionTask                 terms/dense/position_task.py:42-50
  RotationTask                 terms/dense/rotation_task.py:25-80 (axis-angle rows, geometric Jacobian)
  PositionLimit                terms/dense/position_limit.py:204-220 (residual_mode "abs")
  SmoothnessTask               terms/dense/smoothness_task.py:41-50
  SceneCollisionTask           terms/dense/scene_collision_task.py:276-356 (robot spheres against a
                               PrimitiveScene of obstacle spheres and axis-aligned boxes, geom/scene.py)
  LMOptimizerConfig            opt/lm_optimizer.py:502-530 (+ OptimizerConfig, opt/optimizer.py:159-174)
  LMOptimizer                  opt/optimizer.py:177-333 lifecycle: solve / reset / step, costs

The whole solve of a batch runs on the GPU (librkhip.so through robokit_amd.native): terms are
compiled into the library's term descriptor, buffers are sized at the first solve's batch size,
and each `step` is one fused LM iteration continuing the previous one on the device.  There is
no CPU fallback.  Options of the reference this path does not implement raise
NotImplementedError instead of silently changing the algorithm.
"""
from __future__ import annotations

from dataclasses import dataclass, replace
from enum import Enum
from typing import Callable, Dict, List, Literal, Optional, Sequence, Tuple, Union

import numpy as np

from . import native as N
from .robo.robot_spec import RobotSpec

IterationCallback = Callable[[int, "VarValues", np.ndarray], None]
MAX_POSE_TARGETS = 4
MAX_POSITION_TARGETS = 12

_CONTEXTS: Dict[int, N.Context] = {}


def _device_ordinal(device) -> int:
    if device is None:
        return 0
    if isinstance(device, int):
        return device
    s = str(device)
    if ":" in s:
        return int(s.split(":")[1])
    if s in ("cuda", "hip", "gpu"):
        return 0
    raise ValueError(f"unsupported device {device!r}: this library runs on a HIP GPU ('hip:N' / 'cuda:N')")


def get_context(device=None) -> N.Context:
    """The library context (HIP stream + resident robots) of a device, created on first use."""
    d = _device_ordinal(device)
    if d not in _CONTEXTS:
        _CONTEXTS[d] = N.Context(d)
    return _CONTEXTS[d]


# ----------------------------------------------------------------------------- variables
class VarKey(Enum):
    """Variable keys (opt/var_values.py:73-85); this path optimises the robot configuration."""

    ROBOT = 0


class Robot:
    """A kinematic model (RobotSpec) with the state factory of robo/robot.py:105-130.
    ``collision_spheres`` = (link indices [n], centres in the link frames [n, 3], radii [n]): the
    robot's collision spheres (RobotSpec.collision_spheres_link_indices /
    local_collision_sphere_centers / radii in the reference)."""

    def __init__(self, spec: RobotSpec, collision_spheres=None):
        spec.validate()
        self.spec = spec
        self.collision_spheres = None
        if collision_spheres is not None:
            links, centers, radii = collision_spheres
            links = np.asarray(links, np.int32).reshape(-1)
            centers = np.asarray(centers, np.float32).reshape(-1, 3)
            radii = np.asarray(radii, np.float32).reshape(-1)
            if not (len(links) == len(centers) == len(radii)):
                raise ValueError("collision spheres need one link, centre and radius each")
            if np.any(links < 0) or np.any(links >= spec.num_links):
                raise ValueError("collision sphere link out of range")
            self.collision_spheres = (links, centers, radii)

    @property
    def has_collision_spheres(self) -> bool:
        return self.collision_spheres is not None and len(self.collision_spheres[0]) > 0

    @staticmethod
    def load(urdf: str) -> "Robot":
        """Robot from URDF text or a path to a URDF file."""
        return Robot(RobotSpec.parse_urdf(urdf))

    @property
    def link_names(self) -> List[str]:
        return list(self.spec.link_names)

    def state(self, q: Optional[np.ndarray] = None) -> "RobotState":
        """Robot state for joint positions ``q`` [na] or [B, na] (default: zero_q)."""
        if q is None:
            q = self.spec.zero_q
        return RobotState(self, q)

    def forward_kinematics(self, state: "RobotState", device=None) -> np.ndarray:
        """World pose of every link [B, num_links, 7] (x, y, z, qw, qx, qy, qz)."""
        return get_context(device).fk(self.spec, state.q)


class RobotState:
    """Joint positions of a batch, q [B, na] float32 (robo/robot_state.py:41-181)."""

    def __init__(self, robot: Robot, q: np.ndarray):
        q = np.array(q, dtype=np.float32)  # a copy: solves write the state in place
        if q.ndim == 1:
            q = q[None]
        if q.ndim != 2 or q.shape[1] != robot.spec.num_actuated_joints:
            raise ValueError(f"q must be [B, {robot.spec.num_actuated_joints}], got {q.shape}")
        self.robot = robot
        self.q = np.ascontiguousarray(q)

    @property
    def batch_size(self) -> int:
        return self.q.shape[0]

    @property
    def tangent_dim(self) -> int:
        return self.q.shape[1]

    def clone(self) -> "RobotState":
        return RobotState(self.robot, self.q.copy())


class VarValues:
    """The variables of a batch of problems (opt/var_values.py:24-342); here only VarKey.ROBOT."""

    def __init__(self, values: Dict[VarKey, RobotState]):
        if set(values) != {VarKey.ROBOT}:
            raise NotImplementedError("this path optimises the robot configuration only (VarKey.ROBOT)")
        self._values = dict(values)

    def get(self, key: VarKey) -> RobotState:
        return self._values[key]

    def __getitem__(self, key: VarKey) -> RobotState:
        return self._values[key]

    @property
    def batch_size(self) -> int:
        return self._values[VarKey.ROBOT].batch_size

    @property
    def tangent_dim(self) -> int:
        return self._values[VarKey.ROBOT].tangent_dim

    def clone(self) -> "VarValues":
        return VarValues({k: v.clone() for k, v in self._values.items()})


# ----------------------------------------------------------------------------- terms
def _frames(frame_index) -> List[int]:
    return [int(frame_index)] if np.isscalar(frame_index) else [int(f) for f in frame_index]


def _targets(T, nframes: int, width: int) -> np.ndarray:
    """targets as [B or 1, nframes, width] float32 (one per frame, optionally per problem)"""
    t = np.asarray(T, dtype=np.float32)
    if t.ndim == 1:
        t = t[None, None]
    elif t.ndim == 2:
        t = t[None] if (nframes > 1 and t.shape[0] == nframes) else t[:, None]
    if t.ndim != 3 or t.shape[1] != nframes or t.shape[2] not in (width, 7):
        raise ValueError(f"T_world_target must hold {nframes} target(s) of {width} (or 7) floats, got {t.shape}")
    return t[..., :width] if width == 3 else t


def _per_frame_weights(w, nframes: int, name: str) -> np.ndarray:
    a = np.asarray(w, dtype=np.float32)
    if a.ndim == 0:
        return np.full(nframes, float(a), np.float32)
    a = a.reshape(-1)
    if a.size != nframes:
        raise ValueError(f"{name} needs one value or one per frame ({nframes}), got {a.size}")
    return a


class Context(Enum):
    """Base of the keys that stand for a helper's values in task arguments (terms/task.py:13-15)."""


def _has_context(*args) -> bool:
    return any(isinstance(a, Context) or (isinstance(a, tuple) and any(isinstance(x, Context) for x in a))
               for a in args)


class Task:
    """A cost term.  Built without a robot, or with a Context placeholder among its arguments, it
    is a template (the IK presets' form, helpers/ik/presets.py:13-18): ``resolve`` swaps the
    placeholders for a helper's values and builds the concrete term for a robot."""

    name: str = ""
    _template: Optional[dict] = None

    @property
    def is_template(self) -> bool:
        return self._template is not None

    def _as_template(self, **kwargs) -> None:
        self._template = kwargs
        self.robot, self.name = None, kwargs.get("name", "")
        self.weight = kwargs.get("weight")

    def resolve(self, robot: "Robot", context: Dict["Context", object]) -> "Task":
        """the concrete term for ``robot`` with every Context placeholder replaced from ``context``"""
        if self._template is None:
            return self
        sub = lambda v: context[v] if isinstance(v, Context) else v  # noqa: E731
        args = {k: (tuple(sub(x) for x in v) if isinstance(v, tuple) else sub(v)) for k, v in self._template.items()}
        return type(self)(robot, **args)


class FrameTask(Task):
    """Pose residual r = w * log(T_ee^-1 T_target) per frame ([v; w] rows, terms/dense/frame_task.py:505-554)."""

    def __init__(self, robot: Optional[Robot] = None, frame_index: Union[int, Sequence[int], Context, None] = None,
                 T_world_target=None, position_weight: Union[float, Sequence[float]] = 1.0,
                 orientation_weight: Union[float, Sequence[float]] = 1.0,
                 residual_mode: Literal["log", "distance"] = "log", trajectory_frames: int = 1, name: str = ""):
        if robot is None or frame_index is None or T_world_target is None or _has_context(frame_index, T_world_target):
            self._as_template(frame_index=frame_index, T_world_target=T_world_target, position_weight=position_weight,
                              orientation_weight=orientation_weight, residual_mode=residual_mode,
                              trajectory_frames=trajectory_frames, name=name)
            return
        if residual_mode not in {"log", "distance"}:
            raise ValueError(f"Unsupported residual_mode: {residual_mode}")
        if residual_mode != "log":
            raise NotImplementedError("FrameTask residual_mode 'distance' is not on this path (log only)")
        if trajectory_frames != 1:
            raise NotImplementedError("trajectory FrameTask is not on this path")
        self.robot, self.name, self.residual_mode = robot, name, residual_mode
        self.frame_indices = _frames(frame_index)
        k = len(self.frame_indices)
        for f in self.frame_indices:
            if not 0 <= f < robot.spec.num_links:
                raise ValueError(f"frame_index {f} out of range (robot has {robot.spec.num_links} links)")
        self.position_weight, self.orientation_weight = position_weight, orientation_weight
        pw = _per_frame_weights(position_weight, k, "position_weight")
        ow = _per_frame_weights(orientation_weight, k, "orientation_weight")
        self.weights = np.stack([np.repeat(pw, 3).reshape(k, 3), np.repeat(ow, 3).reshape(k, 3)], 1).reshape(k, 6)
        self.T_world_target = _targets(T_world_target, k, 7)

    def set_target(self, T_world_target) -> None:
        self.T_world_target = _targets(T_world_target, len(self.frame_indices), 7)

    @property
    def residual_dim(self) -> int:
        return 6 * len(self.frame_indices)


class PositionTask(Task):
    """Point residual r = w * (p_target - p_link) per frame (terms/dense/position_task.py:188-212)."""

    def __init__(self, robot: Optional[Robot] = None, frame_index: Union[int, Sequence[int], Context, None] = None,
                 T_world_target=None, weight: Union[float, Sequence[float]] = 20.0, fixed_target: bool = False,
                 name: str = ""):
        if robot is None or frame_index is None or T_world_target is None or _has_context(frame_index, T_world_target):
            self._as_template(frame_index=frame_index, T_world_target=T_world_target, weight=weight,
                              fixed_target=fixed_target, name=name)
            return
        self.robot, self.name, self.weight, self.fixed_target = robot, name, weight, fixed_target
        self.frame_indices = _frames(frame_index)
        k = len(self.frame_indices)
        for f in self.frame_indices:
            if not 0 <= f < robot.spec.num_links:
                raise ValueError(f"frame_index {f} out of range (robot has {robot.spec.num_links} links)")
        w = np.asarray(weight, dtype=np.float32)
        if w.ndim == 0:
            self.weights = np.full((k, 3), float(w), np.float32)
        elif w.size == 3:
            self.weights = np.tile(w.reshape(1, 3), (k, 1))
        elif w.size == 3 * k:
            self.weights = w.reshape(k, 3)
        else:
            raise ValueError(f"weight must be a scalar, 3 values or 3 per frame, got {w.size}")
        self.T_world_target = _targets(T_world_target, k, 3)

    def set_target(self, T_world_target) -> None:
        self.T_world_target = _targets(T_world_target, len(self.frame_indices), 3)

    @property
    def residual_dim(self) -> int:
        return 3 * len(self.frame_indices)


class RotationTask(Task):
    """Orientation residual r = w * axis_angle(q_link * q_target^-1) per frame (rotation_task.py:323-349)."""

    def __init__(self, robot: Optional[Robot] = None, frame_index: Union[int, Sequence[int], Context, None] = None,
                 T_world_target=None, weight: Union[float, Sequence[float]] = 10.0, fixed_target: bool = False,
                 name: str = ""):
        if robot is None or frame_index is None or T_world_target is None or _has_context(frame_index, T_world_target):
            self._as_template(frame_index=frame_index, T_world_target=T_world_target, weight=weight,
                              fixed_target=fixed_target, name=name)
            return
        self.robot, self.name, self.weight, self.fixed_target = robot, name, weight, fixed_target
        self.frame_indices = _frames(frame_index)
        k = len(self.frame_indices)
        for f in self.frame_indices:
            if not 0 <= f < robot.spec.num_links:
                raise ValueError(f"frame_index {f} out of range (robot has {robot.spec.num_links} links)")
        w = np.asarray(weight, dtype=np.float32)
        if w.ndim == 0:
            self.weights = np.full((k, 3), float(w), np.float32)
        elif w.size == 3:
            self.weights = np.tile(w.reshape(1, 3), (k, 1))
        elif w.size == 3 * k:
            self.weights = w.reshape(k, 3)
        else:
            raise ValueError(f"weight must be a scalar, 3 values or 3 per frame, got {w.size}")
        self.T_world_target = _targets(T_world_target, k, 7)

    def set_target(self, T_world_target) -> None:
        self.T_world_target = _targets(T_world_target, len(self.frame_indices), 7)

    @property
    def residual_dim(self) -> int:
        return 3 * len(self.frame_indices)


class PositionLimit(Task):
    """Joint-limit hinge r_i = w_i (max(0, q_i - hi_i) + max(0, lo_i - q_i)) (position_limit.py:508-534)."""

    def __init__(self, robot: Optional[Robot] = None, weight: Optional[Union[float, Sequence[float]]] = None,
                 residual_mode: Literal["abs", "sqrt_abs", "exp_barrier"] = "abs", name: str = ""):
        if robot is None:
            self._as_template(weight=weight, residual_mode=residual_mode, name=name)
            return
        if residual_mode not in {"abs", "sqrt_abs", "exp_barrier"}:
            raise ValueError(f"Unsupported residual_mode: {residual_mode}")
        if residual_mode != "abs":
            raise NotImplementedError("PositionLimit residual_mode must be 'abs' on this path")
        self.robot, self.name, self.weight, self.residual_mode = robot, name, weight, residual_mode
        na = robot.spec.num_actuated_joints
        w = np.full(na, 1.0 if weight is None else 0.0, np.float32)
        if weight is not None:
            w[:] = np.asarray(weight, dtype=np.float32)
        self.weights = w

    @property
    def residual_dim(self) -> int:
        return self.robot.spec.num_actuated_joints


class SmoothnessTask(Task):
    """Temporal smoothness r = w (q - q_prev) (terms/dense/smoothness_task.py:350-365)."""

    def __init__(self, robot: Optional[Robot] = None, prev_var: Union[RobotState, Context, None] = None,
                 weight: Union[float, Sequence[float]] = 1.0, name: str = ""):
        if robot is None or _has_context(prev_var):
            self._as_template(prev_var=prev_var, weight=weight, name=name)
            return
        self.robot, self.name, self.weight = robot, name, weight
        na = robot.spec.num_actuated_joints
        w = np.zeros(na, np.float32)
        w[:] = np.asarray(weight, dtype=np.float32)
        self.weights = w
        self.prev_var = prev_var.clone() if prev_var is not None else None

    def set_prev(self, prev_var: RobotState) -> None:
        self.prev_var = prev_var.clone()

    @property
    def residual_dim(self) -> int:
        return self.robot.spec.num_actuated_joints


class PrimitiveScene:
    """A world scene of obstacle spheres and axis-aligned boxes: the primitive form of the
    reference's WarpScene with SphereGeom / BoxGeom members (geom/scene.py:132-256,
    geom/geoms.py:372-429), queried by its signed distance (geom/sdf_kernels.py:256-295)."""

    def __init__(self, spheres: Optional[np.ndarray] = None, boxes: Optional[np.ndarray] = None):
        self.spheres = np.zeros((0, 4), np.float32) if spheres is None else np.asarray(spheres, np.float32).reshape(-1, 4)
        self.boxes = np.zeros((0, 6), np.float32) if boxes is None else np.asarray(boxes, np.float32).reshape(-1, 6)

    def add_spheres(self, centers, radii) -> "PrimitiveScene":
        c = np.asarray(centers, np.float32).reshape(-1, 3)
        r = np.asarray(radii, np.float32).reshape(-1, 1)
        self.spheres = np.concatenate([self.spheres, np.concatenate([c, np.broadcast_to(r, (len(c), 1))], 1)])
        return self

    def add_boxes(self, centers, half_extents) -> "PrimitiveScene":
        c = np.asarray(centers, np.float32).reshape(-1, 3)
        h = np.broadcast_to(np.asarray(half_extents, np.float32).reshape(-1, 3), c.shape)
        self.boxes = np.concatenate([self.boxes, np.concatenate([c, h], 1)])
        return self

    @property
    def geoms(self) -> int:
        return len(self.spheres) + len(self.boxes)


class SceneCollisionTask(Task):
    """Collision rows r_k = w_k pen(sd_k) per robot sphere, sd_k the scene distance of the sphere
    centre minus its radius, pen the smooth / plain / surface-distance penalty
    (terms/dense/scene_collision_task.py:276-356, terms/sphere_collision_kernels.py:19-42)."""

    def __init__(self, robot: Optional[Robot] = None, scene: Optional[PrimitiveScene] = None, geometry=None,
                 scene_indices=None, use_link_bounding_sphere_filter: bool = False,
                 weight: Union[float, Sequence[float]] = 1.0, margin: Optional[float] = 0.02,
                 sphere_indices: Optional[Sequence[int]] = None, residual_mode: Literal["abs", "sqrt_abs"] = "abs",
                 residual_eps: float = 1e-6, penalty: Literal["smooth", "plain", "surface_distance"] = "smooth",
                 skip_trajectory_endpoints: bool = True, movable_only: bool = False, name: str = ""):
        if penalty not in {"smooth", "plain", "surface_distance"}:
            raise ValueError(f"Unsupported penalty: {penalty}")
        if robot is None or scene is None or _has_context(scene):
            self._as_template(scene=scene, weight=weight, margin=margin, sphere_indices=sphere_indices,
                              residual_mode=residual_mode, penalty=penalty, movable_only=movable_only, name=name)
            return
        if geometry is not None or scene_indices is not None or use_link_bounding_sphere_filter:
            raise NotImplementedError("geometry filters, per-query scenes and bounding-sphere filters are not on "
                                      "this path")
        if residual_mode != "abs":
            raise NotImplementedError("SceneCollisionTask residual_mode must be 'abs' on this path")
        if not robot.has_collision_spheres:
            raise RuntimeError("Robot has no collision spheres. Build it with collision_spheres=(links, centers, radii).")
        self.robot, self.name, self.weight, self.margin, self.penalty = robot, name, weight, float(margin), penalty
        links, centers, radii = robot.collision_spheres
        idx = np.arange(len(links))
        if sphere_indices is None and movable_only:
            idx = np.flatnonzero(robot.spec.link_parent_joint_indices[links] >= 0)
        elif sphere_indices is not None:
            idx = np.asarray(sphere_indices, np.int32).reshape(-1)
            if np.any(idx < 0) or np.any(idx >= len(links)):
                raise ValueError("sphere_indices contains out-of-range indices.")
        w = np.asarray(weight, np.float32)
        if w.ndim and w.shape != (len(idx),):
            raise ValueError(f"Expected weight shape ({len(idx)},), got {w.shape}.")
        self.sphere_links, self.sphere_centers, self.sphere_radii = links[idx], centers[idx], radii[idx]
        self.weights = np.broadcast_to(w, (len(idx),)).astype(np.float32)
        self.scene = scene

    @property
    def residual_dim(self) -> int:
        return len(self.sphere_links)


# ----------------------------------------------------------------------------- optimizer
@dataclass(frozen=True)
class OptimizerConfig:
    max_iter: int = 100
    early_stopping_interval: int = 1
    cost_tol: float = 1e-9
    cuda_graph_mode: Literal["none", "iter", "full"] = "none"


@dataclass(frozen=True)
class LMOptimizerConfig(OptimizerConfig):
    lm_lambda: float = 1e-3
    lambda_factor: float = 2.0
    lambda_min: float = 1e-5
    lambda_max: float = 1e10
    rho_min: float = 1e-3
    gain_ratio_epsilon: float = 1e-8
    linear_solver: Literal["cholesky", "cg", "torch_cholesky"] = "cholesky"
    cg_max_iter: Optional[int] = None
    cg_tol: float = 1e-6
    cg_check_interval: int = 4
    threads_per_problem: int = 0
    """Kernel choice (no reference counterpart): 0 library default / autotuned, 1 one lane per problem,
    -64 one wavefront per problem, -32 two problems per wavefront, 4..64 lanes per problem staged in LDS."""

    def build(self, terms: Sequence[Task], device=None, active_dof_mask=None) -> "LMOptimizer":
        return LMOptimizer(terms, self, device, active_dof_mask)

    def params(self) -> N.LMParams:
        return N.LMParams(self.lm_lambda, self.lambda_factor, self.lambda_min, self.lambda_max, self.rho_min,
                          self.gain_ratio_epsilon)


class LMOptimizer:
    """Levenberg-Marquardt over a batch of independent problems, one fused GPU kernel per step."""

    _config_type = LMOptimizerConfig

    def __init__(self, terms: Union[Task, Sequence[Task]], config: Optional[LMOptimizerConfig] = None,
                 device=None, active_dof_mask=None):
        self.terms: List[Task] = list(terms) if isinstance(terms, (list, tuple)) else [terms]
        self.config = config if config is not None else LMOptimizerConfig()
        if self.config.linear_solver != "cholesky":
            raise NotImplementedError("linear_solver must be 'cholesky' on this path")
        if active_dof_mask is not None:
            raise NotImplementedError("active_dof_mask is not supported on this path")
        self.device = device
        self._ctx = get_context(device)
        self._plan()
        self._batch: Optional[N.Batch] = None
        self.max_batch_size = 0
        self.batch_size = 0
        self.costs: Optional[np.ndarray] = None
        self._iter = 0
        self._task_stop = False
        self._var: Optional[VarValues] = None

    # -- term structure
    def _plan(self) -> None:
        if any(t.is_template for t in self.terms):
            raise ValueError("a term template (no robot, or Context placeholders) must be resolved first")
        robots = {id(t.robot) for t in self.terms}
        if len(robots) != 1:
            raise ValueError("all terms of one optimizer must act on the same robot")
        self.robot: Robot = self.terms[0].robot
        self._frame = [t for t in self.terms if isinstance(t, FrameTask)]
        self._pos = [t for t in self.terms if isinstance(t, PositionTask)]
        self._rot = [t for t in self.terms if isinstance(t, RotationTask)]
        lim = [t for t in self.terms if isinstance(t, PositionLimit)]
        sm = [t for t in self.terms if isinstance(t, SmoothnessTask)]
        col = [t for t in self.terms if isinstance(t, SceneCollisionTask)]
        other = [t for t in self.terms if not isinstance(t, (FrameTask, PositionTask, RotationTask, PositionLimit,
                                                             SmoothnessTask, SceneCollisionTask))]
        if other:
            raise NotImplementedError(f"terms not on this path: {[type(t).__name__ for t in other]}")
        if len(lim) > 1 or len(sm) > 1 or len(col) > 1:
            raise ValueError("at most one PositionLimit, one SmoothnessTask and one SceneCollisionTask per optimizer")
        self._collision = col[0] if col else None
        self._limit = lim[0] if lim else None
        self._smooth = sm[0] if sm else None
        pose_links = [f for t in self._frame for f in t.frame_indices]
        pos_links = [f for t in self._pos for f in t.frame_indices]
        rot_links = [f for t in self._rot for f in t.frame_indices]
        if len(pose_links) + len(rot_links) > MAX_POSE_TARGETS or len(pos_links) > MAX_POSITION_TARGETS:
            raise ValueError(f"at most {MAX_POSE_TARGETS} pose + rotation and {MAX_POSITION_TARGETS} position targets")
        self.termspec = N.TermSpec(
            pose_links=pose_links,
            pose_weights=np.concatenate([t.weights for t in self._frame]) if self._frame else None,
            pos_links=pos_links,
            pos_weights=np.concatenate([t.weights for t in self._pos]) if self._pos else None,
            limit_weights=self._limit.weights if self._limit else None,
            smooth_weights=self._smooth.weights if self._smooth else None,
            rot_links=rot_links,
            rot_weights=np.concatenate([t.weights for t in self._rot]) if self._rot else None,
        )
        if self._collision is not None:
            c = self._collision
            self.termspec.coll_links = list(c.sphere_links)
            self.termspec.coll_centers, self.termspec.coll_radii = c.sphere_centers, c.sphere_radii
            self.termspec.coll_weights, self.termspec.coll_margin = c.weights, c.margin
            self.termspec.coll_penalty = c.penalty
            self.termspec.obstacle_spheres, self.termspec.obstacle_boxes = c.scene.spheres, c.scene.boxes

    @property
    def residual_dim(self) -> int:
        return sum(t.residual_dim for t in self.terms)

    def _gather_targets(self, B: int):
        def stack(tasks, w):
            if not tasks:
                return None
            parts = []
            for t in tasks:
                tg = t.T_world_target
                if tg.shape[0] not in (1, B):
                    raise ValueError(f"{type(t).__name__} has {tg.shape[0]} targets for a batch of {B}")
                parts.append(np.broadcast_to(tg, (B,) + tg.shape[1:]))
            return np.ascontiguousarray(np.concatenate(parts, axis=1), dtype=np.float32)

        self._rot_targets = stack(self._rot, 7)
        q_prev = None
        if self._smooth is not None:
            if self._smooth.prev_var is None:
                raise ValueError("SmoothnessTask needs prev_var")
            q_prev = np.broadcast_to(self._smooth.prev_var.q, (B, self.robot.spec.num_actuated_joints))
        return stack(self._frame, 7), stack(self._pos, 3), q_prev

    # -- lifecycle (opt/optimizer.py:216-333)
    def reset(self, var: VarValues, num_steps: Optional[int] = None) -> None:
        B = var.batch_size
        state = var.get(VarKey.ROBOT)
        if state.robot.spec is not self.robot.spec and state.tangent_dim != self.robot.spec.num_actuated_joints:
            raise ValueError("var does not match the optimizer's robot")
        if self._batch is None:
            self._batch = N.Batch(self._ctx, self.robot.spec, self.termspec, B)
            self.max_batch_size = B
        if B > self.max_batch_size:
            raise ValueError(f"batch size {B} exceeds the {self.max_batch_size} this optimizer was built for")
        pose, pos, q_prev = self._gather_targets(B)
        self._batch.set(q=state.q, pose_targets=pose, pos_targets=pos, q_prev=q_prev, n=B,
                        rot_targets=self._rot_targets)
        self.batch_size = B
        self._num_steps = self.config.max_iter if num_steps is None else num_steps
        self._var = var
        self._iter = 0
        self._task_stop = False
        self.costs = None

    def step(self, var: VarValues) -> None:
        """One LM iteration in place (continues the device state of the previous step)."""
        self._run(var, 1)

    def _run(self, var: VarValues, iters: int) -> None:
        if var is not self._var:
            raise ValueError("step() must follow reset() on the same VarValues")
        self._batch.solve(iters, self.config.params(), tpp=self.config.threads_per_problem, cont=self._iter > 0)
        self._iter += iters
        out = self._batch.get()
        var.get(VarKey.ROBOT).q[...] = out["q"]
        self.costs = out["cost"]

    def solve(self, var: VarValues, iteration_callback: Optional[IterationCallback] = None
              ) -> Tuple[VarValues, np.ndarray]:
        """Optimise ``var`` in place; returns it with the final cost per batch element."""
        cfg = self.config
        if cfg.cuda_graph_mode == "full" and iteration_callback is not None:
            raise ValueError("iteration_callback is incompatible with full-solve CUDA graph capture")
        self.reset(var)
        if iteration_callback is not None:
            # the cost at the start (the reference's step-0 callback sees the seed cost)
            self._batch.solve(0, cfg.params(), tpp=cfg.threads_per_problem)
            iteration_callback(0, var, self._batch.get()["cost"])
        interval = cfg.early_stopping_interval
        if iteration_callback is None and (interval <= 0 or cfg.cuda_graph_mode == "full"):
            self._run(var, cfg.max_iter)
            return var, self.costs
        chunk = 1 if iteration_callback is not None else interval
        done = 0
        while done < cfg.max_iter:
            n = min(chunk, cfg.max_iter - done)
            self._run(var, n)
            done += n
            if iteration_callback is not None:
                iteration_callback(done, var, self.costs)
            if interval > 0 and done % interval == 0 and bool(np.all(self.costs <= cfg.cost_tol)):
                break
        return var, self.costs


def ik_metrics(robot: Robot, q: np.ndarray, frame_indices: Sequence[int], T_world_target: np.ndarray,
               pos_threshold: float = 1e-3, ang_threshold: float = 1e-2, device=None):
    """Per-target position / angle errors and success (helpers/ik/ik.py:59-82)."""
    frames = _frames(frame_indices)
    tg = _targets(T_world_target, len(frames), 7)
    q = np.atleast_2d(np.asarray(q, np.float32))
    b = N.Batch(get_context(device), robot.spec,
                N.TermSpec(pose_links=frames, pose_weights=np.ones((len(frames), 6), np.float32)), q.shape[0])
    try:
        b.set(q=q, pose_targets=np.broadcast_to(tg, (q.shape[0],) + tg.shape[1:]))
        return b.metrics(pos_threshold, ang_threshold)
    finally:
        b.close()

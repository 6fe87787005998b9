"""RoboKit's LM solver and residual-term API, on the MI355X library.

The classes mirror the reference's names, arguments, defaults and error behaviour so that a
caller of the reference's Levenberg-Marquardt path can switch without edits:

  VarKey, VarValues, VectorVar opt/var_values.py:24-342 (column blocks per key; the terms read
                               VarKey.ROBOT, other keys are carried through a solve unchanged)
  Robot, RobotState            robo/robot.py:105-130, robo/robot_state.py:41-476 (q, batch)
  FrameTask                    terms/dense/frame_task.py:72-82 (residual_mode "log")
  PositionTask                 terms/dense/position_task.py:42-50
  RotationTask                 terms/dense/rotation_task.py:25-80 (axis-angle rows, geometric Jacobian)
  PositionLimit                terms/dense/position_limit.py:204-220 (residual_mode "abs")
  SmoothnessTask               terms/dense/smoothness_task.py:41-50
  SceneCollisionTask           terms/dense/scene_collision_task.py:276-356 (robot spheres against a
                               PrimitiveScene of obstacle spheres and axis-aligned boxes, geom/scene.py)
  TrajectoryTask               terms/trajectory_task.py:65-135 (a FrameTask goal on the last frame)
  TrajectorySmoothnessTask     terms/sparse/trajectory_smoothness_task.py:58-120 (orders 1, 2)
  TrajectoryCollisionTask      terms/sparse/trajectory_collision_task.py:848-900 (penalty "plain")
  TrajectoryLimitTask          terms/sparse/trajectory_limit_task.py:29-160 (stencil "forward", orders 0-2)
  LMOptimizerConfig            opt/lm_optimizer.py:502-530 (+ OptimizerConfig, opt/optimizer.py:159-174;
                               linear_solver "cg" on trajectories, the banded Cholesky otherwise)
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
    """Variable keys (opt/var_values.py:73-85).  The terms of this path read ROBOT; a VarValues may
    carry the others (and keys of a caller's own Enum) beside it."""

    ROBOT = 0
    TIMING = 1
    ACTION = 2
    CONTACTS = 3
    QPOS_OFFSET = 4
    MOUNT_OFFSET = 5
    STATE = 6
    SEGMENT_STEP = 7


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
    """Joint positions of a batch, q [B, na] float32, or of a batch of trajectories, q [B, T, na]
    (robo/robot_state.py:41-181; a trajectory's tangent is its T * na joint values)."""

    def __init__(self, robot: Robot, q: np.ndarray):
        q = np.array(q, dtype=np.float32)  # a copy: solves write the state in place
        if q.ndim == 1:
            q = q[None]
        if q.ndim not in (2, 3) or q.shape[-1] != robot.spec.num_actuated_joints:
            raise ValueError(f"q must be [B, {robot.spec.num_actuated_joints}] or [B, T, "
                             f"{robot.spec.num_actuated_joints}], got {q.shape}")
        self.robot = robot
        self.q = np.ascontiguousarray(q)

    @property
    def batch_size(self) -> int:
        return self.q.shape[0]

    @property
    def num_frames(self) -> int:
        """trajectory frames (1 for a configuration batch)"""
        return self.q.shape[1] if self.q.ndim == 3 else 1

    @property
    def tangent_dim(self) -> int:
        return int(np.prod(self.q.shape[1:]))

    def view(self, batch_size: int, offset: int = 0) -> "RobotState":
        """elements offset .. offset + batch_size, sharing storage (opt/var_values.py:33-36)"""
        if offset < 0 or offset + batch_size > self.batch_size:
            raise ValueError(f"view [{offset}, {offset + batch_size}) outside a batch of {self.batch_size}")
        v = RobotState.__new__(RobotState)
        v.robot, v.q = self.robot, self.q[offset:offset + batch_size]
        return v

    def clone(self) -> "RobotState":
        return RobotState(self.robot, self.q.copy())


    # host forms of the Var protocol (opt/var_values.py:24-70; robo/robot_state.py:238-333, 448-476).
    # The fused kernels integrate and accept on the device; these serve callers that compose
    # variables on the host.
    def integrate(self, velocity: np.ndarray, out: Optional["RobotState"] = None,
                  tangent_mask: Optional[np.ndarray] = None, weight_decay: float = 0.0) -> "RobotState":
        """q + mask * (velocity - weight_decay * q), no clamp (robot_state.py:59, 238-312)"""
        v = np.asarray(velocity, np.float32).reshape(self.q.shape)
        m = np.float32(1.0) if tangent_mask is None else np.asarray(tangent_mask, np.float32).reshape(self.q.shape[1:])
        q = self.q + m * (v - np.float32(weight_decay) * self.q)
        if out is None:
            return RobotState(self.robot, q)
        out.q[...] = q
        return out

    def difference(self, other: "RobotState") -> np.ndarray:
        """[B, tangent_dim] velocities that integrate moves each element by to reach ``other``"""
        return (other.q - self.q).reshape(self.batch_size, -1)

    def gather(self, indices, out: Optional["RobotState"] = None) -> "RobotState":
        q = self.q[np.asarray(indices, np.int64)]
        if out is None:
            return RobotState(self.robot, q)
        out.q[...] = q
        return out

    def accept(self, accept_mask, proposed: "RobotState") -> "RobotState":
        """rows where accept_mask[i] == 1 take the proposal (robot_kernels.py:743-767)"""
        m = np.asarray(accept_mask).reshape(-1).astype(bool)
        self.q[m] = proposed.q[m]
        return self


class VectorVar:
    """A batched Euclidean vector variable x [B, d] float32 (opt/var_values.py:251-342): a move is
    ``unit`` times the velocity, at most ``max_step`` per entry when that is positive."""

    def __init__(self, x: np.ndarray, unit: float = 1.0, max_step: float = 0.0):
        x = np.array(x, dtype=np.float32)
        if x.ndim != 2:
            raise ValueError("VectorVar requires a 2D float32 array")
        self.x, self.unit, self.max_step = np.ascontiguousarray(x), float(unit), float(max_step)

    @property
    def tangent_dim(self) -> int:
        return int(self.x.shape[1])

    @property
    def batch_size(self) -> int:
        return int(self.x.shape[0])

    def clone(self) -> "VectorVar":
        return VectorVar(self.x.copy(), self.unit, self.max_step)

    def view(self, batch_size: int, offset: int = 0) -> "VectorVar":
        v = VectorVar.__new__(VectorVar)
        v.x, v.unit, v.max_step = self.x[offset:offset + batch_size], self.unit, self.max_step
        return v

    def integrate(self, velocity: np.ndarray, out: Optional["VectorVar"] = None,
                  tangent_mask: Optional[np.ndarray] = None, weight_decay: float = 0.0) -> "VectorVar":
        v = np.asarray(velocity, np.float32)
        if v.shape != self.x.shape:
            raise ValueError(f"Expected velocity shape {self.x.shape}, but got {v.shape}")
        m = np.float32(1.0) if tangent_mask is None else np.asarray(tangent_mask, np.float32)
        move = np.float32(self.unit) * (m * (v - np.float32(weight_decay) * self.x))
        if self.max_step > 0:
            move = np.clip(move, -np.float32(self.max_step), np.float32(self.max_step))
        if out is None:
            out = VectorVar(self.x, self.unit, self.max_step)
        out.x[...] = self.x + move
        return out

    def difference(self, other: "VectorVar") -> np.ndarray:
        return other.x - self.x

    def gather(self, indices, out: Optional["VectorVar"] = None) -> "VectorVar":
        x = self.x[np.asarray(indices, np.int64)]
        if out is None:
            return VectorVar(x, self.unit, self.max_step)
        out.x[...] = x
        return out

    def accept(self, accept_mask, proposed: "VectorVar") -> "VectorVar":
        m = np.asarray(accept_mask).reshape(-1).astype(bool)
        self.x[m] = proposed.x[m]
        return self


class VarValues:
    """Variables keyed by an Enum, combined into tangent-column blocks in insertion order
    (opt/var_values.py:88-248).  The LM path optimises the key its terms read (VarKey.ROBOT);
    a variable that no term reads has all-zero Jacobian columns, so its damped step is exactly
    zero and it is carried through a solve unchanged, as in the reference."""

    def __init__(self, named_vars: Dict[Enum, Union[RobotState, VectorVar]]):
        if not named_vars:
            raise ValueError("VarValues requires at least one variable")
        self._values = dict(named_vars)
        self._offsets: Dict[Enum, int] = {}
        off = 0
        for k, v in self._values.items():
            self._offsets[k] = off
            off += v.tangent_dim
        self._tangent_dim = off
        self._batch_size = next(iter(self._values.values())).batch_size
        if any(v.batch_size != self._batch_size for v in self._values.values()):
            raise ValueError("every variable of a VarValues needs the same batch size")

    @classmethod
    def from_var(cls, var, key: Enum) -> "VarValues":
        return cls({key: var})

    def get(self, key: Enum):
        return self._values[key]

    def __getitem__(self, key: Enum):
        return self._values[key]

    def __contains__(self, key: Enum) -> bool:
        return key in self._values

    def values(self):
        return self._values.values()

    def keys(self):
        return self._values.keys()

    def tangent_offset(self, key: Enum) -> int:
        return self._offsets[key]

    @property
    def batch_size(self) -> int:
        return self._batch_size

    @property
    def tangent_dim(self) -> int:
        return self._tangent_dim

    def clone(self) -> "VarValues":
        return type(self)({k: v.clone() for k, v in self._values.items()})

    def view(self, batch_size: int, offset: int = 0) -> "VarValues":
        return type(self)({k: v.view(batch_size, offset) for k, v in self._values.items()})

    def _block(self, key: Enum, a: np.ndarray) -> np.ndarray:
        o, d = self._offsets[key], self._values[key].tangent_dim
        return a[:, o:o + d]

    def integrate(self, velocity: np.ndarray, out: Optional["VarValues"] = None,
                  tangent_mask: Optional[np.ndarray] = None, weight_decay: float = 0.0) -> "VarValues":
        """each variable moves by its own column block of velocity [B, tangent_dim]"""
        v = np.asarray(velocity, np.float32)
        if v.shape != (self.batch_size, self.tangent_dim):
            raise ValueError(f"Expected velocity shape {(self.batch_size, self.tangent_dim)}, got {v.shape}")
        res = {}
        for k, var in self._values.items():
            o, d = self._offsets[k], var.tangent_dim
            m = None if tangent_mask is None else np.asarray(tangent_mask)[o:o + d]
            res[k] = var.integrate(self._block(k, v).reshape((self.batch_size,) + _var_shape(var)),
                                   out=None if out is None else out.get(k), tangent_mask=m, weight_decay=weight_decay)
        return out if out is not None else type(self)(res)

    def difference(self, other: "VarValues") -> np.ndarray:
        out = np.zeros((self.batch_size, self.tangent_dim), np.float32)
        for k, var in self._values.items():
            self._block(k, out)[...] = var.difference(other.get(k)).reshape(self.batch_size, -1)
        return out

    def gather(self, indices, out: Optional["VarValues"] = None) -> "VarValues":
        if out is None:
            return type(self)({k: v.gather(indices) for k, v in self._values.items()})
        for k, v in self._values.items():
            v.gather(indices, out=out.get(k))
        return out

    def accept(self, accept_mask, proposed: "VarValues") -> "VarValues":
        for k, v in self._values.items():
            v.accept(accept_mask, proposed.get(k))
        return self


def _var_shape(var) -> tuple:
    return var.q.shape[1:] if isinstance(var, RobotState) else var.x.shape[1:]


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
    """A world scene of obstacle spheres and axis-aligned boxes, then up to four signed-distance
    volumes / triangle meshes (robokit_amd.geom.SdfVolume / TriMesh): the reference's WarpScene
    with SphereGeom / BoxGeom / VolumeGeom / MeshGeom members (geom/scene.py:132-256,
    geom/geoms.py:372-660), queried by its signed distance (geom/sdf_kernels.py:151-380)."""

    def __init__(self, spheres: Optional[np.ndarray] = None, boxes: Optional[np.ndarray] = None, geoms=()):
        from .geom import as_geom

        self.spheres = np.zeros((0, 4), np.float32) if spheres is None else np.asarray(spheres, np.float32).reshape(-1, 4)
        self.boxes = np.zeros((0, 6), np.float32) if boxes is None else np.asarray(boxes, np.float32).reshape(-1, 6)
        self.geoms = [as_geom(g) for g in geoms]
        if len(self.geoms) > 4:
            raise ValueError("at most 4 volumes / meshes per scene")

    def add_geometry(self, geom) -> "PrimitiveScene":
        """append a volume (SdfVolume), a mesh (TriMesh) or a (vertices, faces) pair"""
        from .geom import as_geom

        if len(self.geoms) >= 4:
            raise ValueError("at most 4 volumes / meshes per scene")
        self.geoms.append(as_geom(geom))
        return self

    def query_sdf(self, points: np.ndarray, device=None):
        """(signed distance [n], normal [n, 3]) at points [n, 3] on the GPU (WarpScene.query_sdf)"""
        return get_context(device).scene_query(points, self.spheres, self.boxes, self.geoms)

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
    def num_geoms(self) -> int:
        return len(self.spheres) + len(self.boxes) + len(self.geoms)


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


# ----------------------------------------------------------------------------- trajectory terms
class TrajectoryTask(Task):
    """A per-pose task applied to trajectory frames (terms/trajectory_task.py:65-135).  The
    trajectory kernel takes the goal form: a FrameTask on one link, on the last frame."""

    def __init__(self, dense_task: Task, num_frames: Optional[int] = None,
                 frame_indices: Union[Sequence[int], Context, None] = None, name: str = ""):
        self._dense_task, self.name = dense_task, name
        self.frames = frame_indices
        self.num_frames = num_frames
        if num_frames is not None and not isinstance(frame_indices, Context):
            self.frames = list(range(num_frames)) if frame_indices is None else [int(f) % num_frames for f in frame_indices]
        self.robot = getattr(dense_task, "robot", None)
        if dense_task.is_template or num_frames is None or isinstance(frame_indices, Context):
            self._template = {"dense_task": dense_task, "num_frames": num_frames, "frame_indices": frame_indices,
                              "name": name}
            return
        if not isinstance(dense_task, FrameTask) or len(dense_task.frame_indices) != 1:
            raise NotImplementedError("trajectory tasks on this path: a FrameTask on one link (the goal pose)")
        if self.frames != [num_frames - 1]:
            raise NotImplementedError("the trajectory kernel applies the FrameTask to the last frame only "
                                      "(frame_indices=[-1])")

    def resolve(self, robot: "Robot", context: Dict["Context", object]) -> "Task":
        if self._template is None:
            return self
        t = self._template
        dense = t["dense_task"].resolve(robot, context)
        fr = context[t["frame_indices"]] if isinstance(t["frame_indices"], Context) else t["frame_indices"]
        return TrajectoryTask(dense, t["num_frames"], fr, t["name"])

    @property
    def dense_task(self) -> Task:
        return self._dense_task

    @property
    def goal_link(self) -> int:
        return self._dense_task.frame_indices[0]

    @property
    def residual_dim(self) -> int:
        return self._dense_task.residual_dim * len(self.frames)


def _scalar_weight(w, name: str) -> float:
    a = np.asarray(w, np.float64).reshape(-1)
    if a.size == 0 or np.any(a != a[0]):
        raise NotImplementedError(f"{name}: the trajectory kernel takes one weight for every dof")
    return float(a[0])


class TrajectorySmoothnessTask(Task):
    """Finite differences of order 1 (velocity) or 2 (acceleration) over the frames,
    r = w / dt^order * sum_j (-1)^(order-j) C(order, j) q[t+j]
    (terms/sparse/trajectory_smoothness_task.py:58-120, 506-624)."""

    def __init__(self, robot: Optional[Robot] = None, num_frames: Optional[int] = None,
                 weight: Union[float, Sequence[float], np.ndarray] = 1.0, order: int = 1, dt: float = 1.0,
                 reference_q=None, base_weight=None, smooth_boundary: bool = True, virtual_hold_frames: int = 0,
                 history=None, name: str = ""):
        if order < 1 or (num_frames is not None and order >= num_frames):
            raise ValueError(f"order must be positive and smaller than num_frames={num_frames}, got {order}")
        if base_weight is not None and order != 1:
            raise ValueError("base_weight is only supported for order=1.")
        if robot is None or num_frames is None or _has_context(dt):
            self._as_template(num_frames=num_frames, weight=weight, order=order, dt=dt, name=name)
            return
        if order > 2:
            raise NotImplementedError("the trajectory kernel has velocity and acceleration rows (order 1, 2)")
        if reference_q is not None or base_weight is not None or virtual_hold_frames or history is not None \
                or not smooth_boundary:
            raise NotImplementedError("reference_q, base_weight, virtual hold frames and history are not on this path")
        if not dt > 0:
            raise ValueError("dt must be positive")
        self.robot, self.name, self.num_frames, self.order, self.dt = robot, name, int(num_frames), int(order), float(dt)
        self.weight = _scalar_weight(weight, "TrajectorySmoothnessTask weight")

    @property
    def residual_dim(self) -> int:
        return (self.num_frames - self.order) * self.robot.spec.num_actuated_joints


class TrajectoryCollisionTask(Task):
    """Robot spheres against a primitive scene at every frame: r = w max(margin - sd, 0)
    (penalty "plain", terms/sparse/trajectory_collision_task.py:848-1100, sphere_collision_kernels.py:24-27)."""

    def __init__(self, robot: Optional[Robot] = None, scene: Optional[PrimitiveScene] = None,
                 num_frames: Optional[int] = None, weight: Union[float, Sequence[float]] = 20.0,
                 margin: Union[float, Sequence[float]] = 0.1, max_dist: float = 1e6, sphere_indices=None,
                 scene_indices=None, active_actuated_indices=None, sweep_steps: int = 0,
                 penalty: Literal["smooth", "plain", "surface_distance"] = "plain", geometry=None,
                 frame_sphere_mask=None, speed_dt=None, free_goal_frame: bool = False, name: str = "",
                 goal_radius: float = 0.0, goal_distance=None):
        if robot is None or scene is None or num_frames is None or _has_context(scene):
            self._as_template(scene=scene, num_frames=num_frames, weight=weight, margin=margin,
                              sphere_indices=sphere_indices, penalty=penalty, name=name)
            return
        if penalty != "plain":
            raise NotImplementedError("the trajectory kernel's collision rows use the plain hinge")
        if sweep_steps or scene_indices is not None or active_actuated_indices is not None or geometry is not None \
                or frame_sphere_mask is not None or speed_dt is not None or free_goal_frame or goal_radius \
                or goal_distance is not None:
            raise NotImplementedError("swept spheres, per-query scenes, frame masks and goal tapering are not on "
                                      "this path")
        if not robot.has_collision_spheres:
            raise RuntimeError("Robot has no collision spheres. Build it with collision_spheres=(links, centers, radii).")
        self.robot, self.name, self.scene, self.num_frames = robot, name, scene, int(num_frames)
        links, centers, radii = robot.collision_spheres
        idx = np.arange(len(links)) if sphere_indices is None else np.asarray(sphere_indices, np.int64).reshape(-1)
        if np.any(idx < 0) or np.any(idx >= len(links)):
            raise ValueError("sphere_indices contains out-of-range indices.")
        self.sphere_links, self.sphere_centers, self.sphere_radii = links[idx], centers[idx], radii[idx]
        self.weight = _scalar_weight(weight, "TrajectoryCollisionTask weight")
        self.margin = _scalar_weight(margin, "TrajectoryCollisionTask margin")

    @property
    def residual_dim(self) -> int:
        return self.num_frames * len(self.sphere_links)


class TrajectoryLimitTask(Task):
    """Bounds on the joint values (order 0), velocities (1) and accelerations (2) of every window of
    order + 1 frames: r = w max(lower - v, v - upper, 0), v the forward difference / dt^order
    (terms/sparse/trajectory_limit_task.py:29-537, stencil "forward")."""

    def __init__(self, robot: Optional[Robot] = None, num_frames: Optional[int] = None,
                 order: Union[int, Sequence[int]] = 1, lower=None, upper=None,
                 weight: Union[float, Sequence[float], np.ndarray] = 1.0, dt: float = 1.0,
                 stencil: Literal["central", "forward"] = "central", dt_power=None,
                 activation: Sequence[float] = (0.0, 0.0, 0.0, 0.0), normalize: bool = False, name: str = ""):
        if robot is None or num_frames is None or _has_context(upper, dt):
            self._as_template(num_frames=num_frames, order=order, lower=lower, upper=upper, weight=weight, dt=dt,
                              stencil=stencil, name=name)
            return
        if stencil != "forward":
            raise NotImplementedError("the trajectory kernel's bound rows use forward stencils (stencil='forward')")
        if dt_power is not None or any(activation) or normalize:
            raise NotImplementedError("dt_power, activation and normalize are not on this path")
        orders = [int(order)] if np.isscalar(order) else [int(o) for o in order]
        if not orders or any(o not in (0, 1, 2) for o in orders) or len(set(orders)) != len(orders):
            raise NotImplementedError("bound rows of orders 0 (position), 1 (velocity), 2 (acceleration)")
        if not dt > 0:
            raise ValueError("dt must be positive")
        na, k = robot.spec.num_actuated_joints, len(orders)
        lim = np.asarray(robot.spec.actuated_joint_limits, np.float64)
        vel = getattr(robot.spec, "actuated_joint_velocity_limits", None)

        def per(v, what):  # bounds / weights broadcast over the joints, a leading order axis if several
            a = np.asarray(v, np.float64)
            if a.ndim == 0 or a.shape == (na,):
                return np.broadcast_to(a, (k, na)).copy()
            if a.shape == (k,):
                return np.repeat(a.reshape(k, 1), na, 1)
            if a.shape == (k, na):
                return a.copy()
            raise ValueError(f"{what} must broadcast to [orders {k}, dofs {na}], got {a.shape}")

        if upper is None:  # the robot's position and velocity limits (trajectory_limit_task.py:136-140)
            rows = []
            for o in orders:
                if o == 0:
                    rows.append(lim[:, 1])
                elif o == 1 and vel is not None:
                    rows.append(np.asarray(vel, np.float64))
                else:
                    raise ValueError(f"TrajectoryLimitTask needs upper for order {o}")
            up = np.stack(rows)
        else:
            up = per(np.stack([np.broadcast_to(np.asarray(r, np.float64), (na,)) for r in upper])
                     if isinstance(upper, tuple) else upper, "upper")
        lo = (np.stack([lim[:, 0] if o == 0 else -up[i] for i, o in enumerate(orders)]) if lower is None
              else per(lower, "lower"))
        w = per(weight, "weight")
        lo, up, w = lo.astype(np.float32), up.astype(np.float32), w.astype(np.float32)
        if not np.all(lo <= up):
            raise ValueError("lower and upper must be ordered, non-NaN bounds")
        self.robot, self.name, self.num_frames, self.orders, self.dt = robot, name, int(num_frames), orders, float(dt)
        self.lower, self.upper, self.weights = lo, up, w

    @property
    def residual_dim(self) -> int:
        return sum(self.num_frames - o for o in self.orders) * self.robot.spec.num_actuated_joints


_TRAJ_TERMS = (TrajectoryTask, TrajectorySmoothnessTask, TrajectoryCollisionTask, TrajectoryLimitTask)


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
        if self.config.linear_solver not in ("cholesky", "cg"):
            raise NotImplementedError("linear_solver must be 'cholesky' or 'cg' on this path")
        if active_dof_mask is not None:
            raise NotImplementedError("active_dof_mask is not supported on this path")
        self.device = device
        _device_ordinal(device)  # an unsupported device fails here; the context opens at the first reset
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
        self._traj = None
        if any(isinstance(t, _TRAJ_TERMS) for t in self.terms):
            self._plan_trajectory()
            return
        if self.config.linear_solver != "cholesky":
            raise NotImplementedError("linear_solver 'cg' is on the trajectory path; dense problems use 'cholesky'")
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
            self.termspec.geoms = list(c.scene.geoms)

    def _plan_trajectory(self) -> None:
        """the trajectory problem set of the terms: goal FrameTask on the last frame, velocity /
        acceleration differences, the plain collision hinge and the bound rows, one dt shared by
        every term that has one (the trajectory kernel's descriptor, rk_traj_desc)"""
        from .workloads import TrajBatch, TrajScene  # the problem-set containers of the trajectory binding

        other = [t for t in self.terms if not isinstance(t, _TRAJ_TERMS)]
        if other:
            raise NotImplementedError(f"trajectory problems take trajectory terms only, got {[type(t).__name__ for t in other]}")
        T = {t.num_frames for t in self.terms}
        if len(T) != 1:
            raise ValueError(f"the trajectory terms disagree on num_frames: {sorted(T)}")
        T = T.pop()
        goal = [t for t in self.terms if isinstance(t, TrajectoryTask)]
        smooth = [t for t in self.terms if isinstance(t, TrajectorySmoothnessTask)]
        coll = [t for t in self.terms if isinstance(t, TrajectoryCollisionTask)]
        lims = [t for t in self.terms if isinstance(t, TrajectoryLimitTask)]
        if len(goal) > 1 or len(coll) > 1:
            raise NotImplementedError("one goal TrajectoryTask and one TrajectoryCollisionTask per trajectory optimizer")
        wv = wa = 0.0
        for t in smooth:
            if (t.order == 1 and wv) or (t.order == 2 and wa):
                raise ValueError(f"two TrajectorySmoothnessTasks of order {t.order}")
            if t.order == 1:
                wv = t.weight
            else:
                wa = t.weight
        dts = {t.dt for t in smooth + lims}
        if len(dts) > 1:
            raise NotImplementedError(f"the trajectory terms use one dt, got {sorted(dts)}")
        dt = dts.pop() if dts else 1.0
        spec = self.robot.spec
        na = spec.num_actuated_joints
        lim_orders, lim_w, lim_lo, lim_hi = 0, None, None, None
        if lims:
            lim_w, lim_lo, lim_hi = (np.zeros((3, na), np.float32) for _ in range(3))
            for t in lims:
                for i, o in enumerate(t.orders):
                    if lim_orders >> o & 1:
                        raise ValueError(f"two TrajectoryLimitTask rows of order {o}")
                    lim_orders |= 1 << o
                    lim_w[o], lim_lo[o], lim_hi[o] = t.weights[i], t.lower[i], t.upper[i]
        if coll:
            c = coll[0]
            scene = TrajScene(c.sphere_links.astype(np.int32), c.sphere_centers, c.sphere_radii,
                              c.scene.spheres, c.scene.boxes, list(c.scene.geoms))
        else:
            scene = TrajScene(np.zeros(0, np.int32), np.zeros((0, 3), np.float32), np.zeros(0, np.float32),
                              np.zeros((0, 4), np.float32), np.zeros((0, 6), np.float32))
        self._goal = goal[0] if goal else None
        D = T * na
        cg = 0
        if self.config.linear_solver == "cg":  # the reference's default budget, opt/lm_optimizer.py:577
            cg = self.config.cg_max_iter if self.config.cg_max_iter is not None else max(4, min(64, D // 4))
        self._traj = TrajBatch(spec, self._goal.goal_link if self._goal else spec.num_links - 1, T,
                               np.zeros((0, T, na), np.float32), np.zeros((0, 7), np.float32), scene, wv=wv, wa=wa,
                               dt=dt, coll_w=coll[0].weight if coll else 0.0, margin=coll[0].margin if coll else 0.0,
                               lim_orders=lim_orders, lim_w=lim_w, lim_lo=lim_lo, lim_hi=lim_hi, cg_iters=int(cg))
        self._goal_weights = (self._goal.dense_task.weights[0] if self._goal else np.zeros(6, np.float32)).astype(np.float32)

    @property
    def _ctx(self) -> N.Context:
        return get_context(self.device)

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
        if self._traj is not None:
            self._reset_trajectory(var, state, B, num_steps)
            return
        if state.q.ndim != 2:
            raise ValueError("a configuration batch needs q [B, na]; trajectories need trajectory terms")
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

    def _reset_trajectory(self, var: VarValues, state: RobotState, B: int, num_steps: Optional[int]) -> None:
        tb = self._traj
        if state.q.ndim != 3 or state.q.shape[1:] != (tb.T, self.robot.spec.num_actuated_joints):
            raise ValueError(f"trajectory terms need q [B, {tb.T}, {self.robot.spec.num_actuated_joints}], "
                             f"got {state.q.shape}")
        if self._batch is None:
            self._batch = N.TrajBatch(self._ctx, tb, B, goal_weights=self._goal_weights)
            self.max_batch_size = B
        if B > self.max_batch_size:
            raise ValueError(f"batch size {B} exceeds the {self.max_batch_size} this optimizer was built for")
        if self._goal is not None:
            g = self._goal.dense_task.T_world_target.reshape(-1, 7)
            if g.shape[0] not in (1, B):
                raise ValueError(f"TrajectoryTask has {g.shape[0]} goals for a batch of {B}")
            goals = np.broadcast_to(g, (B, 7))
        else:
            goals = np.broadcast_to(np.array([0, 0, 0, 1, 0, 0, 0], np.float32), (B, 7))
        self._batch.set(q=state.q, goals=goals, n=B)
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
        tpp = 0 if self._traj is not None else self.config.threads_per_problem
        self._batch.solve(iters, self.config.params(), tpp=tpp, cont=self._iter > 0)
        self._iter += iters
        out = self._batch.get()
        var.get(VarKey.ROBOT).q[...] = out["q"].reshape(var.get(VarKey.ROBOT).q.shape)
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
            self._batch.solve(0, cfg.params(), tpp=0 if self._traj is not None else cfg.threads_per_problem)
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

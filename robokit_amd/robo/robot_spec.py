"""Robot topology arrays: the kinematic-tree description every kernel reads.

Mirrors ``RobotSpec`` (robo/robot_spec.py:45-138) field for field for the parts the LM
path reads: joint types/axes/twists, parent joint indices and transforms, actuated and
mimic maps, limits, link -> parent joint, plus the derived ``joints_to_actuated_mapping``
(robot_spec.py:1120-1132), ``link_ancestor_joints_mask`` (:1134-1143) and
``topological_order_joint_indices`` (:1044-1069).

yourdfpy is not available, so ``parse_urdf`` reads URDF XML itself with the same
ordering rules the reference gets from yourdfpy (robot_spec.py:284-618): joints and links
in file order, actuated joints = non-fixed joints without ``<mimic>`` in joint order.
"""
from __future__ import annotations

import math
import xml.etree.ElementTree as ET
from dataclasses import dataclass, field
from enum import IntEnum
from pathlib import Path
from typing import List, Optional, Sequence, Tuple, Union

import numpy as np


class JointType(IntEnum):
    """robot_spec.py:35-41."""

    FIXED = 0
    PRISMATIC = 1
    REVOLUTE = 2


def narrow_limits_to_float32(limits: np.ndarray) -> np.ndarray:
    """Cast ``[lo, hi]`` inward to float32 (robo/robot_spec_tensors.py:17-31)."""
    limits = np.asarray(limits, dtype=np.float64)
    narrowed = limits.astype(np.float32)
    if narrowed.size == 0:
        return narrowed
    lo, hi = narrowed[:, 0], narrowed[:, 1]
    np.copyto(lo, np.nextafter(lo, np.float32(np.inf)), where=lo < limits[:, 0])
    np.copyto(hi, np.nextafter(hi, np.float32(-np.inf)), where=hi > limits[:, 1])
    return narrowed


def rpy_to_matrix(rpy: Sequence[float]) -> np.ndarray:
    """URDF fixed-axis roll-pitch-yaw: R = Rz(yaw) Ry(pitch) Rx(roll)."""
    r, p, y = (float(v) for v in rpy)
    cr, sr, cp, sp, cy, sy = math.cos(r), math.sin(r), math.cos(p), math.sin(p), math.cos(y), math.sin(y)
    return np.array(
        [
            [cy * cp, cy * sp * sr - sy * cr, cy * sp * cr + sy * sr],
            [sy * cp, sy * sp * sr + cy * cr, sy * sp * cr - cy * sr],
            [-sp, cp * sr, cp * cr],
        ]
    )


def matrix_to_quaternion(R: np.ndarray) -> np.ndarray:
    """Rotation matrix -> unit quaternion wxyz with w >= 0 (largest-pivot method)."""
    R = np.asarray(R, dtype=np.float64)
    m00, m01, m02, m10, m11, m12, m20, m21, m22 = R.reshape(9)
    cands = np.array(
        [
            [1 + m00 + m11 + m22, m21 - m12, m02 - m20, m10 - m01],
            [m21 - m12, 1 + m00 - m11 - m22, m10 + m01, m02 + m20],
            [m02 - m20, m10 + m01, 1 - m00 + m11 - m22, m12 + m21],
            [m10 - m01, m20 + m02, m21 + m12, 1 - m00 - m11 + m22],
        ]
    )
    k = int(np.argmax(np.diag(cands)))
    q = cands[k] / (2.0 * math.sqrt(max(cands[k, k], 1e-300)))
    q = q / np.linalg.norm(q)
    return -q if q[0] < 0 else q


def pose_from_xyz_rpy(xyz: Sequence[float], rpy: Sequence[float]) -> np.ndarray:
    return np.concatenate([np.asarray(xyz, dtype=np.float64), matrix_to_quaternion(rpy_to_matrix(rpy))])


@dataclass
class RobotSpec:
    """Kinematic description (robo/robot_spec.py:45-138).  Arrays are float64 on the host;
    the device copy is float32 (limits narrowed inward)."""

    name: str
    joint_names: List[str]
    actuated_joint_names: List[str]
    joint_types: np.ndarray  # [nj] int32
    joint_axes: np.ndarray  # [nj, 3]
    actuated_joint_indices: np.ndarray  # [nj] int32, -1 if not actuated
    parent_joint_indices: np.ndarray  # [nj] int32, -1 for a root joint
    parent_joint_transforms: np.ndarray  # [nj, 7] xyz wxyz
    mimic_actuated_joint_indices: np.ndarray  # [nj] int32
    mimic_multipliers: np.ndarray  # [nj]
    mimic_offsets: np.ndarray  # [nj]
    actuated_joint_limits: np.ndarray  # [na, 2]
    link_names: List[str]
    link_parent_joint_indices: np.ndarray  # [nl] int32
    base_link_name: str = ""
    ee_link_names: List[str] = field(default_factory=list)
    # collision spheres (link-local centres), robot_spec.py:228-259
    local_collision_sphere_centers: np.ndarray = field(default_factory=lambda: np.zeros((0, 3)))
    collision_sphere_radii: np.ndarray = field(default_factory=lambda: np.zeros((0,)))
    collision_spheres_link_indices: np.ndarray = field(default_factory=lambda: np.zeros((0,), np.int32))

    # ------------------------------------------------------------------ sizes
    @property
    def num_joints(self) -> int:
        return len(self.joint_names)

    @property
    def num_actuated_joints(self) -> int:
        return len(self.actuated_joint_names)

    @property
    def num_links(self) -> int:
        return len(self.link_names)

    @property
    def zero_q(self) -> np.ndarray:
        return np.zeros(self.num_actuated_joints, dtype=np.float32)

    @property
    def joint_twists(self) -> np.ndarray:
        """revolute [0,0,0,axis], prismatic [axis,0,0,0], fixed 0 (robot_spec.py:320-331)."""
        tw = np.zeros((self.num_joints, 6))
        for j, t in enumerate(self.joint_types):
            if t == JointType.REVOLUTE:
                tw[j, 3:] = self.joint_axes[j]
            elif t == JointType.PRISMATIC:
                tw[j, :3] = self.joint_axes[j]
        return tw

    @property
    def joints_to_actuated_mapping(self) -> np.ndarray:
        """robot_spec.py:1120-1132."""
        m = np.zeros((self.num_joints, self.num_actuated_joints), dtype=np.float32)
        for j in range(self.num_joints):
            a, mi = int(self.actuated_joint_indices[j]), int(self.mimic_actuated_joint_indices[j])
            if a >= 0:
                m[j, a] = 1.0
            elif mi >= 0:
                m[j, mi] = self.mimic_multipliers[j]
        return m

    @property
    def link_ancestor_joints_mask(self) -> np.ndarray:
        """robot_spec.py:1134-1143."""
        mask = np.zeros((self.num_links, self.num_joints), dtype=bool)
        for link in range(self.num_links):
            j = int(self.link_parent_joint_indices[link])
            while j != -1:
                mask[link, j] = True
                j = int(self.parent_joint_indices[j])
        return mask

    @property
    def topological_order_joint_indices(self) -> np.ndarray:
        """Kahn order over parent edges (robot_spec.py:1044-1069)."""
        deps: List[List[int]] = [[] for _ in range(self.num_joints)]
        indeg = [0] * self.num_joints
        for j, p in enumerate(self.parent_joint_indices):
            if p >= 0:
                deps[int(p)].append(j)
                indeg[j] += 1
        queue = [j for j in range(self.num_joints) if indeg[j] == 0]
        order = []
        while queue:
            j = queue.pop(0)
            order.append(j)
            for d in deps[j]:
                indeg[d] -= 1
                if indeg[d] == 0:
                    queue.append(d)
        if len(order) != self.num_joints:
            raise ValueError("Circular dependency detected in joint relationships")
        return np.asarray(order, dtype=np.int32)

    @property
    def joint_depths(self) -> np.ndarray:
        depth = np.zeros(self.num_joints, dtype=np.int32)
        for j in self.topological_order_joint_indices:
            p = int(self.parent_joint_indices[j])
            depth[j] = 0 if p < 0 else depth[p] + 1
        return depth

    @property
    def actuated_joint_limits_f32(self) -> np.ndarray:
        return narrow_limits_to_float32(self.actuated_joint_limits)

    def validate(self) -> None:
        nj = self.num_joints
        for name, arr, shape in (
            ("joint_types", self.joint_types, (nj,)),
            ("joint_axes", self.joint_axes, (nj, 3)),
            ("parent_joint_transforms", self.parent_joint_transforms, (nj, 7)),
            ("parent_joint_indices", self.parent_joint_indices, (nj,)),
            ("actuated_joint_limits", self.actuated_joint_limits, (self.num_actuated_joints, 2)),
            ("link_parent_joint_indices", self.link_parent_joint_indices, (self.num_links,)),
        ):
            if np.shape(arr) != shape:
                raise ValueError(f"{name} has shape {np.shape(arr)}, expected {shape}")
        self.topological_order_joint_indices  # raises on cycles

    # ------------------------------------------------------------------ builders
    @staticmethod
    def from_joints(
        name: str,
        link_names: Sequence[str],
        joints: Sequence[dict],
        base_link_name: Optional[str] = None,
    ) -> "RobotSpec":
        """Build from joint dicts ``{name, type, parent, child, xyz, rpy|quat, axis, lower, upper,
        mimic: (joint, multiplier, offset) | None}`` in file order; links in ``link_names`` order."""
        joint_names = [j["name"] for j in joints]
        type_map = {"revolute": JointType.REVOLUTE, "continuous": JointType.REVOLUTE,
                    "prismatic": JointType.PRISMATIC, "fixed": JointType.FIXED}
        actuated = [j["name"] for j in joints if j["type"] != "fixed" and not j.get("mimic")]
        child_to_joint = {j["child"]: i for i, j in enumerate(joints)}
        nj = len(joints)
        types = np.zeros(nj, np.int32)
        axes = np.zeros((nj, 3))
        act_idx = -np.ones(nj, np.int32)
        parent_idx = -np.ones(nj, np.int32)
        parent_T = np.zeros((nj, 7))
        mimic_idx = -np.ones(nj, np.int32)
        mimic_mul = np.ones(nj)
        mimic_off = np.zeros(nj)
        limits = []
        for i, j in enumerate(joints):
            if j["type"] not in type_map:
                raise ValueError(f"unsupported joint type {j['type']!r} for joint {j['name']!r}")
            t = type_map[j["type"]]
            types[i] = int(t)
            if t != JointType.FIXED:
                ax = np.asarray(j.get("axis", (1.0, 0.0, 0.0)), dtype=np.float64)
                axes[i] = ax
            mim = j.get("mimic")
            if mim:
                mimic_idx[i] = actuated.index(mim[0])
                mimic_mul[i] = 1.0 if mim[1] is None else float(mim[1])
                mimic_off[i] = 0.0 if mim[2] is None else float(mim[2])
            elif t != JointType.FIXED:
                act_idx[i] = actuated.index(j["name"])
            parent_idx[i] = child_to_joint.get(j["parent"], -1)
            if "quat" in j:
                parent_T[i] = np.concatenate([np.asarray(j.get("xyz", (0, 0, 0)), float), np.asarray(j["quat"], float)])
            else:
                parent_T[i] = pose_from_xyz_rpy(j.get("xyz", (0, 0, 0)), j.get("rpy", (0, 0, 0)))
        for nm in actuated:
            j = joints[joint_names.index(nm)]
            lo, hi = j.get("lower"), j.get("upper")
            if lo is None or hi is None:
                lo, hi = (-math.pi, math.pi) if j["type"] in ("revolute", "continuous") else (-1.0, 1.0)
            limits.append((float(lo), float(hi)))
        link_parent = np.array([child_to_joint.get(ln, -1) for ln in link_names], np.int32)
        if base_link_name is None:
            children = {j["child"] for j in joints}
            base_link_name = next(ln for ln in link_names if ln not in children)
        parents = {j["parent"] for j in joints}
        spec = RobotSpec(
            name=name,
            joint_names=joint_names,
            actuated_joint_names=actuated,
            joint_types=types,
            joint_axes=axes,
            actuated_joint_indices=act_idx,
            parent_joint_indices=parent_idx,
            parent_joint_transforms=parent_T,
            mimic_actuated_joint_indices=mimic_idx,
            mimic_multipliers=mimic_mul,
            mimic_offsets=mimic_off,
            actuated_joint_limits=np.asarray(limits, dtype=np.float64).reshape(-1, 2),
            link_names=list(link_names),
            link_parent_joint_indices=link_parent,
            base_link_name=base_link_name,
            ee_link_names=[ln for ln in link_names if ln not in parents],
        )
        spec.validate()
        return spec

    @staticmethod
    def parse_urdf(source: Union[str, Path]) -> "RobotSpec":
        """Parse URDF text or a path (robot_spec.py:284-618 ordering rules)."""
        text = Path(source).read_text() if isinstance(source, Path) or (
            isinstance(source, str) and not source.lstrip().startswith("<")
        ) else source
        root = ET.fromstring(text)
        link_names = [el.get("name") for el in root.findall("link")]
        joints = []
        for el in root.findall("joint"):
            origin = el.find("origin")
            axis = el.find("axis")
            limit = el.find("limit")
            mimic = el.find("mimic")
            parse = lambda s, d: [float(v) for v in s.split()] if s is not None else d  # noqa: E731
            jd = {
                "name": el.get("name"),
                "type": el.get("type"),
                "parent": el.find("parent").get("link"),
                "child": el.find("child").get("link"),
                "xyz": parse(origin.get("xyz") if origin is not None else None, [0.0, 0.0, 0.0]),
                "rpy": parse(origin.get("rpy") if origin is not None else None, [0.0, 0.0, 0.0]),
                "axis": parse(axis.get("xyz") if axis is not None else None, [1.0, 0.0, 0.0]),
                "lower": float(limit.get("lower")) if limit is not None and limit.get("lower") is not None else None,
                "upper": float(limit.get("upper")) if limit is not None and limit.get("upper") is not None else None,
                "mimic": None,
            }
            if mimic is not None:
                jd["mimic"] = (
                    mimic.get("joint"),
                    float(mimic.get("multiplier")) if mimic.get("multiplier") is not None else None,
                    float(mimic.get("offset")) if mimic.get("offset") is not None else None,
                )
            joints.append(jd)
        return RobotSpec.from_joints(root.get("name", "robot"), link_names, joints)

    @staticmethod
    def from_modified_dh(
        name: str,
        d: Sequence[float],
        a: Sequence[float],
        alpha: Sequence[float],
        limits: Union[Tuple[float, float], Sequence[Tuple[float, float]]],
    ) -> "RobotSpec":
        """Revolute serial chain from a modified (Craig) DH table:
        T_i = RotX(alpha_i) TransX(a_i) RotZ(q_i) TransZ(d_i).  TransZ(d_i) commutes with
        RotZ(q_i), so joint i's parent transform is RotX(alpha_i) TransX(a_i) TransZ(d_i) and its
        axis is +z.  Links: ``link0`` (base) .. ``link{n}`` (child of joint n)."""
        n = len(d)
        if not (len(a) == len(alpha) == n):
            raise ValueError("d, a, alpha must have equal length")
        lim = list(limits) if isinstance(limits[0], (tuple, list, np.ndarray)) else [tuple(limits)] * n
        joints = []
        for i in range(n):
            Rx = rpy_to_matrix((alpha[i], 0.0, 0.0))
            xyz = np.array([a[i], 0.0, 0.0]) + Rx @ np.array([0.0, 0.0, d[i]])
            joints.append(
                {
                    "name": f"joint{i + 1}",
                    "type": "revolute",
                    "parent": f"link{i}",
                    "child": f"link{i + 1}",
                    "xyz": xyz,
                    "quat": matrix_to_quaternion(Rx),
                    "axis": (0.0, 0.0, 1.0),
                    "lower": lim[i][0],
                    "upper": lim[i][1],
                }
            )
        return RobotSpec.from_joints(name, [f"link{i}" for i in range(n + 1)], joints, "link0")

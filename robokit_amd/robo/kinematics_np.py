"""Float64 numpy forward kinematics over a ``RobotSpec``.

Host-side utility for building synthetic targets (targets = FK of random joints) and for
unit conversions; it is not on the solve path, which runs only in the HIP library.  The
kinematic model is the reference's (robo/robot_kernels.py:264-297 joint transform,
parent-chain composition), written with rotation matrices instead of quaternions.
"""
from __future__ import annotations

import numpy as np

from robokit_amd.robo.robot_spec import JointType, RobotSpec


def quat_to_matrix(q: np.ndarray) -> np.ndarray:
    q = np.asarray(q, dtype=np.float64)
    w, x, y, z = q[..., 0], q[..., 1], q[..., 2], q[..., 3]
    s = 2.0 / (w * w + x * x + y * y + z * z)
    R = np.empty(q.shape[:-1] + (3, 3))
    R[..., 0, 0] = 1 - s * (y * y + z * z)
    R[..., 0, 1] = s * (x * y - z * w)
    R[..., 0, 2] = s * (x * z + y * w)
    R[..., 1, 0] = s * (x * y + z * w)
    R[..., 1, 1] = 1 - s * (x * x + z * z)
    R[..., 1, 2] = s * (y * z - x * w)
    R[..., 2, 0] = s * (x * z - y * w)
    R[..., 2, 1] = s * (y * z + x * w)
    R[..., 2, 2] = 1 - s * (x * x + y * y)
    return R


def matrix_to_quat(R: np.ndarray) -> np.ndarray:
    """Batched rotation matrix -> wxyz quaternion with w >= 0."""
    R = np.asarray(R, dtype=np.float64)
    shp = R.shape[:-2]
    R = R.reshape(-1, 3, 3)
    out = np.empty((R.shape[0], 4))
    m = R
    cands = np.stack(
        [
            np.stack([1 + m[:, 0, 0] + m[:, 1, 1] + m[:, 2, 2], m[:, 2, 1] - m[:, 1, 2], m[:, 0, 2] - m[:, 2, 0], m[:, 1, 0] - m[:, 0, 1]], -1),
            np.stack([m[:, 2, 1] - m[:, 1, 2], 1 + m[:, 0, 0] - m[:, 1, 1] - m[:, 2, 2], m[:, 1, 0] + m[:, 0, 1], m[:, 0, 2] + m[:, 2, 0]], -1),
            np.stack([m[:, 0, 2] - m[:, 2, 0], m[:, 1, 0] + m[:, 0, 1], 1 - m[:, 0, 0] + m[:, 1, 1] - m[:, 2, 2], m[:, 1, 2] + m[:, 2, 1]], -1),
            np.stack([m[:, 1, 0] - m[:, 0, 1], m[:, 2, 0] + m[:, 0, 2], m[:, 2, 1] + m[:, 1, 2], 1 - m[:, 0, 0] - m[:, 1, 1] + m[:, 2, 2]], -1),
        ],
        1,
    )  # [n, 4, 4]
    diag = np.stack([cands[:, k, k] for k in range(4)], -1)
    k = np.argmax(diag, axis=-1)
    sel = cands[np.arange(R.shape[0]), k]
    out = sel / (2.0 * np.sqrt(np.maximum(diag[np.arange(R.shape[0]), k], 1e-300)))[:, None]
    out /= np.linalg.norm(out, axis=-1, keepdims=True)
    out[out[:, 0] < 0] *= -1
    return out.reshape(shp + (4,))


def axis_angle_matrix(axis: np.ndarray, angle: np.ndarray) -> np.ndarray:
    """Rodrigues: rotation by ``angle`` [...] about unit ``axis`` [3]."""
    a = np.asarray(axis, dtype=np.float64)
    K = np.array([[0, -a[2], a[1]], [a[2], 0, -a[0]], [-a[1], a[0], 0]])
    th = np.asarray(angle, dtype=np.float64)[..., None, None]
    return np.eye(3) + np.sin(th) * K + (1 - np.cos(th)) * (K @ K)


def forward_kinematics(spec: RobotSpec, q: np.ndarray) -> np.ndarray:
    """World poses of every link, ``[..., num_links, 7]`` xyz-wxyz, float64, identity base."""
    q = np.asarray(q, dtype=np.float64)
    batch = q.shape[:-1]
    qf = q.reshape(-1, spec.num_actuated_joints)
    n = qf.shape[0]
    nj = spec.num_joints
    Rw = np.zeros((nj, n, 3, 3))
    tw = np.zeros((nj, n, 3))
    for j in spec.topological_order_joint_indices:
        j = int(j)
        if spec.mimic_actuated_joint_indices[j] >= 0:
            pos = qf[:, spec.mimic_actuated_joint_indices[j]] * spec.mimic_multipliers[j] + spec.mimic_offsets[j]
        elif spec.actuated_joint_indices[j] >= 0:
            pos = qf[:, spec.actuated_joint_indices[j]]
        else:
            pos = np.zeros(n)
        P = spec.parent_joint_transforms[j]
        Rp = quat_to_matrix(P[3:7])
        tp = P[:3]
        if spec.joint_types[j] == JointType.REVOLUTE:
            Rj = axis_angle_matrix(spec.joint_axes[j], pos)
            tj = np.zeros((n, 3))
        elif spec.joint_types[j] == JointType.PRISMATIC:
            Rj = np.broadcast_to(np.eye(3), (n, 3, 3))
            tj = pos[:, None] * spec.joint_axes[j][None]
        else:
            Rj = np.broadcast_to(np.eye(3), (n, 3, 3))
            tj = np.zeros((n, 3))
        Rl = Rp @ Rj
        tl = tp[None] + (Rp @ tj[..., None])[..., 0]
        p = int(spec.parent_joint_indices[j])
        if p < 0:
            Rw[j], tw[j] = Rl, tl
        else:
            Rw[j] = Rw[p] @ Rl
            tw[j] = tw[p] + (Rw[p] @ tl[..., None])[..., 0]
    out = np.zeros((n, spec.num_links, 7))
    out[..., 3] = 1.0
    for link in range(spec.num_links):
        j = int(spec.link_parent_joint_indices[link])
        if j >= 0:
            out[:, link, :3] = tw[j]
            out[:, link, 3:] = matrix_to_quat(Rw[j])
    return out.reshape(batch + (spec.num_links, 7))

"""Robots and seeded synthetic inputs for the batched-LM workloads.

* ``panda_dh()`` -- 7-DOF revolute chain from the modified-DH table
  d=[0.333,0,0.316,0,0.384,0,0.107], a=[0,0,0,0.0825,-0.0825,0,0.088],
  alpha=[0,-pi/2,pi/2,pi/2,-pi/2,pi/2,pi/2], limits +-2.8 rad.  Its end link equals the
  Franka Panda ``panda_link8`` (checked against the reference URDF in the tests).
* ``ur10()`` -- 6-DOF UR10 serial arm (joint origins of ur_description), limits +-2.8 rad,
  end link ``ee_link``; its FK is checked against the reference's UR10 FK ground truth.
* ``humanoid29(seed)`` -- 29-DOF fixed-base tree: pelvis root, 3-DOF waist to the chest,
  2 x 6-DOF legs from the pelvis, 2 x 7-DOF arms from the chest; every joint offset is
  L*u (L ~ U(0.1, 0.4) m, u a random unit vector), every axis a random unit vector,
  limits +-2.8 rad.

Every random array is drawn from its own child stream of ``np.random.SeedSequence(seed)``.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field
from typing import Dict, List, Optional, Sequence

import numpy as np

from robokit_amd.robo.kinematics_np import forward_kinematics
from robokit_amd.robo.robot_spec import RobotSpec

PI = math.pi
LIMIT = 2.8


def _streams(seed: int, n: int) -> List[np.random.Generator]:
    return [np.random.default_rng(s) for s in np.random.SeedSequence(seed).spawn(n)]


def panda_dh() -> RobotSpec:
    return RobotSpec.from_modified_dh(
        "panda_dh",
        d=[0.333, 0.0, 0.316, 0.0, 0.384, 0.0, 0.107],
        a=[0.0, 0.0, 0.0, 0.0825, -0.0825, 0.0, 0.088],
        alpha=[0.0, -PI / 2, PI / 2, PI / 2, -PI / 2, PI / 2, PI / 2],
        limits=(-LIMIT, LIMIT),
    )


def ur10(limit: float = LIMIT) -> RobotSpec:
    """UR10 joint origins (ur_description), base_link -> ee_link."""
    rev = lambda name, parent, child, xyz, rpy, axis: {  # noqa: E731
        "name": name, "type": "revolute", "parent": parent, "child": child,
        "xyz": xyz, "rpy": rpy, "axis": axis, "lower": -limit, "upper": limit,
    }
    joints = [
        rev("shoulder_pan_joint", "base_link", "shoulder_link", (0, 0, 0.1273), (0, 0, 0), (0, 0, 1)),
        rev("shoulder_lift_joint", "shoulder_link", "upper_arm_link", (0, 0.220941, 0), (0, PI / 2, 0), (0, 1, 0)),
        rev("elbow_joint", "upper_arm_link", "forearm_link", (0, -0.1719, 0.612), (0, 0, 0), (0, 1, 0)),
        rev("wrist_1_joint", "forearm_link", "wrist_1_link", (0, 0, 0.5723), (0, PI / 2, 0), (0, 1, 0)),
        rev("wrist_2_joint", "wrist_1_link", "wrist_2_link", (0, 0.1149, 0), (0, 0, 0), (0, 0, 1)),
        rev("wrist_3_joint", "wrist_2_link", "wrist_3_link", (0, 0, 0.1157), (0, 0, 0), (0, 1, 0)),
        {"name": "ee_fixed_joint", "type": "fixed", "parent": "wrist_3_link", "child": "ee_link",
         "xyz": (0, 0.0922, 0), "rpy": (0, 0, PI / 2)},
    ]
    links = ["base_link", "shoulder_link", "upper_arm_link", "forearm_link", "wrist_1_link", "wrist_2_link",
             "wrist_3_link", "ee_link"]
    return RobotSpec.from_joints("ur10", links, joints, "base_link")


def panda_tcp() -> RobotSpec:
    """The Franka Panda arm as the reference's IK benchmark builds it (finger joints fixed,
    scripts/benchmark/ik/helper.py:47-61): the seven revolute joints of franka_description with
    their limits, link7 -> panda_hand as one fixed joint (0.107 m along z, -pi/4 about z), and
    panda_hand -> panda_hand_tcp (0.1034 m along z), the benchmark's target link.  Its FK is checked
    against the reference's Panda ground truth for panda_hand."""
    rev = lambda i, xyz, rpy, lo, hi: {  # noqa: E731
        "name": f"panda_joint{i}", "type": "revolute", "parent": f"panda_link{i - 1}", "child": f"panda_link{i}",
        "xyz": xyz, "rpy": rpy, "axis": (0, 0, 1), "lower": lo, "upper": hi,
    }
    h = PI / 2
    joints = [
        rev(1, (0, 0, 0.333), (0, 0, 0), -2.8973, 2.8973),
        rev(2, (0, 0, 0), (-h, 0, 0), -1.7628, 1.7628),
        rev(3, (0, -0.316, 0), (h, 0, 0), -2.8973, 2.8973),
        rev(4, (0.0825, 0, 0), (h, 0, 0), -3.0718, -0.0698),
        rev(5, (-0.0825, 0.384, 0), (-h, 0, 0), -2.8973, 2.8973),
        rev(6, (0, 0, 0), (h, 0, 0), -0.0175, 3.7525),
        rev(7, (0.088, 0, 0), (h, 0, 0), -2.8973, 2.8973),
        {"name": "panda_hand_joint", "type": "fixed", "parent": "panda_link7", "child": "panda_hand",
         "xyz": (0, 0, 0.107), "rpy": (0, 0, -PI / 4)},
        {"name": "panda_hand_tcp_joint", "type": "fixed", "parent": "panda_hand", "child": "panda_hand_tcp",
         "xyz": (0, 0, 0.1034), "rpy": (0, 0, 0)},
    ]
    links = [f"panda_link{i}" for i in range(8)] + ["panda_hand", "panda_hand_tcp"]
    return RobotSpec.from_joints("panda_tcp", links, joints, "panda_link0")


HUMANOID_LIMBS = (("waist", "pelvis", 3), ("left_leg", "pelvis", 6), ("right_leg", "pelvis", 6),
                  ("left_arm", "chest", 7), ("right_arm", "chest", 7))


def humanoid29(seed: int = 0) -> RobotSpec:
    rng_len, rng_dir, rng_axis = _streams(seed, 3)
    joints = []
    links = ["pelvis"]
    for limb, parent, n in HUMANOID_LIMBS:
        prev = parent
        for k in range(n):
            child = "chest" if (limb == "waist" and k == n - 1) else f"{limb}_link{k + 1}"
            L = rng_len.uniform(0.1, 0.4)
            u = rng_dir.normal(size=3)
            u /= np.linalg.norm(u)
            ax = rng_axis.normal(size=3)
            ax /= np.linalg.norm(ax)
            joints.append({"name": f"{limb}_joint{k + 1}", "type": "revolute", "parent": prev, "child": child,
                           "xyz": L * u, "rpy": (0.0, 0.0, 0.0), "axis": ax, "lower": -LIMIT, "upper": LIMIT})
            links.append(child)
            prev = child
    return RobotSpec.from_joints(f"humanoid29_s{seed}", links, joints, "pelvis")


def humanoid_ee_links(spec: RobotSpec) -> List[int]:
    """Pose-target links: both hands, then both feet (the last link of each limb)."""
    return [spec.link_names.index(n) for n in ("left_arm_link7", "right_arm_link7", "left_leg_link6", "right_leg_link6")]


def humanoid_position_links(spec: RobotSpec) -> List[int]:
    """12 link-position targets: links 2, 4, 6 of each leg and 3, 5, 7 of each arm."""
    names = [f"{s}_leg_link{k}" for s in ("left", "right") for k in (2, 4, 6)]
    names += [f"{s}_arm_link{k}" for s in ("left", "right") for k in (3, 5, 7)]
    return [spec.link_names.index(n) for n in names]


def uniform_q(rng: np.random.Generator, spec: RobotSpec, n: int) -> np.ndarray:
    lim = spec.actuated_joint_limits
    return rng.uniform(lim[:, 0], lim[:, 1], size=(n, spec.num_actuated_joints)).astype(np.float32)


@dataclass
class IKBatch:
    """One homogeneous IK batch: a robot, its pose-target links, seeds and targets."""

    spec: RobotSpec
    links: List[int]
    q_seed: np.ndarray  # [B, na] float32
    targets: np.ndarray  # [B, k, 7] float32
    q_target: np.ndarray  # [B, na] the joints the targets were made from
    meta: Dict = field(default_factory=dict)

    @property
    def size(self) -> int:
        return self.q_seed.shape[0]


def ik_batch(spec: RobotSpec, links: List[int], n: int, seed: int) -> IKBatch:
    r_tgt, r_seed = _streams(seed, 2)
    q_tgt = uniform_q(r_tgt, spec, n)
    q_seed = uniform_q(r_seed, spec, n)
    T = forward_kinematics(spec, q_tgt.astype(np.float64))[:, links, :].astype(np.float32)
    return IKBatch(spec, list(links), q_seed, T, q_tgt)


def config0(seed: int = 0, batch: int = 1024) -> IKBatch:
    """Plain LM IK on the 7-DOF chain: targets = FK of uniform in-limit joints."""
    spec = panda_dh()
    return ik_batch(spec, [spec.link_names.index("link7")], batch, seed)


@dataclass
class MixedBatch:
    """Uneven batch: buckets of different robots, plus a random problem order."""

    buckets: List[IKBatch]
    order: np.ndarray  # [total] (bucket, index) pairs in arrival order

    @property
    def size(self) -> int:
        return sum(b.size for b in self.buckets)


def config1(seed: int = 0, total: int = 16384, humanoid_seed: int = 0) -> MixedBatch:
    """40 % 6-DOF UR10, 40 % 7-DOF chain, 20 % 29-DOF humanoid with 4 pose targets."""
    n6 = int(round(0.4 * total))
    n7 = int(round(0.4 * total))
    n29 = total - n6 - n7
    s = np.random.SeedSequence(seed).spawn(4)
    u = ur10()
    p = panda_dh()
    h = humanoid29(humanoid_seed)
    buckets = [
        ik_batch(u, [u.link_names.index("ee_link")], n6, int(s[0].generate_state(1)[0])),
        ik_batch(p, [p.link_names.index("link7")], n7, int(s[1].generate_state(1)[0])),
        ik_batch(h, humanoid_ee_links(h), n29, int(s[2].generate_state(1)[0])),
    ]
    ids = np.concatenate([np.stack([np.full(b.size, i), np.arange(b.size)], 1) for i, b in enumerate(buckets)])
    perm = np.random.default_rng(s[3]).permutation(len(ids))
    return MixedBatch(buckets, ids[perm])


@dataclass
class RetargetClip:
    """Smooth joint trajectory of the humanoid and noisy link-position targets per frame."""

    spec: RobotSpec
    links: List[int]
    q_traj: np.ndarray  # [F, na]
    targets: np.ndarray  # [F, 12, 7] (positions in xyz, identity rotation)
    fps: float = 30.0


def config2(seed: int = 0, frames: int = 4096, humanoid_seed: int = 0) -> RetargetClip:
    """q_j(t) = sum_{k=1..5} 0.1 sin(2 pi f_k t + phi_k), f ~ U(0.1, 2) Hz, phi ~ U(0, 2 pi): a total
    amplitude of 0.5 rad shared by the five sinusoids (at most ~0.16 rad per frame at 30 fps; 0.5 rad
    each would move joints up to ~1 rad per frame), 30 fps; targets = FK positions + N(0, (5 mm)^2)."""
    spec = humanoid29(humanoid_seed)
    r_f, r_p, r_n = _streams(seed, 3)
    na = spec.num_actuated_joints
    f = r_f.uniform(0.1, 2.0, size=(na, 5))
    ph = r_p.uniform(0.0, 2 * PI, size=(na, 5))
    t = np.arange(frames) / 30.0
    q = (0.1 * np.sin(2 * PI * f[None] * t[:, None, None] + ph[None])).sum(-1)
    q = np.clip(q, -LIMIT, LIMIT)
    links = humanoid_position_links(spec)
    P = forward_kinematics(spec, q)[:, links, :3]
    P = P + r_n.normal(0.0, 0.005, size=P.shape)
    tg = np.zeros((frames, len(links), 7), np.float32)
    tg[..., :3] = P
    tg[..., 3] = 1.0
    return RetargetClip(spec, links, q.astype(np.float32), tg)


@dataclass
class ServeTrace:
    """Poisson request stream over problem pools: request i arrives at t[i] with n[i] problems of
    kind[i], taken from that kind's pool at offset off[i]."""

    kinds: List[str]
    pools: Dict[str, object]  # kind -> IKBatch, or (RetargetClip, seeds, q_prev) for "retarget"
    t: np.ndarray
    kind: np.ndarray
    n: np.ndarray
    off: np.ndarray

    @property
    def size(self) -> int:
        return len(self.t)


def config4(seed: int = 0, rate: float = 2000.0, seconds: float = 60.0, pool: int = 4096,
            kinds: Sequence[str] = ("ik6", "ik7", "ik29", "retarget", "trajopt")) -> ServeTrace:
    """Poisson arrivals at `rate` requests/s for `seconds`; problems per request floor(exp(U(0, ln 257)))
    in [1, 256]; request kind uniform over `kinds` (IK 6/7/29-DOF as config 1, one retargeting frame
    of config 2 per problem seeded at the clip's previous frame, T = 64 trajectories of config 3)."""
    s = np.random.SeedSequence(seed).spawn(8)
    u, p, h = ur10(), panda_dh(), humanoid29(0)
    pools: Dict[str, object] = {}
    if "ik6" in kinds:
        pools["ik6"] = ik_batch(u, [u.link_names.index("ee_link")], pool, int(s[0].generate_state(1)[0]))
    if "ik7" in kinds:
        pools["ik7"] = ik_batch(p, [p.link_names.index("link7")], pool, int(s[1].generate_state(1)[0]))
    if "ik29" in kinds:
        pools["ik29"] = ik_batch(h, humanoid_ee_links(h), pool, int(s[2].generate_state(1)[0]))
    if "retarget" in kinds:
        clip = config2(seed=int(s[3].generate_state(1)[0]) % 1000, frames=pool + 1)
        prev = clip.q_traj[:-1] + np.random.default_rng(s[4]).normal(0, 0.02, clip.q_traj[:-1].shape)
        pools["retarget"] = (clip, prev.astype(np.float32))
    if "trajopt" in kinds:
        pools["trajopt"] = config3(seed=int(s[6].generate_state(1)[0]) % 1000, batch=pool, T=64)
    rng = np.random.default_rng(s[5])
    n_req = int(rate * seconds * 1.2) + 16
    t = np.cumsum(rng.exponential(1.0 / rate, n_req))
    t = t[t < seconds]
    m = len(t)
    kind = rng.integers(0, len(kinds), m)
    n = np.floor(np.exp(rng.uniform(0, np.log(257), m))).astype(np.int64).clip(1, 256)
    off = rng.integers(0, pool - 256, m)
    return ServeTrace(list(kinds), pools, t, kind, n, off)


@dataclass
class TrajScene:
    """Collision model of a serial arm and a primitive scene (world frame)."""

    sphere_links: np.ndarray  # [m] int32, link of each robot sphere
    sphere_centers: np.ndarray  # [m, 3] float32, in the link frame
    sphere_radii: np.ndarray  # [m] float32
    obstacle_spheres: np.ndarray  # [ns, 4] float32: centre, radius
    obstacle_boxes: np.ndarray  # [nb, 6] float32: centre, half extents (axis-aligned)
    geoms: list = field(default_factory=list)  # volumes / meshes (robokit_amd.geom), after the boxes


@dataclass
class TrajBatch:
    """Trajectory problems: seeds [B, T, na], goal pose of the last frame [B, 7]."""

    spec: RobotSpec
    goal_link: int
    T: int
    q_seed: np.ndarray
    goals: np.ndarray
    scene: TrajScene
    wv: float = 1e-2
    wa: float = 1e-2
    dt: float = 1.0
    coll_w: float = 20.0
    margin: float = 0.02
    # TrajectoryLimitTask bound rows (forward stencils): bit o of lim_orders for order o (0 position,
    # 1 velocity, 2 acceleration); weight / lower / upper per order and dof [3, na]
    lim_orders: int = 0
    lim_w: Optional[np.ndarray] = None
    lim_lo: Optional[np.ndarray] = None
    lim_hi: Optional[np.ndarray] = None
    # linear solver: 0 the band Cholesky, n > 0 n conjugate-gradient iterations per LM step
    # (LMOptimizerConfig(linear_solver="cg", cg_max_iter=n); the reference's default budget is
    # max(4, min(64, T * na // 4)), cg_default_iters)
    cg_iters: int = 0

    @property
    def cg_default_iters(self) -> int:
        """opt/lm_optimizer.py:577: max(4, min(64, tangent_dim // 4))"""
        return max(4, min(64, self.T * self.spec.num_actuated_joints // 4))

    @property
    def size(self) -> int:
        return self.q_seed.shape[0]


def arm_spheres(spec: RobotSpec, per_link: int, rng: np.random.Generator, r_lo=0.03, r_hi=0.08):
    """`per_link` spheres on every link of a serial arm, spaced along the segment to the next
    joint's origin (the last link: 0.1 m along its z axis); radii uniform in [r_lo, r_hi]."""
    nl = spec.num_links
    pT = np.asarray(spec.parent_joint_transforms)
    child = {int(spec.link_parent_joint_indices[l]): l for l in range(nl)}
    links, centres = [], []
    for l in range(nl):
        j_in = int(spec.link_parent_joint_indices[l])
        nxt = [j for j in range(spec.num_joints) if int(spec.parent_joint_indices[j]) == j_in]
        seg = pT[nxt[0], :3] if nxt else np.array([0.0, 0.0, 0.1])
        for i in range(per_link):
            links.append(l)
            centres.append(seg * (i + 0.5) / per_link)
    radii = rng.uniform(r_lo, r_hi, len(links))
    del child
    return np.asarray(links, np.int32), np.asarray(centres, np.float32), radii.astype(np.float32)


def config3(seed: int = 0, batch: int = 512, T: int = 64, n_obs_spheres: int = 32, n_obs_boxes: int = 8,
            arm: str = "panda", spheres_per_link: int = 6) -> TrajBatch:
    """Collision-aware trajectory LM on the 7-DOF chain (arm "ur10": the 6-DOF UR10): 6 robot
    spheres per link (radius 3-8 cm; 48 on the 7-DOF chain); 32 obstacle spheres (radius 5-15 cm)
    + 8 axis-aligned boxes (half extents 5-15 cm), centres uniform in the 1.5 m cube
    [-0.75, 0.75]^2 x [0, 1.5] around the base, none within 0.3 m of the base's vertical axis
    below 0.5 m; seeds = straight joint-space lines between two uniform in-limit configurations;
    goal = FK of a third."""
    spec = {"panda": panda_dh, "ur10": ur10}[arm]()
    r_sph, r_obs, r_q = _streams(seed, 3)
    links, centres, radii = arm_spheres(spec, spheres_per_link, r_sph)

    def centres_ok(n):
        out = []
        while len(out) < n:
            c = r_obs.uniform([-0.75, -0.75, 0.0], [0.75, 0.75, 1.5])
            if np.hypot(c[0], c[1]) < 0.3 and c[2] < 0.5:
                continue
            out.append(c)
        return np.asarray(out)

    obs_s = np.concatenate([centres_ok(n_obs_spheres), r_obs.uniform(0.05, 0.15, (n_obs_spheres, 1))], 1)
    obs_b = np.concatenate([centres_ok(n_obs_boxes), r_obs.uniform(0.05, 0.15, (n_obs_boxes, 3))], 1)
    scene = TrajScene(links, centres, radii, obs_s.astype(np.float32), obs_b.astype(np.float32))
    qa, qb, qg = uniform_q(r_q, spec, batch), uniform_q(r_q, spec, batch), uniform_q(r_q, spec, batch)
    s = np.linspace(0.0, 1.0, T, dtype=np.float32)[None, :, None]
    q_seed = (qa[:, None] * (1 - s) + qb[:, None] * s).astype(np.float32)
    goal_link = spec.num_links - 1
    goals = forward_kinematics(spec, qg.astype(np.float64))[:, goal_link].astype(np.float32)
    return TrajBatch(spec, goal_link, T, q_seed, goals, scene)

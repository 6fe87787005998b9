"""ctypes binding of librkhip.so (include/rkhip.h).

There is no fallback: if the library is missing or the device is unavailable, the calls
raise.  The library is built in-tree by ``robokit_amd.build``.
"""
from __future__ import annotations

import atexit
import ctypes as C
import weakref
from dataclasses import dataclass
from pathlib import Path
from typing import Dict, List, Optional, Sequence

import numpy as np

PKG = Path(__file__).resolve().parent
LIB_PATH = PKG / "librkhip.so"

RK_OK, RK_EINVAL, RK_EHIP, RK_ECAPACITY, RK_ENOMEM, RK_EBUSY = range(6)
MODE_FUSED, MODE_PLAIN = 0, 1
TPP_WAVE = -64  # one wavefront per problem, register rows
TPP_HALF = -32  # two problems per wavefront (32 lanes each), register rows
TPP_QUAD = -4  # four lanes per problem, register rows (6- and 7-dof serial arms, one pose target)
TPP_SUPER = -33  # as TPP_HALF, normal equations solved in block-arrow form (limbs side by side, then the root)
TPP_TRAJ = -128  # a trajectory: two wavefronts per problem (group kernel times)
FROM_SEED, CONTINUE, USE_GRAPH, ASYNC, NOWAIT = 0x100, 0x200, 0x1, 0x2, 0x400
SE3_OPS = {"compose": (0, 7), "inverse": (1, 7), "log": (2, 6), "jlog": (3, 36), "adjoint": (4, 36), "apply": (5, 3)}


class RKError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"[rk {code}] {msg}")
        self.code = code


class RobotDesc(C.Structure):
    _fields_ = [
        ("nj", C.c_int32), ("nl", C.c_int32), ("na", C.c_int32),
        ("joint_type", C.c_void_p), ("joint_axis", C.c_void_p), ("parent_T", C.c_void_p),
        ("parent_joint", C.c_void_p), ("act_idx", C.c_void_p), ("mimic_idx", C.c_void_p),
        ("mimic_mul", C.c_void_p), ("mimic_off", C.c_void_p), ("link_parent_joint", C.c_void_p),
        ("limits", C.c_void_p),
    ]


class TermsDesc(C.Structure):
    _fields_ = [
        ("robot", C.c_int32),
        ("n_pose", C.c_int32), ("pose_links", C.c_void_p), ("pose_weights", C.c_void_p),
        ("n_pos", C.c_int32), ("pos_links", C.c_void_p), ("pos_weights", C.c_void_p),
        ("limit", C.c_int32), ("limit_weights", C.c_void_p),
        ("smooth", C.c_int32), ("smooth_weights", C.c_void_p),
        ("n_rot", C.c_int32), ("rot_links", C.c_void_p), ("rot_weights", C.c_void_p),
        ("n_coll", C.c_int32), ("coll_links", C.c_void_p), ("coll_centers", C.c_void_p),
        ("coll_radii", C.c_void_p), ("coll_weights", C.c_void_p), ("coll_margin", C.c_float),
        ("coll_penalty", C.c_int32), ("n_obs_spheres", C.c_int32), ("obs_spheres", C.c_void_p),
        ("n_obs_boxes", C.c_int32), ("obs_boxes", C.c_void_p), ("n_geoms", C.c_int32), ("geoms", C.c_void_p),
    ]


class VolumeDescC(C.Structure):
    _fields_ = [("values", C.c_void_p), ("nx", C.c_int32), ("ny", C.c_int32), ("nz", C.c_int32),
                ("origin", C.c_float * 3), ("voxel_size", C.c_float), ("aabb_min", C.c_float * 3),
                ("aabb_max", C.c_float * 3), ("padding", C.c_float)]


class MeshDescC(C.Structure):
    _fields_ = [("vertices", C.c_void_p), ("n_vertices", C.c_int32), ("faces", C.c_void_p), ("n_faces", C.c_int32)]

COLL_PENALTY = {"smooth": 0, "plain": 1, "surface_distance": 2}


class LMParamsC(C.Structure):
    _fields_ = [(n, C.c_float) for n in
                ("lm_lambda", "lambda_factor", "lambda_min", "lambda_max", "rho_min", "gain_ratio_epsilon")]


class TrajDescC(C.Structure):
    _fields_ = [("robot", C.c_int32), ("frames", C.c_int32), ("goal_link", C.c_int32), ("goal_weights", C.c_void_p),
                ("vel_weight", C.c_float), ("acc_weight", C.c_float), ("dt", C.c_float), ("n_spheres", C.c_int32),
                ("sphere_links", C.c_void_p), ("sphere_centers", C.c_void_p), ("sphere_radii", C.c_void_p),
                ("coll_weight", C.c_float), ("coll_margin", C.c_float), ("n_obs_spheres", C.c_int32),
                ("obs_spheres", C.c_void_p), ("n_obs_boxes", C.c_int32), ("obs_boxes", C.c_void_p),
                ("limit_orders", C.c_int32), ("limit_weights", C.c_void_p), ("limit_lower", C.c_void_p),
                ("limit_upper", C.c_void_p), ("cg_max_iter", C.c_int32), ("n_geoms", C.c_int32), ("geoms", C.c_void_p)]


def traj_desc(robot_id: int, tb, goal_weights=None, ctx: Optional["Context"] = None):
    """(rk_traj_desc, arrays it points at) of a workloads.TrajBatch-shaped problem set: robot,
    frames, scene (with its volumes / meshes uploaded to ``ctx``), weights, and the optional bound
    rows (lim_orders bit o, lim_w / lim_lo / lim_hi [3, na]) and the CG budget (cg_iters: 0 = band
    Cholesky)"""
    sc = tb.scene
    na = tb.spec.num_actuated_joints
    arrs = [_f32(np.ones(6) if goal_weights is None else goal_weights), _i32(sc.sphere_links),
            _f32(sc.sphere_centers), _f32(sc.sphere_radii), _f32(np.asarray(sc.obstacle_spheres).reshape(-1, 4)),
            _f32(np.asarray(sc.obstacle_boxes).reshape(-1, 6))]
    orders = int(getattr(tb, "lim_orders", 0) or 0)
    lim = [None, None, None]
    if orders:
        lim = [_f32(np.broadcast_to(np.asarray(getattr(tb, k), np.float32), (3, na))) for k in ("lim_w", "lim_lo", "lim_hi")]
    arrs += lim
    d = TrajDescC(robot_id, tb.T, tb.goal_link, _p(arrs[0]), tb.wv, tb.wa, tb.dt, len(arrs[1]), _p(arrs[1]),
                  _p(arrs[2]), _p(arrs[3]), tb.coll_w, tb.margin, len(arrs[4]), _p(arrs[4]), len(arrs[5]), _p(arrs[5]),
                  orders, _p(lim[0]), _p(lim[1]), _p(lim[2]), int(getattr(tb, "cg_iters", 0) or 0))
    geoms = list(getattr(sc, "geoms", None) or [])
    if geoms:
        if ctx is None:
            raise ValueError("a scene with volumes / meshes needs the context they are uploaded to")
        ids = _i32([ctx.geom(g) for g in geoms])
        arrs.append(ids)
        d.n_geoms, d.geoms = len(geoms), _p(ids)
    return d, arrs


class RequestC(C.Structure):
    _fields_ = [("terms", C.c_int32), ("n", C.c_int32), ("iters", C.c_int32), ("params", LMParamsC),
                ("q", C.c_void_p), ("pose_targets", C.c_void_p), ("pos_targets", C.c_void_p),
                ("q_prev", C.c_void_p), ("session", C.c_int64)]


class ServeStatsC(C.Structure):
    _fields_ = [(n, C.c_int64) for n in ("submitted", "completed", "failed", "queued", "launches", "problems",
                                         "converged")] + \
               [(n, C.c_double) for n in ("p50_ms", "p99_ms", "mean_ms", "max_ms", "mean_problems_per_launch")] + \
               [(n, C.c_int64) for n in ("budget_bytes", "peak_bytes", "bytes_in_use")] + [("span_s", C.c_double)]


_lib: Optional[C.CDLL] = None

# Live handles, closed at interpreter exit (groups, then batches, then contexts) while the HIP
# runtime is still up: objects freed during interpreter teardown would call into it afterwards.
_LIVE = {"group": weakref.WeakSet(), "server": weakref.WeakSet(), "batch": weakref.WeakSet(),
         "context": weakref.WeakSet()}


@atexit.register
def _close_all() -> None:
    for kind in ("group", "batch", "server", "context"):
        for obj in list(_LIVE.get(kind, ())):
            try:
                obj.close()
            except Exception:
                pass

EXPORTS = [
    "rk_last_error", "rk_abi_version", "rk_context_create", "rk_context_destroy", "rk_context_sync",
    "rk_context_mem", "rk_context_mark", "rk_context_elapsed", "rk_robot_upload", "rk_fk", "rk_fk_base", "rk_link_jacobian",
    "rk_link_jacobian_base", "rk_copy_probe", "rk_se3_op",
    "rk_batch_create",
    "rk_batch_destroy", "rk_batch_set", "rk_batch_set_rotations", "rk_batch_select", "rk_batch_select_scored",
    "rk_batch_copy_q", "rk_batch_advance", "rk_batch_solve", "rk_batch_get", "rk_batch_get_iterates", "rk_batch_metrics",
    "rk_batch_metrics_noalloc", "rk_group_create",
    "rk_group_destroy", "rk_group_set_tuning", "rk_group_autotune", "rk_group_solve", "rk_group_kernel_times",
    "rk_server_create", "rk_server_add_terms", "rk_server_add_traj", "rk_server_submit", "rk_server_poll", "rk_server_end_session",
    "rk_server_stats", "rk_server_destroy", "rk_traj_batch_create", "rk_traj_metrics",
    "rk_geom_upload_volume", "rk_geom_upload_mesh", "rk_scene_query",
]


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise RuntimeError(f"{LIB_PATH} is not built; run `python -m robokit_amd.build` (hipcc, gfx950)")
        L = C.CDLL(str(LIB_PATH))
        L.rk_last_error.restype = C.c_char_p
        for name in EXPORTS:
            if name != "rk_last_error":
                getattr(L, name).restype = C.c_int32
        _lib = L
    return _lib


def _check(rc: int) -> None:
    if rc != RK_OK:
        msg = lib().rk_last_error().decode()
        if rc in (RK_EINVAL, RK_ECAPACITY):
            raise ValueError(msg)
        raise RKError(rc, msg)


def _p(a: Optional[np.ndarray]):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def _f32(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.float32)


def _i32(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.int32)


@dataclass
class LMParams:
    """LMOptimizerConfig damping fields (opt/lm_optimizer.py:502-523 defaults)."""

    lm_lambda: float = 1e-3
    lambda_factor: float = 2.0
    lambda_min: float = 1e-5
    lambda_max: float = 1e10
    rho_min: float = 1e-3
    gain_ratio_epsilon: float = 1e-8

    def c(self) -> LMParamsC:
        return LMParamsC(self.lm_lambda, self.lambda_factor, self.lambda_min, self.lambda_max, self.rho_min,
                         self.gain_ratio_epsilon)


class Context:
    """One HIP device + stream + the resident robot models (rk_context)."""

    def __init__(self, device: int = 0):
        self._h = C.c_void_p()
        _check(lib().rk_context_create(int(device), C.byref(self._h)))
        _LIVE["context"].add(self)
        self._children = weakref.WeakSet()  # batches / groups / servers on this context
        self.device = device
        self._robots: Dict[int, int] = {}
        self._geoms: Dict[int, int] = {}
        self._keep: List = []

    def close(self) -> None:
        """Destroy the context; objects still built on it are closed first (their handles point
        into it)."""
        if self._h:
            kids = list(self._children)
            for k in sorted(kids, key=lambda o: 0 if isinstance(o, (Group, Server)) else 1):
                k.close()
            lib().rk_context_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def handle(self):
        return self._h

    def sync(self) -> None:
        _check(lib().rk_context_sync(self._h))

    def mark(self, slot: int) -> None:
        """record HIP event ``slot`` (0..7) on the context stream"""
        _check(lib().rk_context_mark(self._h, int(slot)))

    def elapsed(self, a: int, b: int) -> float:
        """milliseconds from mark ``a`` to mark ``b`` (waits for ``b``)"""
        ms = C.c_float()
        _check(lib().rk_context_elapsed(self._h, int(a), int(b), C.byref(ms)))
        return float(ms.value)

    def mem_bytes(self) -> int:
        v = C.c_int64()
        _check(lib().rk_context_mem(self._h, C.byref(v)))
        return int(v.value)

    def robot(self, spec) -> int:
        """Upload ``spec`` once (resident on the device) and return its id."""
        key = id(spec)
        if key in self._robots:
            return self._robots[key]
        arrs = [
            _i32(spec.joint_types), _f32(spec.joint_axes), _f32(spec.parent_joint_transforms),
            _i32(spec.parent_joint_indices), _i32(spec.actuated_joint_indices),
            _i32(spec.mimic_actuated_joint_indices), _f32(spec.mimic_multipliers), _f32(spec.mimic_offsets),
            _i32(spec.link_parent_joint_indices), _f32(spec.actuated_joint_limits_f32),
        ]
        d = RobotDesc(spec.num_joints, spec.num_links, spec.num_actuated_joints, *[_p(a) for a in arrs])
        rid = C.c_int32()
        _check(lib().rk_robot_upload(self._h, C.byref(d), C.byref(rid)))
        self._robots[key] = rid.value
        self._keep.append(spec)
        return rid.value

    def geom(self, g) -> int:
        """Upload a scene volume / mesh (robokit_amd.geom) once (resident) and return its id."""
        from .geom import SdfVolume, TriMesh

        key = id(g)
        if key in self._geoms:
            return self._geoms[key]
        gid = C.c_int32()
        if isinstance(g, SdfVolume):
            nx, ny, nz = g.shape_xyz
            vals = _f32(g.values)
            d = VolumeDescC(_p(vals), nx, ny, nz, (C.c_float * 3)(*g.origin), float(g.voxel_size),
                            (C.c_float * 3)(*g.aabb_min), (C.c_float * 3)(*g.aabb_max), float(g.padding))
            _check(lib().rk_geom_upload_volume(self._h, C.byref(d), C.byref(gid)))
        elif isinstance(g, TriMesh):
            v, f = _f32(g.vertices), _i32(g.faces)
            d = MeshDescC(_p(v), len(v), _p(f), len(f))
            _check(lib().rk_geom_upload_mesh(self._h, C.byref(d), C.byref(gid)))
        else:
            raise TypeError(f"scene geometries are SdfVolume or TriMesh, got {type(g).__name__}")
        self._geoms[key] = gid.value
        self._keep.append(g)
        return gid.value

    def scene_query(self, points: np.ndarray, spheres: Optional[np.ndarray] = None,
                    boxes: Optional[np.ndarray] = None, geoms: Sequence = ()):
        """(signed distance [n], normal [n, 3]) of the scene of obstacle spheres [k, 4], axis-aligned
        boxes [k, 6] and volumes / meshes at points [n, 3] (rk_scene_query)"""
        pts = _f32(np.asarray(points, np.float32).reshape(-1, 3))
        sp = _f32(np.zeros((0, 4)) if spheres is None else np.asarray(spheres).reshape(-1, 4))
        bx = _f32(np.zeros((0, 6)) if boxes is None else np.asarray(boxes).reshape(-1, 6))
        ids = _i32([self.geom(g) for g in geoms])
        n = len(pts)
        dist = np.zeros(n, np.float32)
        nrm = np.zeros((n, 3), np.float32)
        _check(lib().rk_scene_query(self._h, len(sp), _p(sp), len(bx), _p(bx), len(ids), _p(ids), n, _p(pts), _p(dist),
                                    _p(nrm)))
        return dist, nrm

    def fk(self, spec, q: np.ndarray, T_world_base: Optional[np.ndarray] = None) -> np.ndarray:
        """T_world_link [B, nl, 7]; T_world_base [B, 7] or [7] (None: the identity)"""
        q = _f32(np.atleast_2d(q))
        out = np.zeros((q.shape[0], spec.num_links, 7), np.float32)
        if T_world_base is None:
            _check(lib().rk_fk(self._h, self.robot(spec), q.shape[0], _p(q), _p(out)))
        else:
            base = _f32(np.broadcast_to(np.asarray(T_world_base, np.float32), (q.shape[0], 7)))
            _check(lib().rk_fk_base(self._h, self.robot(spec), q.shape[0], _p(q), _p(base), _p(out)))
        return out

    def link_jacobian(self, spec, q: np.ndarray, link: int, frame: str = "body",
                      T_world_base: Optional[np.ndarray] = None) -> np.ndarray:
        """J [B, 6, na]; with T_world_base ([B, 7] or [7]) the floating-base Jacobian [B, 6, 6 + na],
        base columns first"""
        q = _f32(np.atleast_2d(q))
        fr = 0 if frame == "spatial" else 1
        if T_world_base is None:
            out = np.zeros((q.shape[0], 6, spec.num_actuated_joints), np.float32)
            _check(lib().rk_link_jacobian(self._h, self.robot(spec), q.shape[0], _p(q), int(link), fr, _p(out)))
        else:
            base = _f32(np.broadcast_to(np.asarray(T_world_base, np.float32), (q.shape[0], 7)))
            out = np.zeros((q.shape[0], 6, 6 + spec.num_actuated_joints), np.float32)
            _check(lib().rk_link_jacobian_base(self._h, self.robot(spec), q.shape[0], _p(q), _p(base), int(link), fr,
                                               _p(out)))
        return out

    def copy_probe(self, nbytes: int) -> float:
        """Streaming copy of nbytes (counter calibration); returns the timed launch's ms."""
        ms = C.c_float()
        _check(lib().rk_copy_probe(self._h, C.c_int64(int(nbytes)), C.byref(ms)))
        return float(ms.value)

    def se3(self, op: str, a: np.ndarray, b: Optional[np.ndarray] = None) -> np.ndarray:
        code, w = SE3_OPS[op]
        a = _f32(np.atleast_2d(a))
        bb = None if b is None else _f32(np.atleast_2d(b))
        out = np.zeros((a.shape[0], w), np.float32)
        _check(lib().rk_se3_op(self._h, code, a.shape[0], _p(a), _p(bb), _p(out)))
        return out.reshape(a.shape[0], 6, 6) if w == 36 else out


@dataclass
class TermSpec:
    """Uniform residual-term structure of a batch (row order: pose, position, rotation, limit,
    smoothness, collision)."""

    pose_links: Sequence[int] = ()
    pose_weights: Optional[np.ndarray] = None  # [n_pose, 6]
    pos_links: Sequence[int] = ()
    pos_weights: Optional[np.ndarray] = None  # [n_pos, 3]
    limit_weights: Optional[np.ndarray] = None  # [na] or None (no limit term)
    smooth_weights: Optional[np.ndarray] = None  # [na] or None (no smoothness term)
    rot_links: Sequence[int] = ()
    rot_weights: Optional[np.ndarray] = None  # [n_rot, 3]
    # SceneCollisionTask: robot spheres (link, centre in the link frame, radius, weight) against a
    # primitive scene (obstacle spheres [n, 4], axis-aligned boxes [n, 6]), rows after smoothness
    coll_links: Sequence[int] = ()
    coll_centers: Optional[np.ndarray] = None  # [n_coll, 3]
    coll_radii: Optional[np.ndarray] = None  # [n_coll]
    coll_weights: Optional[np.ndarray] = None  # [n_coll] (None: ones)
    coll_margin: float = 0.02
    coll_penalty: str = "smooth"  # smooth / plain / surface_distance
    obstacle_spheres: Optional[np.ndarray] = None  # [n, 4] centre, radius
    obstacle_boxes: Optional[np.ndarray] = None  # [n, 6] centre, half extents
    geoms: Sequence = ()  # scene volumes / meshes (robokit_amd.geom), after the boxes

    def rows(self, na: int) -> int:
        return 6 * len(self.pose_links) + 3 * len(self.pos_links) + 3 * len(self.rot_links) + \
            na * (self.limit_weights is not None) + na * (self.smooth_weights is not None) + len(self.coll_links)

    def desc(self, robot_id: int, na: int, ctx: Optional["Context"] = None):
        """(rk_terms_desc, the arrays it points at); the scene's volumes / meshes are uploaded to ``ctx``"""
        arrs = [
            _i32(list(self.pose_links)),
            _f32(self.pose_weights if self.pose_weights is not None else np.zeros((0, 6))),
            _i32(list(self.pos_links)),
            _f32(self.pos_weights if self.pos_weights is not None else np.zeros((0, 3))),
            _f32(self.limit_weights if self.limit_weights is not None else np.zeros(na)),
            _f32(self.smooth_weights if self.smooth_weights is not None else np.zeros(na)),
            _i32(list(self.rot_links)),
            _f32(self.rot_weights if self.rot_weights is not None else np.zeros((0, 3))),
        ]
        nc = len(self.coll_links)
        if self.coll_penalty not in COLL_PENALTY:
            raise ValueError(f"Unsupported penalty: {self.coll_penalty}")
        arrs += [
            _i32(list(self.coll_links)),
            _f32(np.zeros((0, 3)) if self.coll_centers is None else np.asarray(self.coll_centers).reshape(-1, 3)),
            _f32(np.zeros(0) if self.coll_radii is None else self.coll_radii),
            None if self.coll_weights is None else _f32(np.broadcast_to(np.asarray(self.coll_weights, np.float32), (nc,))),
            _f32(np.zeros((0, 4)) if self.obstacle_spheres is None else np.asarray(self.obstacle_spheres).reshape(-1, 4)),
            _f32(np.zeros((0, 6)) if self.obstacle_boxes is None else np.asarray(self.obstacle_boxes).reshape(-1, 6)),
        ]
        if nc and (arrs[9].shape[0] != nc or arrs[10].shape[0] != nc):
            raise ValueError(f"collision spheres: {nc} links but {arrs[9].shape[0]} centres / {arrs[10].shape[0]} radii")
        a = arrs
        d = TermsDesc(robot_id, len(self.pose_links), _p(a[0]), _p(a[1]), len(self.pos_links), _p(a[2]), _p(a[3]),
                      int(self.limit_weights is not None), _p(a[4]), int(self.smooth_weights is not None), _p(a[5]),
                      len(self.rot_links), _p(a[6]), _p(a[7]),
                      nc, _p(a[8]), _p(a[9]), _p(a[10]), _p(a[11]), float(self.coll_margin),
                      COLL_PENALTY[self.coll_penalty], a[12].shape[0], _p(a[12]), a[13].shape[0], _p(a[13]))
        if self.geoms:
            if ctx is None:
                raise ValueError("a scene with volumes / meshes needs the context they are uploaded to")
            ids = _i32([ctx.geom(g) for g in self.geoms])
            arrs.append(ids)
            d.n_geoms, d.geoms = len(ids), _p(ids)
        return d, arrs


class Batch:
    """Device-resident SoA buffers for up to ``capacity`` problems of one term structure."""

    def __init__(self, ctx: Context, spec, terms: TermSpec, capacity: int, trace_iters: int = 0):
        self.ctx, self.spec, self.terms = ctx, spec, terms
        self.capacity, self.trace_iters = int(capacity), int(trace_iters)
        na = spec.num_actuated_joints
        self._na = na  # variables per problem
        d, self._arrs = terms.desc(ctx.robot(spec), na, ctx)
        self._h = C.c_void_p()
        _check(lib().rk_batch_create(ctx.handle, C.byref(d), self.capacity, self.trace_iters, C.byref(self._h)))
        _LIVE["batch"].add(self)
        ctx._children.add(self)
        self._groups = weakref.WeakSet()
        self.n = 0

    def close(self) -> None:
        if self._h:
            for g in list(self._groups):  # a group holds this batch's handle
                g.close()
            lib().rk_batch_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def handle(self):
        return self._h

    def set(self, q: Optional[np.ndarray] = None, pose_targets: Optional[np.ndarray] = None,
            pos_targets: Optional[np.ndarray] = None, q_prev: Optional[np.ndarray] = None, n: Optional[int] = None,
            rot_targets: Optional[np.ndarray] = None):
        """rot_targets [B, n_rot, 4] quaternions (w, x, y, z), or [B, n_rot, 7] poses (rotation used)"""
        arrs = [None if x is None else _f32(x) for x in (q, pose_targets, pos_targets, q_prev)]
        if n is None:
            n = next(x.shape[0] for x in arrs + [rot_targets] if x is not None)
        if arrs[2] is not None and arrs[2].shape[-1] == 7:
            arrs[2] = _f32(arrs[2][..., :3])
        _check(lib().rk_batch_set(self._h, int(n), *[_p(x) for x in arrs]))
        self.n = int(n)
        if rot_targets is not None:
            rt = _f32(rot_targets)
            if rt.shape[-1] == 7:
                rt = _f32(rt[..., 3:])
            _check(lib().rk_batch_set_rotations(self._h, int(n), _p(rt)))

    def solve(self, iters: int, params: LMParams, mode: int = MODE_FUSED, tpp: int = 0, from_seed: bool = False,
              cont: bool = False, wait: bool = True) -> float:
        """Run ``iters`` LM iterations; ``cont`` continues the previous solve (q, damping, traces).
        ``wait=False`` only enqueues it (returns 0.0; a later synchronous call waits)."""
        ms = C.c_float()
        m = int(mode) | (FROM_SEED if from_seed else 0) | (CONTINUE if cont else 0) | (0 if wait else NOWAIT)
        _check(lib().rk_batch_solve(self._h, int(iters), C.byref(params.c()), m, int(tpp), C.byref(ms)))
        return float(ms.value) if wait else 0.0

    def get(self, traces: bool = False):
        na = self._na
        q = np.zeros((self.n, na), np.float32)
        cost = np.zeros(self.n, np.float32)
        lam = np.zeros(self.n, np.float32)
        ct = np.zeros((self.n, self.trace_iters + 1), np.float32) if traces else None
        at = np.zeros((self.n, self.trace_iters), np.int32) if traces else None
        _check(lib().rk_batch_get(self._h, _p(q), _p(cost), _p(lam), _p(ct), _p(at)))
        out = dict(q=q, cost=cost, lam=lam)
        if traces:
            out.update(cost_trace=ct, accept=at)
        return out

    def select(self, width_in: int, dst: "Batch", width_out: int, rows: bool = False,
               scores: Optional["Batch"] = None) -> Optional[np.ndarray]:
        """Keep each problem's best ``width_out`` of its ``width_in`` candidates (rows
        i * width_in + j) in ``dst``, best first: lowest cost at the current iterate (or the cost of
        ``scores``, a batch of this size holding each row's score), NaN last, equal costs in
        candidate order.  Returns the source row of each ``dst`` row if ``rows``."""
        n_out = self.n // int(width_in) * int(width_out)
        out = np.zeros(n_out, np.int32) if rows else None
        if scores is None:
            _check(lib().rk_batch_select(self._h, int(width_in), dst.handle, int(width_out), _p(out)))
        else:
            _check(lib().rk_batch_select_scored(self._h, scores.handle, int(width_in), dst.handle, int(width_out),
                                                _p(out)))
        dst.n = n_out
        return out

    def copy_q_to(self, dst: "Batch") -> None:
        """dst's q <- this batch's live q, device to device (dst takes this batch's size)"""
        _check(lib().rk_batch_copy_q(self._h, dst.handle))
        dst.n = self.n

    def advance(self) -> None:
        """Warm start on the device: q_prev <- q and seeds <- q (next frame of a sequence)."""
        _check(lib().rk_batch_advance(self._h))

    def iterates(self):
        """q before each traced step [B, trace_iters, na] and the damping it used [B, trace_iters]."""
        na = self._na
        qt = np.zeros((self.n, self.trace_iters, na), np.float32)
        lt = np.zeros((self.n, self.trace_iters), np.float32)
        _check(lib().rk_batch_get_iterates(self._h, _p(qt), _p(lt)))
        return qt, lt

    def metrics(self, pos_thr: float = 1e-3, ang_thr: float = 1e-2, noalloc: bool = False):
        """(pos_err [B, k], ang_err [B, k], success [B]) at the current q.  ``noalloc`` stages the
        device results in the library's static array (rk_batch_metrics_noalloc: no device
        allocation, at most RK_NOALLOC_MAX_PROBLEMS problems) instead of per-call buffers."""
        k = len(self.terms.pose_links)
        pe = np.zeros((self.n, k), np.float32)
        ae = np.zeros((self.n, k), np.float32)
        ok = np.zeros(self.n, np.int32)
        fn = lib().rk_batch_metrics_noalloc if noalloc else lib().rk_batch_metrics
        _check(fn(self._h, C.c_float(pos_thr), C.c_float(ang_thr), _p(pe), _p(ae), _p(ok)))
        return pe, ae, ok.astype(bool)


class Group:
    """Uneven batch: several Batches solved by one persistent kernel (rk_group); TrajBatches of any
    horizons join it, one launch per arm class beside the IK kernel."""

    def __init__(self, ctx: Context, batches: Sequence[Batch]):
        self.ctx, self.batches = ctx, list(batches)
        arr = (C.c_void_p * len(self.batches))(*[b.handle.value for b in self.batches])
        self._h = C.c_void_p()
        _check(lib().rk_group_create(ctx.handle, len(self.batches), arr, C.byref(self._h)))
        _LIVE["group"].add(self)
        ctx._children.add(self)
        for b in self.batches:
            b._groups.add(self)

    def close(self) -> None:
        if self._h:
            lib().rk_group_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_tuning(self, tpp: Optional[Sequence[int]] = None, wg_per_cu: int = 0, iters_per_launch: int = 0) -> None:
        t = None if tpp is None else _i32(list(tpp))
        _check(lib().rk_group_set_tuning(self._h, _p(t), int(wg_per_cu), int(iters_per_launch)))

    def autotune(self, iters: int, params: LMParams) -> List[int]:
        out = np.zeros(len(self.batches), np.int32)
        _check(lib().rk_group_autotune(self._h, int(iters), C.byref(params.c()), _p(out)))
        return out.tolist()

    def solve(self, iters: int, params: LMParams, use_graph: bool = True, from_seed: bool = False,
              sync: bool = True) -> float:
        """Solve every batch; returns the event-timed milliseconds, or 0.0 with ``sync=False`` (the
        solve is only enqueued: ``Context.sync`` or a later synchronous call waits for it)."""
        ms = C.c_float()
        flags = (USE_GRAPH if use_graph else 0) | (FROM_SEED if from_seed else 0) | (0 if sync else ASYNC)
        _check(lib().rk_group_solve(self._h, int(iters), C.byref(params.c()), flags, C.byref(ms)))
        return float(ms.value) if sync else 0.0

    def kernel_times(self) -> List[tuple]:
        """[(threads-per-problem code, ms, batch indices)] of each kernel of the last eager solve
        (TPP_TRAJ for a trajectory class)."""
        cap = 16
        tpp = np.zeros(cap, np.int32)
        ms = np.zeros(cap, np.float32)
        mask = np.zeros(cap, np.int32)
        n = C.c_int32()
        _check(lib().rk_group_kernel_times(self._h, cap, _p(tpp), _p(ms), _p(mask), C.byref(n)))
        return [(int(tpp[k]), float(ms[k]), [i for i in range(32) if (int(mask[k]) >> i) & 1])
                for k in range(min(n.value, cap))]


class Server:
    """Continuous-batching request server (rk_server): submit returns a ticket at once; results are
    collected with poll.  Requests of one term structure and iteration count that are queued when
    a launch stream frees up share one launch; launches on different streams overlap."""

    def __init__(self, ctx: Context, mem_budget: int, streams: int = 4, max_launch_problems: int = 4096,
                 pos_thr: float = 1e-3, ang_thr: float = 1e-2):
        self.ctx = ctx
        self._h = C.c_void_p()
        _check(lib().rk_server_create(ctx.handle, C.c_int64(int(mem_budget)), int(streams), int(max_launch_problems),
                                      C.c_float(pos_thr), C.c_float(ang_thr), C.byref(self._h)))
        _LIVE["server"].add(self)
        ctx._children.add(self)
        self._terms: List[tuple] = []  # (na, n_pose, n_pos)
        self._pending: Dict[int, tuple] = {}  # ticket -> (term, n)

    def close(self) -> None:
        if self._h:
            lib().rk_server_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def add_terms(self, spec, terms: TermSpec) -> int:
        na = spec.num_actuated_joints
        d, arrs = terms.desc(self.ctx.robot(spec), na, self.ctx)
        tid = C.c_int32()
        _check(lib().rk_server_add_terms(self._h, C.byref(d), C.byref(tid)))
        self._terms.append((na, len(terms.pose_links), len(terms.pos_links)))
        return tid.value

    def add_traj(self, tb, goal_weights=None) -> int:
        """register a trajectory problem set (workloads.TrajBatch shape: robot, frames, scene)"""
        d, arrs = traj_desc(self.ctx.robot(tb.spec), tb, goal_weights, self.ctx)
        tid = C.c_int32()
        _check(lib().rk_server_add_traj(self._h, C.byref(d), C.byref(tid)))
        self._terms.append((tb.T * tb.spec.num_actuated_joints, 1, 0))
        return tid.value

    def submit(self, term: int, iters: int, params: LMParams, q: Optional[np.ndarray] = None,
               pose_targets: Optional[np.ndarray] = None, pos_targets: Optional[np.ndarray] = None,
               q_prev: Optional[np.ndarray] = None, session: int = 0, n: Optional[int] = None) -> int:
        arrs = [None if x is None else _f32(x) for x in (q, pose_targets, pos_targets, q_prev)]
        if arrs[2] is not None and arrs[2].shape[-1] == 7:
            arrs[2] = _f32(arrs[2][..., :3])
        if n is None:
            n = next(x.shape[0] for x in arrs if x is not None)
        na = self._terms[term][0]
        for i in (0, 3):
            if arrs[i] is not None and arrs[i].size != int(n) * na:
                raise ValueError(f"q / q_prev must hold {n} x {na} values, got {arrs[i].shape}")
        req = RequestC(int(term), int(n), int(iters), params.c(), *[_p(x) for x in arrs], int(session))
        t = C.c_int64()
        _check(lib().rk_server_submit(self._h, C.byref(req), C.byref(t)))
        self._pending[t.value] = (int(term), int(n))
        return t.value

    def poll(self, ticket: int, wait_ms: int = -1) -> Optional[dict]:
        """Results of a ticket (None while it is still running with wait_ms >= 0)."""
        term, n = self._pending[ticket]
        na, npose, _ = self._terms[term]
        q = np.zeros((n, na), np.float32)
        cost = np.zeros(n, np.float32)
        ok = np.zeros(n, np.int32)
        pe = np.zeros((n, max(npose, 1)), np.float32)
        ae = np.zeros((n, max(npose, 1)), np.float32)
        lat = C.c_float()
        rc = lib().rk_server_poll(self._h, C.c_int64(int(ticket)), int(wait_ms), _p(q), _p(cost), _p(ok), _p(pe), _p(ae),
                                  C.byref(lat))
        if rc == RK_EBUSY:
            return None
        del self._pending[ticket]
        _check(rc)
        return dict(q=q, cost=cost, success=ok.astype(bool), pos_err=pe[:, :npose], ang_err=ae[:, :npose],
                    latency_ms=float(lat.value))

    def end_session(self, session: int) -> None:
        _check(lib().rk_server_end_session(self._h, C.c_int64(int(session))))

    def stats(self) -> dict:
        st = ServeStatsC()
        _check(lib().rk_server_stats(self._h, C.byref(st)))
        return {name: getattr(st, name) for name, _ in ServeStatsC._fields_}


class TrajBatch(Batch):
    """Trajectory problems (rk_traj_batch_create): q [B, T, na] per problem, a goal pose for the
    last frame, finite-difference smoothness and a sphere-SDF collision hinge against a primitive
    scene; solved by the banded trajectory kernel."""

    def __init__(self, ctx: Context, tb, capacity: int, trace_iters: int = 0, goal_weights=None):
        self.ctx, self.spec = ctx, tb.spec
        self.T, self.robot_na = tb.T, tb.spec.num_actuated_joints
        self._na = self.T * self.robot_na
        self.terms = TermSpec(pose_links=[tb.goal_link], pose_weights=np.ones((1, 6), np.float32))
        self.capacity, self.trace_iters = int(capacity), int(trace_iters)
        d, self._arrs = traj_desc(ctx.robot(tb.spec), tb, goal_weights, ctx)
        self._h = C.c_void_p()
        _check(lib().rk_traj_batch_create(ctx.handle, C.byref(d), self.capacity, self.trace_iters, C.byref(self._h)))
        _LIVE["batch"].add(self)
        ctx._children.add(self)
        self._groups = weakref.WeakSet()
        self.n = 0

    def set(self, q: Optional[np.ndarray] = None, goals: Optional[np.ndarray] = None, n: Optional[int] = None):
        qq = None if q is None else _f32(q).reshape(len(q), -1)
        gg = None if goals is None else _f32(goals).reshape(len(goals), 1, 7)
        super().set(q=qq, pose_targets=gg, n=n)

    def get(self, traces: bool = False):
        out = super().get(traces)
        out["q"] = out["q"].reshape(self.n, self.T, self.robot_na)
        return out

    def iterates(self):
        qt, lt = super().iterates()
        return qt.reshape(self.n, self.trace_iters, self.T, self.robot_na), lt

    def metrics(self, *args):
        pe, ae, sd = (np.zeros(self.n, np.float32) for _ in range(3))
        _check(lib().rk_traj_metrics(self._h, _p(pe), _p(ae), _p(sd)))
        return pe, ae, sd

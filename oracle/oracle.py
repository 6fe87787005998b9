"""ctypes wrapper over the CPU oracle (liboracle_f32.so / liboracle_f64.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg, to check results.  The product package (robokit_amd) never imports it.

The C source (rk_oracle.c) restates the reference's Warp-CPU LM path; see its header for
the file:line each function follows.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
from dataclasses import dataclass
from pathlib import Path
from typing import List, Optional, Sequence

import numpy as np

HERE = Path(__file__).resolve().parent
_LIBS = {}

FRAME, POSITION, LIMIT, SMOOTH, ROTATION, COLLISION = 0, 1, 2, 3, 4, 5
PENALTY = {"smooth": 0, "plain": 1, "surface_distance": 2}


def build() -> None:
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)


def _lib(dtype: str):
    if dtype not in _LIBS:
        path = HERE / f"liboracle_{dtype}.so"
        if not path.exists():
            build()
        _LIBS[dtype] = C.CDLL(str(path))
    return _LIBS[dtype]


class _Robot(C.Structure):
    _fields_ = [
        ("nj", C.c_int), ("nl", C.c_int), ("na", C.c_int),
        ("joint_type", C.c_void_p), ("joint_axis", C.c_void_p), ("parent_T", C.c_void_p),
        ("parent_joint", C.c_void_p), ("act_idx", C.c_void_p), ("mimic_idx", C.c_void_p),
        ("mimic_mul", C.c_void_p), ("mimic_off", C.c_void_p), ("link_parent_joint", C.c_void_p),
        ("limits", C.c_void_p),
    ]


class _Term(C.Structure):
    _fields_ = [
        ("kind", C.c_int), ("nframes", C.c_int), ("links", C.c_void_p), ("targets", C.c_void_p),
        ("weights", C.c_void_p), ("q_prev", C.c_void_p), ("scene", C.c_void_p),
    ]


class _Scene(C.Structure):
    _fields_ = [
        ("nsph", C.c_int), ("sph_link", C.c_void_p), ("sph_c", C.c_void_p), ("sph_r", C.c_void_p),
        ("sph_w", C.c_void_p), ("margin", C.c_float), ("penalty", C.c_int), ("nobs_s", C.c_int),
        ("obs_s", C.c_void_p), ("nobs_b", C.c_int), ("obs_b", C.c_void_p), ("ngeom", C.c_int), ("geoms", C.c_void_p),
    ]


class _Geom(C.Structure):
    _fields_ = [("kind", C.c_int), ("values", C.c_void_p), ("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int),
                ("origin", C.c_float * 3), ("voxel", C.c_float), ("lo", C.c_float * 3), ("hi", C.c_float * 3),
                ("pad", C.c_float), ("verts", C.c_void_p), ("nv", C.c_int), ("faces", C.c_void_p), ("nf", C.c_int)]


def _geoms(geoms, keep: list):
    """(count, pointer) of the C geometry table of robokit_amd.geom volumes / meshes"""
    geoms = list(geoms or [])
    if not geoms:
        return 0, None
    arr = (_Geom * len(geoms))()
    for i, g in enumerate(geoms):
        if hasattr(g, "values"):  # a volume
            v = np.ascontiguousarray(g.values, np.float32)
            nz, ny, nx = v.shape
            keep.append(v)
            arr[i] = _Geom(0, _ptr(v), nx, ny, nz, (C.c_float * 3)(*g.origin), float(g.voxel_size),
                           (C.c_float * 3)(*g.aabb_min), (C.c_float * 3)(*g.aabb_max), float(g.padding), None, 0, None, 0)
        else:  # a mesh
            v = np.ascontiguousarray(g.vertices, np.float32)
            f = np.ascontiguousarray(g.faces, np.int32)
            keep.extend([v, f])
            z3 = (C.c_float * 3)(0, 0, 0)
            arr[i] = _Geom(1, None, 0, 0, 0, z3, 0.0, z3, z3, 0.0, _ptr(v), len(v), _ptr(f), len(f))
    keep.append(arr)
    return len(geoms), C.cast(arr, C.c_void_p)


def scene_query(points: np.ndarray, spheres=None, boxes=None, geoms=(), dtype: str = "f32"):
    """(signed distance [n], normal [n, 3]) of a scene of obstacle spheres, boxes and volumes /
    meshes at points [n, 3]"""
    lib, rt = _lib(dtype), _real(dtype)
    pts = np.ascontiguousarray(np.asarray(points).reshape(-1, 3), rt)
    sp = np.ascontiguousarray(np.zeros((0, 4)) if spheres is None else spheres, np.float32).reshape(-1, 4)
    bx = np.ascontiguousarray(np.zeros((0, 6)) if boxes is None else boxes, np.float32).reshape(-1, 6)
    keep: list = []
    ng, gp = _geoms(geoms, keep)
    n = len(pts)
    dist = np.zeros(n, rt)
    nrm = np.zeros((n, 3), rt)
    lib.ork_scene_query(len(sp), _ptr(sp), len(bx), _ptr(bx), ng, gp, n, _ptr(pts), _ptr(dist), _ptr(nrm))
    return dist, nrm


@dataclass
class CollisionScene:
    """SceneCollisionTask data of an oracle COLLISION term: robot spheres (link, centre in the link
    frame, radius, row weight) against obstacle spheres [n, 4] and axis-aligned boxes [n, 6]."""

    sphere_links: np.ndarray
    sphere_centers: np.ndarray
    sphere_radii: np.ndarray
    weights: np.ndarray
    margin: float = 0.02
    penalty: str = "smooth"
    obstacle_spheres: np.ndarray = None
    obstacle_boxes: np.ndarray = None
    geoms: Sequence = ()  # volumes / meshes after the boxes


class _LMParams(C.Structure):
    _fields_ = [(n, C.c_float) for n in
                ("lm_lambda", "lambda_factor", "lambda_min", "lambda_max", "rho_min", "gain_ratio_epsilon")]


def _ptr(a: Optional[np.ndarray]):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


class OracleRobot:
    """Holds the float32/int32 arrays the C struct points at (keeps them alive)."""

    def __init__(self, spec):
        self.spec = spec
        self.arrays = dict(
            joint_type=np.ascontiguousarray(spec.joint_types, np.int32),
            joint_axis=np.ascontiguousarray(spec.joint_axes, np.float32),
            parent_T=np.ascontiguousarray(spec.parent_joint_transforms, np.float32),
            parent_joint=np.ascontiguousarray(spec.parent_joint_indices, np.int32),
            act_idx=np.ascontiguousarray(spec.actuated_joint_indices, np.int32),
            mimic_idx=np.ascontiguousarray(spec.mimic_actuated_joint_indices, np.int32),
            mimic_mul=np.ascontiguousarray(spec.mimic_multipliers, np.float32),
            mimic_off=np.ascontiguousarray(spec.mimic_offsets, np.float32),
            link_parent_joint=np.ascontiguousarray(spec.link_parent_joint_indices, np.int32),
            limits=np.ascontiguousarray(spec.actuated_joint_limits_f32, np.float32),
        )
        self.struct = _Robot(spec.num_joints, spec.num_links, spec.num_actuated_joints,
                             *[_ptr(self.arrays[k]) for k in ("joint_type", "joint_axis", "parent_T", "parent_joint",
                                                              "act_idx", "mimic_idx", "mimic_mul", "mimic_off",
                                                              "link_parent_joint", "limits")])


@dataclass
class Term:
    kind: int
    links: Sequence[int] = ()
    targets: Optional[np.ndarray] = None  # FRAME [B, k, 7] poses; POSITION [B, k, 3] points (or poses)
    weights: Optional[np.ndarray] = None
    q_prev: Optional[np.ndarray] = None  # [B, na]
    scene: Optional[CollisionScene] = None  # COLLISION


def _terms(terms: Sequence[Term], keep: list):
    arr = (_Term * max(1, len(terms)))()
    for i, t in enumerate(terms):
        links = np.ascontiguousarray(t.links, np.int32)
        tg = None if t.targets is None else np.ascontiguousarray(t.targets, np.float32)
        if tg is not None and tg.shape[-1] == 3:  # points: the C side reads 7-float pose records
            pad = np.zeros(tg.shape[:-1] + (4,), np.float32)
            pad[..., 0] = 1.0
            tg = np.ascontiguousarray(np.concatenate([tg, pad], -1))
        if tg is not None and tg.shape[-1] != 7:
            raise ValueError(f"targets must end in 3 or 7 floats, got {tg.shape}")
        w = None if t.weights is None else np.ascontiguousarray(t.weights, np.float32)
        qp = None if t.q_prev is None else np.ascontiguousarray(t.q_prev, np.float32)
        keep.extend([links, tg, w, qp])
        sc_ptr = None
        if t.kind == COLLISION:
            sc = t.scene
            a = [np.ascontiguousarray(sc.sphere_links, np.int32), np.ascontiguousarray(sc.sphere_centers, np.float32),
                 np.ascontiguousarray(sc.sphere_radii, np.float32),
                 np.ascontiguousarray(np.broadcast_to(np.asarray(sc.weights, np.float32), (len(sc.sphere_links),))),
                 np.ascontiguousarray(np.zeros((0, 4)) if sc.obstacle_spheres is None else sc.obstacle_spheres,
                                      np.float32).reshape(-1, 4),
                 np.ascontiguousarray(np.zeros((0, 6)) if sc.obstacle_boxes is None else sc.obstacle_boxes,
                                      np.float32).reshape(-1, 6)]
            ng, gp = _geoms(sc.geoms, keep)
            st = _Scene(len(a[0]), _ptr(a[0]), _ptr(a[1]), _ptr(a[2]), _ptr(a[3]), float(sc.margin),
                        PENALTY[sc.penalty], len(a[4]), _ptr(a[4]), len(a[5]), _ptr(a[5]), ng, gp)
            keep.extend(a + [st])
            sc_ptr = C.cast(C.pointer(st), C.c_void_p)
        arr[i] = _Term(t.kind, len(links), _ptr(links), _ptr(tg), _ptr(w), _ptr(qp), sc_ptr)
    return arr


def _real(dtype: str):
    return np.float32 if dtype == "f32" else np.float64


def fk(spec, q: np.ndarray, base: Optional[np.ndarray] = None, dtype: str = "f32"):
    lib, rt = _lib(dtype), _real(dtype)
    rb = OracleRobot(spec)
    q = np.ascontiguousarray(np.atleast_2d(q), rt)
    B = q.shape[0]
    bs = None if base is None else np.ascontiguousarray(np.broadcast_to(base, (B, 7)), rt)
    Tl = np.zeros((B, spec.num_links, 7), rt)
    Tj = np.zeros((B, spec.num_joints, 7), rt)
    lib.ork_fk(C.byref(rb.struct), B, _ptr(q), _ptr(bs), _ptr(Tl), _ptr(Tj))
    return Tl, Tj


def link_jacobian(spec, q: np.ndarray, link: int, frame: str, base: Optional[np.ndarray] = None, dtype: str = "f32"):
    lib, rt = _lib(dtype), _real(dtype)
    rb = OracleRobot(spec)
    q = np.ascontiguousarray(np.atleast_2d(q), rt)
    B = q.shape[0]
    bs = None if base is None else np.ascontiguousarray(np.broadcast_to(base, (B, 7)), rt)
    cols = spec.num_actuated_joints + (6 if base is not None else 0)
    J = np.zeros((B, 6, cols), rt)
    lib.ork_link_jacobian(C.byref(rb.struct), B, _ptr(q), _ptr(bs), int(link), 0 if frame == "spatial" else 1, _ptr(J))
    return J


def se3(op: str, *args, eps: float = 1e-4, dtype: str = "f32"):
    lib, rt = _lib(dtype), _real(dtype)
    a = [np.ascontiguousarray(np.atleast_2d(x), rt) for x in args]
    n = a[0].shape[0]
    out_w = {"compose": 7, "inverse": 7, "apply": 3, "exp": 7, "log": 6, "adjoint": 36, "jlog": 36}[op]
    out = np.zeros((n, out_w), rt)
    fn = getattr(lib, f"ork_se3_{op}")
    if op in ("exp", "log", "jlog"):
        cast = C.c_float if dtype == "f32" else C.c_double
        fn(n, _ptr(a[0]), cast(eps), _ptr(out))
    else:
        fn(n, *[_ptr(x) for x in a], _ptr(out))
    return out.reshape((n, 6, 6)) if out_w == 36 else out


def residual_jacobian(spec, terms: Sequence[Term], b: int, q: np.ndarray, dtype: str = "f32"):
    """Stacked residual r[R], Jacobian J[R, na], H = sum J^T J, g = sum J^T r, cost at q for problem b."""
    lib, rt = _lib(dtype), _real(dtype)
    rb = OracleRobot(spec)
    keep: list = []
    ta = _terms(terms, keep)
    na = spec.num_actuated_joints
    R = sum(6 * len(t.links) if t.kind == FRAME else 3 * len(t.links) if t.kind in (POSITION, ROTATION)
            else len(t.scene.sphere_links) if t.kind == COLLISION else na for t in terms)
    q = np.ascontiguousarray(q, rt)
    r = np.zeros(R, rt)
    J = np.zeros((R, na), rt)
    H = np.zeros((na, na), rt)
    g = np.zeros(na, rt)
    cost = (C.c_float if dtype == "f32" else C.c_double)()
    lib.ork_residual_jacobian(C.byref(rb.struct), len(terms), ta, int(b), _ptr(q), None, _ptr(r), _ptr(J), _ptr(H),
                              _ptr(g), C.byref(cost))
    return r, J, H, g, float(cost.value)


def damped_solve(H: np.ndarray, g: np.ndarray, lam: float, dtype: str = "f32"):
    lib, rt = _lib(dtype), _real(dtype)
    D = H.shape[0]
    H = np.ascontiguousarray(H, rt)
    g = np.ascontiguousarray(g, rt)
    delta = np.zeros(D, rt)
    pred = (C.c_float if dtype == "f32" else C.c_double)()
    cast = C.c_float if dtype == "f32" else C.c_double
    ok = lib.ork_damped_solve(D, _ptr(H), _ptr(g), cast(lam), _ptr(delta), C.byref(pred))
    return delta, float(pred.value), bool(ok)


@dataclass
class LMParams:
    """LMOptimizerConfig fields (opt/lm_optimizer.py:502-523)."""

    lm_lambda: float = 1e-3
    lambda_factor: float = 2.0
    lambda_min: float = 1e-5
    lambda_max: float = 1e10
    rho_min: float = 1e-3
    gain_ratio_epsilon: float = 1e-8

    def struct(self):
        return _LMParams(self.lm_lambda, self.lambda_factor, self.lambda_min, self.lambda_max, self.rho_min,
                         self.gain_ratio_epsilon)


def accept(cost, cost_prop, pred, lam, params: LMParams, dtype: str = "f32"):
    lib, rt = _lib(dtype), _real(dtype)
    cost = np.array(cost, rt, ndmin=1)
    cp = np.array(cost_prop, rt, ndmin=1)
    pr = np.array(pred, rt, ndmin=1)
    lam = np.array(lam, rt, ndmin=1)
    acc = np.zeros(cost.shape[0], np.int32)
    p = params.struct()
    lib.ork_accept(cost.shape[0], _ptr(cost), _ptr(cp), _ptr(pr), C.byref(p), _ptr(lam), _ptr(acc))
    return cost, lam, acc


def _slice_terms(terms: Sequence[Term], lo: int, hi: int) -> List[Term]:
    """the terms of problems lo..hi-1 (per-problem targets and smoothness references sliced)"""
    return [Term(t.kind, t.links, None if t.targets is None else np.asarray(t.targets)[lo:hi], t.weights,
                 None if t.q_prev is None else np.asarray(t.q_prev)[lo:hi], t.scene) for t in terms]


def _chunks(B: int, threads: int):
    threads = max(1, min(int(threads), B))
    cuts = np.linspace(0, B, threads + 1).round().astype(int)
    return [(int(a), int(b)) for a, b in zip(cuts[:-1], cuts[1:]) if b > a]


def _parallel(B: int, threads: int, fn):
    """run fn(lo, hi) over contiguous chunks of the batch on `threads` host threads (each chunk is
    one C call, which releases the GIL) and concatenate the per-key results in order"""
    parts = _chunks(B, threads)
    if len(parts) == 1:
        return fn(*parts[0])
    from concurrent.futures import ThreadPoolExecutor

    with ThreadPoolExecutor(len(parts)) as ex:
        outs = list(ex.map(lambda c: fn(*c), parts))
    return {k: (None if outs[0][k] is None else np.concatenate([o[k] for o in outs])) for k in outs[0]}


def lm_solve(spec, terms: Sequence[Term], q_seed: np.ndarray, iters: int, params: LMParams, dtype: str = "f32",
             traces: bool = True, threads: int = 1):
    """Fixed-iteration LM over a batch; returns dict(q, cost, cost_trace[B, iters+1], accept[B, iters], lam[B, iters]).
    Problems are independent: ``threads`` > 1 solves contiguous chunks of the batch at once (the
    same per-problem arithmetic, so the same bits)."""
    if threads > 1:
        return _parallel(len(q_seed), threads, lambda lo, hi: lm_solve(spec, _slice_terms(terms, lo, hi), q_seed[lo:hi],
                                                                     iters, params, dtype, traces))
    lib, rt = _lib(dtype), _real(dtype)
    rb = OracleRobot(spec)
    keep: list = []
    ta = _terms(terms, keep)
    q = np.ascontiguousarray(q_seed, rt).copy()
    B = q.shape[0]
    cost = np.zeros(B, rt)
    ct = np.zeros((B, iters + 1), rt) if traces else None
    at = np.zeros((B, iters), np.int32) if traces else None
    lt = np.zeros((B, iters), rt) if traces else None
    p = params.struct()
    lib.ork_lm_solve(C.byref(rb.struct), len(terms), ta, C.byref(p), B, int(iters), _ptr(q), None, _ptr(cost), _ptr(ct),
                     _ptr(at), _ptr(lt))
    return dict(q=q, cost=cost, cost_trace=ct, accept=at, lam=lt)


def lm_step(spec, terms: Sequence[Term], q: np.ndarray, lam: np.ndarray, params: LMParams, dtype: str = "f32",
            rows: Optional[np.ndarray] = None, threads: int = 1):
    """One LM step per problem from iterate q[B, na] with damping lam[B].  ``rows`` maps each
    entry to the problem index its targets belong to (default: entry i is problem i)."""
    lib, rt = _lib(dtype), _real(dtype)
    rb = OracleRobot(spec)
    q = np.ascontiguousarray(q, rt)
    B = q.shape[0]
    if rows is not None:
        terms = [Term(t.kind, t.links, None if t.targets is None else np.asarray(t.targets)[rows], t.weights,
                      None if t.q_prev is None else np.asarray(t.q_prev)[rows], t.scene) for t in terms]
    if threads > 1:
        lam_a = np.asarray(lam)
        return _parallel(B, threads, lambda lo, hi: lm_step(spec, _slice_terms(terms, lo, hi), q[lo:hi], lam_a[lo:hi],
                                                            params, dtype))
    keep: list = []
    ta = _terms(terms, keep)
    lam = np.ascontiguousarray(lam, rt)
    out = {k: np.zeros(B, rt) for k in ("cost", "cost_prop", "pred", "lam_new")}
    acc = np.zeros(B, np.int32)
    qn = np.zeros_like(q)
    p = params.struct()
    lib.ork_lm_step(C.byref(rb.struct), len(terms), ta, C.byref(p), B, _ptr(q), _ptr(lam), _ptr(out["cost"]),
                    _ptr(out["cost_prop"]), _ptr(out["pred"]), _ptr(acc), _ptr(qn), _ptr(out["lam_new"]))
    out.update(acc=acc, q_new=qn)
    return out


def ik_metrics(spec, q: np.ndarray, links: Sequence[int], targets: np.ndarray, pos_thr: float = 1e-3,
               ang_thr: float = 1e-2, dtype: str = "f32"):
    lib, rt = _lib(dtype), _real(dtype)
    rb = OracleRobot(spec)
    q = np.ascontiguousarray(q, rt)
    B = q.shape[0]
    links_a = np.ascontiguousarray(links, np.int32)
    tg = np.ascontiguousarray(targets, np.float32)
    pe = np.zeros((B, len(links)), rt)
    ae = np.zeros((B, len(links)), rt)
    ok = np.zeros(B, np.int32)
    lib.ork_ik_metrics(C.byref(rb.struct), B, _ptr(q), None, len(links), _ptr(links_a), _ptr(tg), C.c_float(pos_thr),
                       C.c_float(ang_thr), _ptr(pe), _ptr(ae), _ptr(ok))
    return pe, ae, ok.astype(bool)


def ik_terms(links: Sequence[int], targets: np.ndarray, na: int, limit_weight: float = 1.0,
             pos_w: float = 1.0, ori_w: float = 1.0) -> List[Term]:
    """FrameTask(log, w_pos, w_ori) over all target links + PositionLimit(abs, weight)."""
    k = len(links)
    w = np.tile(np.array([pos_w] * 3 + [ori_w] * 3, np.float32), k)
    return [Term(FRAME, links, targets, w), Term(LIMIT, (), None, np.full(na, limit_weight, np.float32))]


# ----------------------------------------------------------------------------- trajectories
class _Traj(C.Structure):
    _fields_ = [("T", C.c_int), ("goal_link", C.c_int), ("goal_targets", C.c_void_p), ("goal_w", C.c_void_p),
                ("wv", C.c_float), ("wa", C.c_float), ("dt", C.c_float), ("nsph", C.c_int),
                ("sph_link", C.c_void_p), ("sph_c", C.c_void_p), ("sph_r", C.c_void_p), ("coll_w", C.c_float),
                ("margin", C.c_float), ("nobs_s", C.c_int), ("obs_s", C.c_void_p), ("nobs_b", C.c_int),
                ("obs_b", C.c_void_p), ("lim_orders", C.c_int), ("lim_w", C.c_void_p), ("lim_lo", C.c_void_p),
                ("lim_hi", C.c_void_p), ("cg_iters", C.c_int), ("ngeom", C.c_int), ("geoms", C.c_void_p)]


class OracleTraj:
    """The C trajectory description of a workloads.TrajBatch (keeps its arrays alive)."""

    def __init__(self, tb, goal_w=None, wv=None, wa=None, coll_w=None, margin=None):
        sc = tb.scene
        self.keep = [
            np.ascontiguousarray(tb.goals, np.float32),
            np.ascontiguousarray(np.ones(6) if goal_w is None else goal_w, np.float32),
            np.ascontiguousarray(sc.sphere_links, np.int32),
            np.ascontiguousarray(sc.sphere_centers, np.float32),
            np.ascontiguousarray(sc.sphere_radii, np.float32),
            np.ascontiguousarray(sc.obstacle_spheres, np.float32).reshape(-1, 4),
            np.ascontiguousarray(sc.obstacle_boxes, np.float32).reshape(-1, 6),
        ]
        na = tb.spec.num_actuated_joints
        orders = int(getattr(tb, "lim_orders", 0) or 0)
        lim = [np.ascontiguousarray(np.broadcast_to(np.asarray(getattr(tb, n), np.float32), (3, na)))
               for n in ("lim_w", "lim_lo", "lim_hi")] if orders else [None, None, None]
        self.keep += lim
        k = self.keep
        self.struct = _Traj(tb.T, tb.goal_link, _ptr(k[0]), _ptr(k[1]), tb.wv if wv is None else wv,
                            tb.wa if wa is None else wa, tb.dt, len(k[2]), _ptr(k[2]), _ptr(k[3]), _ptr(k[4]),
                            tb.coll_w if coll_w is None else coll_w, tb.margin if margin is None else margin,
                            len(k[5]), _ptr(k[5]), len(k[6]), _ptr(k[6]), orders, _ptr(lim[0]), _ptr(lim[1]),
                            _ptr(lim[2]), int(getattr(tb, "cg_iters", 0) or 0),
                            *_geoms(getattr(sc, "geoms", None), self.keep))
        self.tb = tb

    def rows(self):
        tb, na = self.tb, self.tb.spec.num_actuated_joints
        s = self.struct
        lim = sum((tb.T - o) * na for o in range(3) if (s.lim_orders >> o) & 1)
        return 6 + (tb.T - 1) * na * (s.wv != 0) + (tb.T - 1) * na * (s.wa != 0) + tb.T * s.nsph + lim


def traj_residual_jacobian(ot: OracleTraj, b: int, q: np.ndarray, dtype: str = "f32", jac: bool = True):
    """r [R], dense J [R, T*na] (or None), H, g, cost_terms [7] (goal, velocity, acceleration, collision,
    bounds of orders 0, 1, 2) of problem b at the trajectory q [T, na]."""
    lib, rt = _lib(dtype), _real(dtype)
    rb = OracleRobot(ot.tb.spec)
    D = ot.tb.T * ot.tb.spec.num_actuated_joints
    R = ot.rows()
    q = np.ascontiguousarray(q, rt).reshape(-1)
    r = np.zeros(R, rt)
    J = np.zeros((R, D), rt) if jac else None
    H = np.zeros((D, D), rt)
    g = np.zeros(D, rt)
    ct = np.zeros(7, rt)
    lib.ork_traj_residual_jacobian(C.byref(rb.struct), C.byref(ot.struct), int(b), _ptr(q), _ptr(r), _ptr(J), _ptr(H),
                                   _ptr(g), _ptr(ct))
    return r, J, H, g, ct


def traj_lm_step(ot: OracleTraj, q: np.ndarray, lam: np.ndarray, params: LMParams, dtype: str = "f32", rows=None):
    """One trajectory LM step per problem from q [B, T, na] and lam [B]; ``rows`` maps entries to the
    problem whose goal they use."""
    lib, rt = _lib(dtype), _real(dtype)
    rb = OracleRobot(ot.tb.spec)
    q = np.ascontiguousarray(q, rt)
    B = q.shape[0]
    ot2 = ot
    if rows is not None:
        import copy
        tb = copy.copy(ot.tb)
        tb.goals = np.asarray(ot.tb.goals)[rows]
        ot2 = OracleTraj(tb, goal_w=ot.keep[1], wv=ot.struct.wv, wa=ot.struct.wa, coll_w=ot.struct.coll_w,
                         margin=ot.struct.margin)
    lam = np.ascontiguousarray(lam, rt)
    out = {k: np.zeros(B, rt) for k in ("cost", "cost_prop", "pred", "lam_new")}
    acc = np.zeros(B, np.int32)
    qn = np.zeros_like(q)
    p = params.struct()
    lib.ork_traj_lm_step(C.byref(rb.struct), C.byref(ot2.struct), C.byref(p), B, _ptr(q), _ptr(lam), _ptr(out["cost"]),
                         _ptr(out["cost_prop"]), _ptr(out["pred"]), _ptr(acc), _ptr(qn), _ptr(out["lam_new"]))
    out.update(acc=acc, q_new=qn)
    return out


def traj_delta(ot: OracleTraj, b: int, q: np.ndarray, lam: float, dtype: str = "f32"):
    """(delta [T*na], pred) of problem b's linear step at q with damping lam: conjugate gradients
    when the problem set has cg_iters > 0, else the damped Cholesky"""
    lib, rt = _lib(dtype), _real(dtype)
    rb = OracleRobot(ot.tb.spec)
    q = np.ascontiguousarray(q, rt).reshape(-1)
    delta = np.zeros(q.shape[0], rt)
    pred = (C.c_float if dtype == "f32" else C.c_double)()
    cast = C.c_float if dtype == "f32" else C.c_double
    lib.ork_traj_delta(C.byref(rb.struct), C.byref(ot.struct), int(b), _ptr(q), cast(lam), _ptr(delta), C.byref(pred))
    return delta, float(pred.value)


def traj_lm_solve(ot: OracleTraj, q_seed: np.ndarray, iters: int, params: LMParams, dtype: str = "f32"):
    lib, rt = _lib(dtype), _real(dtype)
    rb = OracleRobot(ot.tb.spec)
    q = np.ascontiguousarray(q_seed, rt).copy()
    B = q.shape[0]
    cost = np.zeros(B, rt)
    ct = np.zeros((B, iters + 1), rt)
    at = np.zeros((B, iters), np.int32)
    p = params.struct()
    lib.ork_traj_lm_solve(C.byref(rb.struct), C.byref(ot.struct), C.byref(p), B, int(iters), _ptr(q), _ptr(cost),
                          _ptr(ct), _ptr(at))
    return dict(q=q, cost=cost, cost_trace=ct, accept=at)


def traj_metrics(ot: OracleTraj, q: np.ndarray, dtype: str = "f32"):
    """goal position / angle error at the last frame [B] and the smallest sphere-scene distance [B]"""
    lib, rt = _lib(dtype), _real(dtype)
    rb = OracleRobot(ot.tb.spec)
    q = np.ascontiguousarray(q, rt)
    B = q.shape[0]
    pe, ae, sd = np.zeros(B, rt), np.zeros(B, rt), np.zeros(B, rt)
    lib.ork_traj_metrics(C.byref(rb.struct), C.byref(ot.struct), B, _ptr(q), _ptr(pe), _ptr(ae), _ptr(sd))
    return pe, ae, sd

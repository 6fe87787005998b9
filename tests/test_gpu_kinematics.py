"""GPU parity of the device SE(3) ops, FK and link Jacobians (through the C ABI).

Checked against the CPU oracle (restatement of lie/se3_kernels.py, robo/robot_kernels.py)
and against the reference's own ground truth (tests/robo/fk_ground_truth.json,
tests/robo/jacobian_ground_truth.json, tests/lie/test_se3.py:39-80 constants), FP32,
tolerance 1e-5 as in the reference tests.
"""
import json

import numpy as np
import pytest

from oracle import oracle as O
from robokit_amd import workloads as W
from tests.conftest import GOLDEN

pytestmark = pytest.mark.gpu


def _pose_close(a, b, atol):
    """poses equal up to the quaternion double cover"""
    d1 = np.abs(a - b).max(-1)
    flip = np.concatenate([b[..., :3], -b[..., 3:]], -1)
    d2 = np.abs(a - flip).max(-1)
    return np.minimum(d1, d2).max() <= atol, np.minimum(d1, d2).max()


def _rand_poses(rng, n):
    q = rng.normal(size=(n, 4))
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    q[q[:, 0] < 0] *= -1
    return np.concatenate([rng.uniform(-2, 2, (n, 3)), q], 1).astype(np.float32)


def test_se3_golden_constants(ctx):
    c = json.loads((GOLDEN / "se3_constants.json").read_text())
    T = np.array([c["T_X_90"]], np.float32)
    np.testing.assert_allclose(ctx.se3("inverse", T)[0], c["T_X_90_inverse"], atol=1e-6)
    np.testing.assert_allclose(ctx.se3("log", T)[0], c["T_X_90_log"], atol=1e-5)
    np.testing.assert_allclose(ctx.se3("adjoint", T)[0], np.array(c["T_X_90_adjoint"]), atol=1e-6)
    np.testing.assert_allclose(ctx.se3("log", np.array([c["T_TRANS"]], np.float32))[0], c["T_TRANS_log"], atol=1e-6)
    ident = np.array([[0, 0, 0, 1, 0, 0, 0]], np.float32)
    np.testing.assert_allclose(ctx.se3("compose", T, ident)[0], T[0], atol=1e-6)
    np.testing.assert_allclose(ctx.se3("apply", T, np.array([[1.0, 0, 0]], np.float32))[0], [2.0, 2.0, 3.0], atol=1e-6)


@pytest.mark.parametrize("op", ["compose", "inverse", "log", "jlog", "adjoint"])
def test_se3_ops_vs_oracle(ctx, op):
    rng = np.random.default_rng(7)
    a = _rand_poses(rng, 512)
    b = _rand_poses(rng, 512)
    # include near-identity rotations (Taylor branches of log / jlog)
    a[:64, 3:] = np.array([1, 0, 0, 0], np.float32) + rng.normal(0, 1e-4, (64, 4)).astype(np.float32)
    a[:64, 3:] /= np.linalg.norm(a[:64, 3:], axis=1, keepdims=True)
    args = (a, b) if op == "compose" else (a,)
    gpu = ctx.se3(op, *args)
    ora = O.se3(op, *args)
    ora64 = O.se3(op, *args, dtype="f64")
    if op in ("compose", "inverse"):
        ok, err = _pose_close(gpu, ora64, 2e-6)
    else:
        err = np.abs(gpu - ora64).max() / max(1.0, np.abs(ora64).max())
        ok = err <= 2e-5
    assert ok, f"{op}: max err {err}"
    # the FP32 oracle is within the same bound of the FP64 twin
    assert np.abs(gpu.astype(np.float64) - ora).max() <= 5e-5 * max(1.0, np.abs(ora).max())


def test_fk_ground_truth(ctx, panda_urdf_spec):
    gt = json.loads((GOLDEN / "fk_ground_truth.json").read_text())
    spec = panda_urdf_spec
    for robot in gt["robots"]:
        if robot["robot_description_name"] != "panda_description":
            continue
        for cfg in robot["configurations"]:
            base = None if cfg["T_world_base"] is None else np.array(cfg["T_world_base"], np.float32)
            T = ctx.fk(spec, np.array(cfg["qpos"], np.float32), T_world_base=base)[0]
            for lp in cfg["link_poses"]:
                if lp["link_name"] not in spec.link_names:
                    continue  # panda_hand_tcp is not in the fixture URDF
                i = spec.link_names.index(lp["link_name"])
                ok, err = _pose_close(T[i][None], np.array([lp["xyz_wxyz"]]), 1e-5)
                assert ok, (cfg["name"], lp["link_name"], err)


@pytest.mark.parametrize("robot", ["panda_urdf", "ur10", "humanoid"])
def test_fk_vs_oracle(ctx, panda_urdf_spec, robot):
    spec = {"panda_urdf": panda_urdf_spec, "ur10": W.ur10(), "humanoid": W.humanoid29(0)}[robot]
    rng = np.random.default_rng(3)
    q = W.uniform_q(rng, spec, 256)
    gpu = ctx.fk(spec, q)
    ora, _ = O.fk(spec, q, dtype="f64")
    ok, err = _pose_close(gpu, ora, 1e-5 * 4)
    assert ok, err


def test_jacobian_ground_truth(ctx, panda_urdf_spec):
    gt = json.loads((GOLDEN / "jacobian_ground_truth.json").read_text())
    spec = panda_urdf_spec
    n = nb = 0
    for robot in gt["robots"]:
        if robot["robot_description_name"] != "panda_description":
            continue
        for cfg in robot["configurations"]:
            base = None if cfg["T_world_base"] is None else np.array(cfg["T_world_base"], np.float32)
            q = np.array(cfg["qpos"], np.float32)
            for lj in cfg["link_jacobians"]:
                if lj["link_name"] not in spec.link_names:
                    continue
                link = spec.link_names.index(lj["link_name"])
                for frame in ("body", "spatial"):
                    J = ctx.link_jacobian(spec, q, link, frame, T_world_base=base)[0]
                    msg = f"{cfg['name']} {lj['link_name']} {frame}"
                    if base is None:
                        np.testing.assert_allclose(J, np.array(lj[f"{frame}_J_joints"]), atol=1e-5, err_msg=msg)
                    else:  # floating base: the six base columns first, then the joints
                        np.testing.assert_allclose(J[:, :6], np.array(lj[f"{frame}_J_base"]), atol=1e-5, err_msg=msg)
                        np.testing.assert_allclose(J[:, 6:], np.array(lj[f"{frame}_J_joints"]), atol=1e-5, err_msg=msg)
                        nb += 1
                    n += 1
    assert n > 0 and nb > 0


@pytest.mark.parametrize("robot", ["ur10", "humanoid"])
def test_jacobian_vs_oracle(ctx, robot):
    spec = {"ur10": W.ur10(), "humanoid": W.humanoid29(0)}[robot]
    rng = np.random.default_rng(5)
    q = W.uniform_q(rng, spec, 64)
    link = spec.num_links - 1
    for frame in ("body", "spatial"):
        gpu = ctx.link_jacobian(spec, q, link, frame)
        ora = O.link_jacobian(spec, q, link, frame, dtype="f64")
        np.testing.assert_allclose(gpu, ora, atol=2e-5, rtol=1e-5)


@pytest.mark.parametrize("robot", ["ur10", "humanoid"])
def test_floating_base_vs_oracle(ctx, robot):
    """FK and the floating-base link Jacobian (base columns first) with a random base pose per
    problem, against the FP64 oracle (robot_kernels.py:327-377, 667-737); the base columns satisfy
    their closed forms: spatial = Ad(T_world_base), body = Ad(T_link^-1 T_world_base)"""
    spec = {"ur10": W.ur10(), "humanoid": W.humanoid29(0)}[robot]
    rng = np.random.default_rng(9)
    B = 64
    q = W.uniform_q(rng, spec, B)
    quat = rng.normal(size=(B, 4))
    quat /= np.linalg.norm(quat, axis=1, keepdims=True)
    base = np.concatenate([rng.uniform(-2, 2, (B, 3)), quat], 1).astype(np.float32)
    gpu = ctx.fk(spec, q, T_world_base=base)
    ora, _ = O.fk(spec, q, base=base, dtype="f64")
    ok, err = _pose_close(gpu, ora, 4e-5)
    assert ok, err
    link = spec.num_links - 1
    for frame in ("body", "spatial"):
        J = ctx.link_jacobian(spec, q, link, frame, T_world_base=base)
        Jo = O.link_jacobian(spec, q, link, frame, base=base, dtype="f64")
        np.testing.assert_allclose(J, Jo, atol=5e-5, rtol=1e-5)
    # closed form of the spatial base block: Ad(T_world_base)
    adj = O.se3("adjoint", base.astype(np.float64), dtype="f64").reshape(B, 6, 6)
    Js = ctx.link_jacobian(spec, q, link, "spatial", T_world_base=base)
    np.testing.assert_allclose(Js[:, :, :6], adj, atol=2e-6)
    # without a base the base-free entry points give the joint columns of the identity base
    ident = np.tile(np.array([0, 0, 0, 1, 0, 0, 0], np.float32), (B, 1))
    np.testing.assert_allclose(ctx.link_jacobian(spec, q, link, "body", T_world_base=ident)[:, :, 6:],
                               ctx.link_jacobian(spec, q, link, "body"), atol=1e-6)

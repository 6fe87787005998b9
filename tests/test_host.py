"""CPU tests of the host side: the C-ABI library's exports, the no-fallback rule, the robot
model parser and the synthetic workloads.  No compute call reaches a GPU here."""
import ast
import ctypes as C
import re
import subprocess
import sys
from pathlib import Path

import numpy as np
import pytest

from robokit_amd import native as N
from robokit_amd import workloads as W
from robokit_amd.robo.robot_spec import RobotSpec
from tests.conftest import GOLDEN, ROOT

HEADER = ROOT / "include" / "rkhip.h"


def _header_functions():
    text = re.sub(r"/\*.*?\*/", "", HEADER.read_text(), flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:const\s+)?\w+\**\s+\**(rk_\w+)\s*\(", text, flags=re.M)))


def test_library_exports_every_declared_symbol():
    names = _header_functions()
    assert len(names) >= 20
    L = C.CDLL(str(N.LIB_PATH))
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing
    assert sorted(N.EXPORTS) == names, "ctypes binding and header disagree"
    assert L.rk_abi_version() >= 1


def test_library_targets_gfx950_only():
    data = N.LIB_PATH.read_bytes()
    assert b"gfx950" in data
    for other in (b"gfx90a", b"gfx942", b"gfx1100"):
        assert b"amdgcn-amd-amdhsa--" + other not in data


def test_no_cpu_fallback_without_a_device():
    """the product raises when no GPU is present (the CPU oracle is never used in its place)"""
    if N.lib() is None:
        pytest.fail("library missing")
    import subprocess
    import sys

    probe = subprocess.run([sys.executable, "-c", "import torch; print(torch.cuda.is_available())"],
                           capture_output=True, text=True, timeout=300)
    if probe.stdout.strip() == "True":
        pytest.skip("a GPU is present")
    with pytest.raises(N.RKError):
        N.Context(0)


def test_package_imports_neither_oracle_nor_torch():
    """robokit_amd runs without torch (tests/test_no_torch_runtime.py:1-24 of the reference) and
    never reaches the test oracle: no module imports the oracle or Warp anywhere, torch only
    inside the torch adapter's function body (imported when a torch caller calls it), and importing
    every module loads no torch"""
    for path in (ROOT / "robokit_amd").rglob("*.py"):
        tree = ast.parse(path.read_text())
        lazy = {id(n) for f in ast.walk(tree) if isinstance(f, (ast.FunctionDef, ast.AsyncFunctionDef))
                for n in ast.walk(f)}
        for node in ast.walk(tree):
            mods = []
            if isinstance(node, ast.Import):
                mods = [a.name for a in node.names]
            elif isinstance(node, ast.ImportFrom):
                mods = [node.module or ""]
            for m in mods:
                top = m.split(".")[0]
                assert top not in ("oracle", "warp"), f"{path}: imports {m}"
                assert top != "torch" or id(node) in lazy, f"{path}: imports {m} at module level"
    code = ("import sys; sys.path.insert(0, %r); import robokit_amd.lm, robokit_amd.ik, robokit_amd.population, "
            "robokit_amd.geom, robokit_amd.serving, robokit_amd.native, robokit_amd.workloads; "
            "print('torch' in sys.modules)") % str(ROOT)
    out = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=120)
    assert out.stdout.strip() == "False", out.stdout + out.stderr
    for path in (ROOT / "robokit_amd" / "csrc").iterdir():
        assert "oracle" not in path.read_text().lower() or path.suffix not in (".hip", ".h", ".cpp"), path


def test_error_codes_match_header():
    text = HEADER.read_text()
    for name, val in (("RK_OK", N.RK_OK), ("RK_EINVAL", N.RK_EINVAL), ("RK_EHIP", N.RK_EHIP),
                      ("RK_ECAPACITY", N.RK_ECAPACITY), ("RK_ENOMEM", N.RK_ENOMEM), ("RK_EBUSY", N.RK_EBUSY)):
        assert re.search(rf"#define {name} {val}\b", text), name
    assert re.search(r"#define RK_FROM_SEED 0x100", text) and N.FROM_SEED == 0x100
    assert re.search(r"#define RK_MODE_PLAIN 1", text) and N.MODE_PLAIN == 1


# ----------------------------------------------------------------------------- robot model
def test_parse_panda_urdf(panda_urdf_spec):
    s = panda_urdf_spec
    assert s.num_links == 12 and s.num_joints == 11 and s.num_actuated_joints == 8
    # the second finger joint mimics the first (multiplier 1)
    mim = np.asarray(s.mimic_actuated_joint_indices)
    assert (mim >= 0).sum() == 1
    M = s.joints_to_actuated_mapping
    assert M.shape == (s.num_joints, s.num_actuated_joints)
    assert np.all(M.sum(0) >= 1) and np.all(M.sum(1) <= 1)
    anc = s.link_ancestor_joints_mask
    assert anc.shape == (s.num_links, s.num_joints)
    assert anc[s.link_names.index("panda_link8")].sum() == 8  # 7 revolute joints + the fixed flange joint
    # parents precede children in the topological order
    order = [int(j) for j in s.topological_order_joint_indices]
    pos = {j: i for i, j in enumerate(order)}
    for j, p in enumerate(s.parent_joint_indices):
        if p >= 0:
            assert pos[p] < pos[j]
    lim = s.actuated_joint_limits_f32
    assert lim.dtype == np.float32 and np.all(lim[:, 0] <= lim[:, 1])


def test_float32_limit_narrowing():
    """robo/robot_spec_tensors.py:17-31: limits narrowed inward to float32 (lo up, hi down)"""
    from robokit_amd.robo.robot_spec import narrow_limits_to_float32

    lim = np.array([[-2.8973000000001, 2.8973000000001], [-0.1, 0.1]])
    n = narrow_limits_to_float32(lim)
    assert n.dtype == np.float32
    assert np.all(n[:, 0].astype(np.float64) >= lim[:, 0]) and np.all(n[:, 1].astype(np.float64) <= lim[:, 1])
    assert np.all(np.abs(n - lim) < 1e-6)


def test_robot_models():
    for spec, na in ((W.panda_dh(), 7), (W.ur10(), 6), (W.humanoid29(0), 29)):
        assert spec.num_actuated_joints == na
        assert np.all(np.asarray(spec.parent_joint_indices) < np.arange(spec.num_joints))
    h = W.humanoid29(0)
    assert len(W.humanoid_ee_links(h)) == 4 and len(W.humanoid_position_links(h)) == 12


# ----------------------------------------------------------------------------- workloads
def test_config1_shape_and_determinism():
    a = W.config1(seed=3, total=1000)
    b = W.config1(seed=3, total=1000)
    assert a.size == 1000 and [bk.size for bk in a.buckets] == [400, 400, 200]
    for x, y in zip(a.buckets, b.buckets):
        np.testing.assert_array_equal(x.q_seed, y.q_seed)
        np.testing.assert_array_equal(x.targets, y.targets)
        lim = x.spec.actuated_joint_limits_f32
        assert np.all(x.q_seed >= lim[:, 0]) and np.all(x.q_seed <= lim[:, 1])
        assert x.targets.shape == (x.size, len(x.links), 7)
        np.testing.assert_allclose(np.linalg.norm(x.targets[..., 3:], axis=-1), 1, atol=1e-6)
    c = W.config1(seed=4, total=1000)
    assert not np.array_equal(a.buckets[0].q_seed, c.buckets[0].q_seed)


def test_config0_targets_reachable():
    from robokit_amd.robo.kinematics_np import forward_kinematics

    b = W.config0(seed=0, batch=16)
    assert b.q_seed.shape == (16, 7) and b.targets.shape == (16, 1, 7)
    T = forward_kinematics(b.spec, b.q_target)
    np.testing.assert_allclose(T[:, b.links[0], :3], b.targets[:, 0, :3], atol=1e-5)


def test_config2_clip():
    clip = W.config2(seed=0, frames=64)
    assert clip.targets.shape == (64, 12, 7) and clip.q_traj.shape == (64, 29)
    # five sinusoids of 0.5 rad at <= 2 Hz: |dq/dt| <= 5 * 0.5 * 2 pi * 2 rad/s, at 30 fps per frame
    assert np.abs(np.diff(clip.q_traj, axis=0)).max() <= 5 * 0.5 * 2 * np.pi * 2 / 30


# ctypes structures of the binding and the C structs of include/rkhip.h they mirror
_STRUCTS = [(N.RobotDesc, "rk_robot_desc"), (N.TermsDesc, "rk_terms_desc"), (N.LMParamsC, "rk_lm_params"),
            (N.TrajDescC, "rk_traj_desc"), (N.RequestC, "rk_request"), (N.ServeStatsC, "rk_serve_stats")]


def test_struct_layouts_match_header(tmp_path):
    """every field of every ctypes structure sits at the offset the C compiler gives it, and the
    sizes agree (gcc on a probe that includes the header)"""
    import shutil
    import subprocess

    if shutil.which("gcc") is None:
        pytest.skip("gcc not available")
    lines = ['#include <stdio.h>', '#include <stddef.h>', f'#include "{HEADER}"', "int main(void) {"]
    for cls, cname in _STRUCTS:
        lines.append(f'  printf("{cname} sizeof %zu\\n", sizeof({cname}));')
        for fname, _ in cls._fields_:
            lines.append(f'  printf("{cname} {fname} %zu\\n", offsetof({cname}, {fname}));')
    lines += ["  return 0;", "}"]
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines) + "\n")
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c11", "-o", str(exe), str(src)], check=True, capture_output=True)
    got = {}
    for line in subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.splitlines():
        cname, field, value = line.split()
        got[(cname, field)] = int(value)
    for cls, cname in _STRUCTS:
        assert C.sizeof(cls) == got[(cname, "sizeof")], cname
        for fname, _ in cls._fields_:
            assert getattr(cls, fname).offset == got[(cname, fname)], (cname, fname)

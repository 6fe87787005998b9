import os
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
sys.path.insert(0, str(ROOT))
GOLDEN = ROOT / "tests" / "golden"


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


@pytest.fixture(scope="session")
def ctx():
    from robokit_amd import native

    c = native.Context(0)
    yield c
    c.close()


@pytest.fixture(scope="session")
def panda_urdf_spec():
    from robokit_amd.robo.robot_spec import RobotSpec

    return RobotSpec.parse_urdf((GOLDEN / "panda_kinematics.urdf").read_text())

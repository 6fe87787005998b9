"""Write the golden fixtures under tests/golden/ from data files the reference's tests hold.

Run in the build container (where the reference checkout is mounted):
    python tests/golden/make_fixtures.py [/root/reference]

Outputs (data only -- inputs and expected outputs):
  panda_kinematics.urdf   the kinematic elements (joints, links) of tests/fixtures/panda.urdf
  fk_ground_truth.json    panda + ur10 entries of tests/robo/fk_ground_truth.json
  jacobian_ground_truth.json  panda + ur10 entries of tests/robo/jacobian_ground_truth.json
  se3_constants.json      the T_X_90 / pure-translation expected values of tests/lie/test_se3.py:39-80
"""
from __future__ import annotations

import json
import sys
import xml.etree.ElementTree as ET
from pathlib import Path

HERE = Path(__file__).resolve().parent
KEEP = ("panda_description", "ur10_description")


def main(ref: Path) -> None:
    tests = ref / "tests"
    # URDF: keep only kinematics (no visual/collision/inertial geometry)
    root = ET.parse(tests / "fixtures" / "panda.urdf").getroot()
    out = ET.Element("robot", {"name": root.get("name", "panda")})
    for el in root:
        if el.tag == "link":
            ET.SubElement(out, "link", {"name": el.get("name")})
        elif el.tag == "joint":
            j = ET.SubElement(out, "joint", {"name": el.get("name"), "type": el.get("type")})
            for sub in el:
                if sub.tag in ("origin", "parent", "child", "axis", "limit", "mimic"):
                    ET.SubElement(j, sub.tag, dict(sub.attrib))
    ET.indent(out)
    (HERE / "panda_kinematics.urdf").write_text(ET.tostring(out, encoding="unicode") + "\n")

    for name in ("fk_ground_truth.json", "jacobian_ground_truth.json"):
        data = json.loads((tests / "robo" / name).read_text())
        data["robots"] = [r for r in data["robots"] if r["robot_description_name"] in KEEP]
        (HERE / name).write_text(json.dumps(data, separators=(",", ":")))

    # expected values stated in tests/lie/test_se3.py:39-80
    se3 = {
        "T_X_90": [1.0, 2.0, 3.0, 0.7071068, 0.7071068, 0.0, 0.0],
        "T_X_90_matrix": [[1, 0, 0, 1], [0, 0, -1, 2], [0, 1, 0, 3], [0, 0, 0, 1]],
        "T_X_90_inverse": [-1.0, -3.0, 2.0, 0.7071068, -0.7071068, 0.0, 0.0],
        "T_X_90_log": [1.0, 3.9269908, 0.7853982, 1.5707964, 0.0, 0.0],
        "T_X_90_adjoint": [
            [1, 0, 0, 0, 2, 3], [0, 0, -1, 3, -1, 0], [0, 1, 0, -2, 0, -1],
            [0, 0, 0, 1, 0, 0], [0, 0, 0, 0, 0, -1], [0, 0, 0, 0, 1, 0],
        ],
        "T_TRANS": [1.0, 2.0, 3.0, 1.0, 0.0, 0.0, 0.0],
        "T_TRANS_log": [1.0, 2.0, 3.0, 0.0, 0.0, 0.0],
    }
    (HERE / "se3_constants.json").write_text(json.dumps(se3, indent=1))


if __name__ == "__main__":
    main(Path(sys.argv[1]) if len(sys.argv) > 1 else Path("/root/reference"))

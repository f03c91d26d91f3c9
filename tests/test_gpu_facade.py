"""The reference's own gtest cases, run through the C++ facade (include/cpvs_amd.hpp) by the
binary tests/cpp/test_facade built in __graft_entry__.build()."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "tests", "cpp", "test_facade")


def _cases(fx, path):
    with open(path, "w") as f:
        for name, arr in fx["arrays"].items():
            f.write(f"image {name} {arr.shape[0]} " + " ".join(repr(float(v)) for v in arr.reshape(-1)) + "\n")
        for e in fx["minmax"]:
            if e[1] == "numLevels":
                f.write(f"levels {e[0]} {e[2]}\n")
            else:
                f.write(f"minmax {e[0]} {e[1]} {e[2]} {e[3]} {e[4]} {e[5]!r} {e[6]!r}\n")
        for name, ndc, try_leaf, exp in fx["traverse"]:
            f.write(f"traverse {name} {ndc[0]!r} {ndc[1]!r} {ndc[2]!r} {int(try_leaf)} {exp}\n")
        for sw in fx["sweeps"]:
            f.write(f"sweep {sw['image']} {sw['x'][0]} {sw['x'][1]} {sw['y'][0]} {sw['y'][1]} {sw['z']!r} {sw['expected']}\n")


def test_reference_gtests_through_cpp_facade(ref_fixtures, tmp_path):
    assert os.path.exists(BIN), "build the facade test with __graft_entry__.build()"
    cases = tmp_path / "cases.txt"
    _cases(ref_fixtures, str(cases))
    r = subprocess.run([BIN, str(cases)], capture_output=True, text=True, timeout=120)
    print(r.stdout[-2000:], r.stderr[-2000:])
    assert r.returncode == 0, r.stdout
    assert "0 failures" in r.stdout

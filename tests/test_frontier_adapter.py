"""The FronTier adapter (include/collision_amd/FronTierAdapter.hpp) on a
small FronTier-like object model: assembleFromInterface's walk order
(dcollid3d.cpp:12-52), point numbering, skipped boundary surfaces and
non-string curves, fixed and movable points, and the state round trip.
Host code only (g++)."""
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_frontier_adapter(tmp_path):
    exe = tmp_path / "frontier_adapter_test"
    src = os.path.join(ROOT, "tests", "cpp", "frontier_adapter_test.cpp")
    subprocess.run(["g++", "-std=c++17", "-O1", "-Wall", "-Werror", "-I", os.path.join(ROOT, "include"),
                    src, "-o", str(exe)], check=True)
    r = subprocess.run([str(exe)], capture_output=True, text=True)
    assert r.returncode == 0 and r.stdout.strip() == "ok", r.stdout + r.stderr

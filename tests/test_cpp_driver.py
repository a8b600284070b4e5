"""A C++ program that uses only include/*.h and libvoxelrestir.so (the
reference's own language and main-loop shape) renders the same frame bits
as the Python binding of the same calls."""
import os
import shutil
import subprocess
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parent.parent
SRC = ROOT / "tests" / "cpp" / "render_driver.cpp"


def _build(tmp_path: Path) -> Path:
    from smallproject_amd._lib import LIBPATH
    cxx = shutil.which("g++")
    if cxx is None:
        pytest.skip("no g++")
    exe = tmp_path / "render_driver"
    subprocess.run([cxx, "-O2", "-std=c++17", "-I", str(ROOT / "include"), str(SRC), str(LIBPATH),
                    f"-Wl,-rpath,{LIBPATH.parent}", "-o", str(exe)], check=True)
    return exe


def test_driver_builds_against_the_headers_and_library(tmp_path):
    exe = _build(tmp_path)
    assert exe.exists()
    out = subprocess.run([str(exe)], capture_output=True, text=True)
    assert out.returncode == 2 and "usage" in out.stderr


@pytest.mark.gpu
def test_driver_frame_matches_the_python_binding(gpu, tmp_path):
    exe = _build(tmp_path)
    W, H, F = 160, 96, 4
    out = subprocess.run([str(exe), str(W), str(H), str(F)], capture_output=True, text=True, timeout=120,
                         env=dict(os.environ))
    assert out.returncode == 0, out.stderr
    fields = out.stdout.split()
    bits = fields[fields.index("bits") + 1]

    from smallproject_amd import camera_path, default_config
    from smallproject_amd.api import World
    w = World(default_config(W, H, point=True, capacity=1 << 14))
    w.scene_terrain(128, 1, 1)
    w.optimize()
    w.scene_lights_above(64, 128, 1, 1, 2, 2, 20, 1, 10)
    w.set_focal(1.7320508)
    w.commit()
    p = w.params
    p.numberOfCandidates, p.spatialTaps, p.spatialRadius, p.numberOfMaxTemporalImportance = 32, 5, 30, 20
    for f in range(F):
        pos, tgt = camera_path(0, 3, f, 8, 128, 1, 1)
        w.look_at(pos, tgt)
        w.render()
        w.commit()
    frame = w.renderer.frame()
    h = 1469598103934665603
    for byte in frame.astype(np.float32).tobytes():
        h = ((h ^ byte) * 1099511628211) & 0xFFFFFFFFFFFFFFFF
    w.close()
    assert bits == f"{h:016x}"
    assert int(fields[fields.index("lights") + 1]) == 64

"""bench.py's command-line contract that holds without a GPU: one GPU only
(--gpus above 1 exits with a message and status 2), and the algorithmic
byte counts and metric name it reports."""
import subprocess
import sys
from pathlib import Path

ROOT = Path(__file__).resolve().parent.parent


def test_more_than_one_gpu_exits_with_a_message():
    res = subprocess.run([sys.executable, str(ROOT / "bench.py"), "--gpus", "2"], capture_output=True, text=True,
                         timeout=120)
    assert res.returncode == 2
    assert "one GPU" in res.stderr
    assert res.stdout == ""


def test_stage_bytes_cover_every_stage():
    sys.path.insert(0, str(ROOT))
    import bench
    from smallproject_amd import _lib
    assert set(bench.STAGE_BYTES) == set(_lib.STAGE_NAMES)
    # SoA planes: hit 12 B; reservoir 36 B; seed 4 B; frame 16 B (per pixel)
    # records: 16-B test word + 16-B sample + 8-B sums; the temporal pass
    # reads the word and the sample of the previous frame's record
    assert bench.STAGE_BYTES == {"primary": 12, "initial": 12 + 16 + 16 + 40 + 4, "resolve": 40 + 12 + 40 + 4 + 16}
    assert bench.METRIC.startswith("1080p fps")

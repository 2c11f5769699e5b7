"""bench.py's contract pieces that need no GPU: the refusal of more than one
GPU, and the committed HBM-traffic summary that fills `roofline.traffic` for
every launch shape and whole-bin threshold the autotuner can pick on the
headline workload."""
import subprocess
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
sys.path.insert(0, str(ROOT))

import bench  # noqa: E402


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, str(ROOT / "bench.py"), "--gpus", "2"], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr


@pytest.mark.parametrize("whole_bin", [8, 16, 32, 64, 128])
def test_traffic_summary_covers_the_tuned_thresholds(whole_bin):
    t = bench.traffic_from_profiles("open2m", 1, whole_bin)
    assert t is not None
    # against 4 B per window-restricted ion id of the 100k-spectrum open batch
    ratio = t / (4 * 63_593_296_046)
    assert 1.0 < ratio < 2.0


def test_traffic_summary_is_tied_to_its_launch_configuration():
    assert bench.traffic_from_profiles("open2m", 0, 32) is None
    assert bench.traffic_from_profiles("no-such-workload", 1, 32) is None

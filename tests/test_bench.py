"""bench.py without a GPU: the algorithmic work per drop of the 4x4 1024-subcarrier link matches
the per-stage counts of the reference algorithm, and a multi-GPU request exits with a message."""

import subprocess
import sys
from pathlib import Path

ROOT = Path(__file__).resolve().parents[1]


def test_flops_per_drop_of_the_metric_link():
    sys.path.insert(0, str(ROOT))
    import bench
    from hermespy_amd.configs import mimo4x4_1024

    cfg = mimo4x4_1024()
    # TDL 23 paths x 15 344 samples x 4 streams, spatial mix 4 x 4 x (15 344 + 92), CSI 23 x 14 x 1024,
    # MMSE 133 complex MACs per resource element; FFTs 5 N log2 N per stream and word (8 streams)
    cmac = 23 * 15344 * 4 + 4 * 4 * (15344 + 92) + 23 * 14 * 1024 + 133 * 14 * 1024
    assert cmac == 1411648 + 246976 + 329728 + 1906688
    fft = 5 * 1024 * 10 * 14 * 8
    assert fft == 5734400
    assert bench.algorithmic_flops_per_drop(cfg) == 8 * cmac + fft == 36894720


def test_more_than_one_gpu_is_refused():
    res = subprocess.run([sys.executable, str(ROOT / "bench.py"), "--gpus", "2"], capture_output=True, text=True,
                         timeout=120)
    assert res.returncode == 2 and "one GPU" in res.stderr and res.stdout == ""

"""bench.py's command-line contract that needs no GPU."""

import json
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_more_than_one_gpu_is_refused_with_a_message():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr
    assert r.stdout.strip() == ""


def test_pcg_byte_model_matches_the_kernels():
    sys.path.insert(0, ROOT)
    import bench
    # 6 words per row for the SpMV, the two axpys, each triangular sweep and
    # 3 for the direction update: 27 words = 216 B (FP64) per row per iteration.
    assert sum(bench.PCG_WORDS.values()) == 27
    args = bench.parse_args([])
    assert args.gpus == 1 and args.size == [256, 128, 128]
    assert json.dumps({"x": None})  # the line is plain JSON

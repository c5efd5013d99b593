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
    # FP64 path: SpMV 36 B per row (s, z, FP64 diag + 3 upper couplings
    # stored as FP32 face weights), axpy 48, forward sweep 48, backward sweep
    # 56 (it also reads r for the fused z.r), direction update 24: 212 B per
    # row per iteration.  Mixed: FP32 vectors and factor, FP64 x, the same
    # matrix, the direction update in the SpMV, the axpy in the forward
    # sweep (which reads r there anyway): 36 + (24 + 28) + 28 = 116 B.
    assert sum(bench.PCG_BYTES["fp64"].values()) == 212
    assert sum(bench.PCG_BYTES["mixed"].values()) == 116
    assert bench.PCG_BYTES["mixed"]["k_axpy_norm"] == 0
    # A true residual (x, b, diag, 3 weights, r) and the forward sweep over it.
    assert bench.TRUE_RESIDUAL_BYTES == (8 + 8 + 8 + 3 * 4 + 4) + 6 * 4
    args = bench.parse_args([])
    assert args.gpus == 1 and args.size == [256, 128, 128]
    assert json.dumps({"x": None})  # the line is plain JSON


def test_cpu_baseline_leaves_stdout_to_the_json_line():
    # The reference engine prints a banner and a log on stdout; bench.py's
    # stdout must carry only its one JSON line.
    code = ("import sys, json; sys.path.insert(0, %r); import bench; from flipengine3d_amd import scenes; "
            "a = bench.parse_args(['--size', '24', '16', '16', '--cpu-seconds', '0.01', '--cpu-threads', '1']); "
            "s = scenes.double_dam_break(24, 16, 16, 1 / 16, seed=3); r = bench.cpu_baseline(s, a); "
            "print(json.dumps(r))") % ROOT
    r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    lines = r.stdout.strip().splitlines()
    assert len(lines) == 1, r.stdout
    assert json.loads(lines[0])["value"] > 0

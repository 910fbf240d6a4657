"""bench.py's contract on the host: one GPU only (a message and exit status 2
for --gpus > 1), and the per-pass byte accounting its roofline uses."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

import bench  # noqa: E402


def test_more_than_one_gpu_exits_with_a_message():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr


def test_pass_bytes():
    """12-pass schedule with the Horner-form RK4 epilogue: FP64 432 B and mixed
    240 B per point and step (mixed stages 1-3 read the FP32 mirror of Y); the
    folded 8-launch schedule drops the four first passes and adds a P write to
    each RK pass: 368 and 208 B; the fused 3-register minimum is 256 / 208 B,
    the Horner-form minimum 176 / 112 B."""
    assert bench.step_bytes("fp64", 12) == 432 and bench.step_bytes("mixed", 12) == 240
    assert bench.step_bytes("fp64", 8) == 368 and bench.step_bytes("mixed", 8) == 208
    assert bench.FUSED_MIN_BYTES == {"fp64": 256, "mixed": 208}
    assert bench.HORNER_MIN_BYTES == {"fp64": 176, "mixed": 112}
    for prec in ("fp64", "mixed"):
        t = bench.PASS_BYTES[prec]
        assert all(len(t[k]) == 4 for k in ("first", "mid", "final"))

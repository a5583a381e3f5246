"""bench.py's command line on the CPU: more than one GPU is refused with a
message (the path tracer is a single-GPU program), and the copy calibration,
build-only and batch-frames-only modes are listed without running the benchmark."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_more_than_one_gpu_is_refused():
    p = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2", "--steps", "1"],
                       capture_output=True, text=True, timeout=300)
    assert p.returncode == 2
    assert "one GPU" in p.stderr


def test_help_lists_the_options():
    p = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--help"], capture_output=True, text=True,
                       timeout=300)
    assert p.returncode == 0
    for opt in ("--gpus", "--steps", "--warmup", "--no-union", "--no-bounce1", "--build-only", "--no-tune", "--traversal", "--batch-traversal",
                "--batch-frames-only", "--copy-calibration"):
        assert opt in p.stdout

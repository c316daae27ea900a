"""bench.py's command line: it measures one GPU, and any other --gpus value
stops with a message before a device is touched."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_more_than_one_gpu_is_refused():
    p = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"),
                        "--gpus", "2"], capture_output=True, text=True,
                       timeout=120)
    assert p.returncode == 2
    assert "one GPU only" in p.stderr and "--gpus 2" in p.stderr
    assert p.stdout == ""

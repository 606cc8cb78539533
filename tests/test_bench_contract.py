"""bench.py's command-line contract on the CPU: one GPU only (a request for
more exits with status 2 and says so, before touching any device)."""
import subprocess
import sys

from .conftest import ROOT


def test_more_than_one_gpu_is_refused():
    out = subprocess.run([sys.executable, str(ROOT / "bench.py"), "--gpus", "2", "--steps", "1", "--warmup", "0"],
                         capture_output=True, text=True, timeout=120)
    assert out.returncode == 2
    assert "one GPU" in out.stderr
    assert out.stdout == ""

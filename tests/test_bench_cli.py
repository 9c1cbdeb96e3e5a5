"""bench.py's command-line contract that needs no GPU: one GPU per run."""
import subprocess
import sys
from pathlib import Path

REPO = Path(__file__).resolve().parent.parent


def test_bench_refuses_more_than_one_gpu():
    """--gpus above 1 exits non-zero with a message saying so, before touching a device."""
    p = subprocess.run([sys.executable, str(REPO / "bench.py"), "--gpus", "2"], capture_output=True, text=True, timeout=120)
    assert p.returncode != 0
    assert "one GPU" in (p.stdout + p.stderr)

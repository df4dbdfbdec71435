"""bench.py refuses more than one GPU and its CLI parses (no device work)."""
import subprocess
import sys

from conftest import ROOT


def test_bench_rejects_multi_gpu():
    r = subprocess.run([sys.executable, str(ROOT / "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode != 0
    assert "one GPU" in (r.stdout + r.stderr)

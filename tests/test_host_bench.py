"""bench.py's command line on the CPU: it parses, and refuses more than one GPU.

Neither case reaches a device call: --help exits in argparse and --gpus 2 returns
before any workload starts.
"""
import subprocess
import sys
from pathlib import Path

ROOT = Path(__file__).resolve().parents[1]


def _bench(*args):
    return subprocess.run([sys.executable, str(ROOT / "bench.py"), *args], cwd=ROOT, capture_output=True,
                          text=True, timeout=120)


def test_bench_parses_and_lists_its_workloads():
    out = _bench("--help")
    assert out.returncode == 0, out.stderr
    for flag in ("--workload", "--history-steps", "--forms", "--cpu-seconds"):
        assert flag in out.stdout


def test_bench_refuses_more_than_one_gpu():
    out = _bench("--gpus", "2")
    assert out.returncode == 2
    assert "one GPU" in out.stderr
    assert out.stdout == ""

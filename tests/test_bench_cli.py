"""bench.py's command line on the CPU: more than one GPU is refused with a message before
anything touches a device, and every workload is a valid choice."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _bench(*args):
    return subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), *args],
                          capture_output=True, text=True, timeout=120)


def test_more_than_one_gpu_is_refused():
    r = _bench("--gpus", "2")
    assert r.returncode == 2
    assert "one GPU" in r.stderr
    assert r.stdout == ""


def test_help_lists_every_workload():
    r = _bench("--help")
    assert r.returncode == 0
    for w in ("scene", "stacks", "stream", "regions", "dropin"):
        assert w in r.stdout

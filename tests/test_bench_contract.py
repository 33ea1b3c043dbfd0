"""bench.py's command-line contract that needs no GPU: a run on more than one GPU is
refused with a message (this build measures one GPU), before anything touches a device."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _bench(*args, env=None):
    return subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), *args],
                          capture_output=True, text=True, timeout=120, env=env)


def test_more_than_one_gpu_is_refused():
    r = _bench("--gpus", "2")
    assert r.returncode == 2
    assert "one GPU" in r.stderr and r.stdout == ""


def test_launcher_world_size_is_refused():
    env = dict(os.environ, WORLD_SIZE="8")
    r = _bench("--gpus", "1", env=env)
    assert r.returncode == 2 and "one GPU" in r.stderr

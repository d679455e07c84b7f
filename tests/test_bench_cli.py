"""bench.py's command-line contract on the CPU: the paths that end before any device work."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _bench(*args):
    env = dict(os.environ)
    env.pop("WORLD_SIZE", None)
    return subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), *args], capture_output=True, text=True,
                          timeout=120, env=env)


def test_more_than_one_gpu_is_refused_with_a_message():
    res = _bench("--gpus", "2")
    assert res.returncode != 0
    assert "one GPU" in res.stderr or "single MI355X" in res.stderr


def test_train_workload_is_a_probe_only():
    res = _bench("--workload", "train")
    assert res.returncode != 0 and "--probe" in res.stderr


def test_help_lists_the_workloads():
    res = _bench("--help")
    assert res.returncode == 0
    for w in ("poisson", "c3", "c4", "train"):
        assert w in res.stdout

"""bench.py's command-line contract on the CPU: more than one GPU (by flag or by a
torch.distributed launch's WORLD_SIZE) is refused with a message and a non-zero status
before anything touches a GPU."""
import os
import subprocess
import sys

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def run_bench(args, env_extra=None):
    env = dict(os.environ)
    env.update(env_extra or {})
    return subprocess.run([sys.executable, os.path.join(REPO, "bench.py"), *args],
                          capture_output=True, text=True, env=env, timeout=300)


def test_more_than_one_gpu_is_refused():
    r = run_bench(["--gpus", "2"])
    assert r.returncode != 0
    assert "one GPU" in r.stderr
    assert r.stdout.strip() == ""


def test_distributed_launch_is_refused():
    r = run_bench(["--gpus", "1"], {"WORLD_SIZE": "4"})
    assert r.returncode != 0
    assert "one GPU" in r.stderr

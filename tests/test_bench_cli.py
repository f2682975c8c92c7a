"""CPU: the bench's command-line contract that needs no GPU.  With --gpus above 1 (or a
WORLD_SIZE above 1 from a multi-process launcher) bench.py exits with status 2 and says on
stderr that it runs on one GPU, before it imports torch or touches a device."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _run(args, env=None):
    return subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), *args], capture_output=True, text=True,
                          timeout=120, env=env)


def test_more_than_one_gpu_is_refused():
    res = _run(["--gpus", "2", "--steps", "1", "--warmup", "0"])
    assert res.returncode == 2
    assert "one GPU" in res.stderr
    assert res.stdout == ""


def test_multi_process_launch_is_refused():
    env = dict(os.environ, WORLD_SIZE="4", RANK="0", LOCAL_RANK="0")
    res = _run(["--steps", "1"], env=env)
    assert res.returncode == 2 and "one GPU" in res.stderr

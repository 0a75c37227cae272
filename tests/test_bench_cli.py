"""bench.py's command line: the one-GPU contract and argument checks run
before anything touches a GPU."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _run(*args):
    return subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), *args], capture_output=True, text=True,
                          timeout=120)


def test_more_than_one_gpu_is_refused():
    p = _run("--gpus", "2")
    assert p.returncode != 0 and "one GPU" in p.stderr


def test_launch_opts_must_be_five_or_six_integers():
    for bad in ("1,2,3", "4,32,8,1,x", "1,2,3,4,5,6,7"):
        p = _run("--launch-opts", bad)
        assert p.returncode != 0 and "five or six integers" in p.stderr, bad


def test_steps_must_be_positive():
    p = _run("--steps", "0")
    assert p.returncode != 0 and "--steps" in p.stderr


def test_streams_must_be_one_to_three():
    for bad in ("4", "-1"):
        p = _run("--streams", bad)
        assert p.returncode != 0 and "--streams" in p.stderr, bad

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


def test_lone_ray_needs_trace_only():
    p = _run("--lone-ray")
    assert p.returncode != 0 and "--trace-only" in p.stderr


def test_trace_only_spec_is_checked():
    for bad, needle in (("nosuch", "preset"), ("hd1080_a0998:x", "numbers"), ("hd1080_a0998:0", "positive"),
                        ("hd1080_a0998:1e-7:fp16", "fp64 or mixed"), ("uhd2160_i60:1e-7:fp64:1.5", "spin"),
                        ("hd1080_a0998:1:fp64:0.5:extra", "preset")):
        p = _run("--trace-only", bad)
        assert p.returncode != 0 and needle in p.stderr, (bad, p.stderr)


def test_profiling_modes_keep_the_one_gpu_contract():
    for extra in (("--trace-only", "hd1080_a0998"), ("--copy-mib", "64")):
        p = _run("--gpus", "2", *extra)
        assert p.returncode != 0 and "one GPU" in p.stderr, extra

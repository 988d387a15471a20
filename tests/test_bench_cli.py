"""The benchmark's command line on the CPU: a request for more than one GPU
ends with a message and exit status 2 before anything touches a device, and
the options parse.  (The GPU legs themselves run in the benchmark.)"""
import json
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def run_bench(*args):
    return subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), *args], capture_output=True, text=True,
                          timeout=120, cwd=ROOT)


def test_more_than_one_gpu_is_refused():
    r = run_bench("--gpus", "2")
    assert r.returncode == 2
    line = json.loads(r.stdout.strip().splitlines()[-1])
    assert line["n_gpus"] == 2 and "one MI355X" in line["error"]


def test_options_parse():
    sys.path.insert(0, ROOT)
    import bench
    argv = sys.argv
    try:
        sys.argv = ["bench.py", "--only-c4", "--c4-eager", "eager_single", "--steps", "2", "--no-cpu"]
        a = bench.parse()
    finally:
        sys.argv = argv
    assert a.only_c4 and a.c4_eager == "eager_single" and a.steps == 2 and a.no_cpu and a.gpus == 1

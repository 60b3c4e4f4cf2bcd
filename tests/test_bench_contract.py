"""bench.py's command-line contract on the host: more than one GPU is refused with
a message before any device work, and --help lists every workload."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def run(*args):
    return subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), *args], capture_output=True, text=True,
                          timeout=120)


def test_more_than_one_gpu_is_refused_with_a_message():
    r = run("--gpus", "2")
    assert r.returncode != 0
    assert "one GPU" in (r.stdout + r.stderr)


def test_help_lists_the_workloads():
    r = run("--help")
    assert r.returncode == 0
    for w in ("plain", "batched", "served", "large", "train", "container", "dwt", "int16", "drop_in"):
        assert w in r.stdout

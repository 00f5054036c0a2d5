"""bench.py command-line contract that needs no GPU: one GPU only, and bad step
counts are refused before anything runs."""
import json
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _run(*args):
    return subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), *args], capture_output=True, text=True,
                          timeout=120, cwd=ROOT)


def test_more_than_one_gpu_is_refused_with_a_message():
    r = _run("--gpus", "2")
    assert r.returncode != 0
    msg = json.loads(r.stdout.strip().splitlines()[-1])
    assert msg["n_gpus"] == 2 and "one GPU" in msg["error"]


def test_zero_steps_is_refused():
    r = _run("--steps", "0")
    assert r.returncode == 2 and "--steps" in r.stderr


def test_traffic_figure_is_keyed_to_the_kernel_sources():
    """profiles/traffic.json carries the hash of the engine sources it was measured
    on; bench.py reports the traffic only when that hash is the current one."""
    sys.path.insert(0, ROOT)
    import bench
    info = json.load(open(os.path.join(ROOT, "profiles", "traffic.json")))["random256"]
    assert len(info["kernel_hash"]) == 16 and info["bytes_per_energy"] > 0
    assert len(bench.kernel_source_hash()) == 16

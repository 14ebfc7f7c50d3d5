"""bench.py's command line without a GPU: it runs on one GPU only and says
so for any other count, and its options parse."""
import json
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _bench(*args):
    return subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), *args], cwd=ROOT,
                          capture_output=True, text=True, timeout=300)


def test_more_than_one_gpu_exits_with_a_message():
    r = _bench("--gpus", "2")
    assert r.returncode == 2
    line = json.loads(r.stdout.strip().splitlines()[-1])
    assert "one GPU" in line["error"]


def test_options_parse():
    r = _bench("--help")
    assert r.returncode == 0
    for opt in ("--steps", "--warmup", "--steps-per-launch", "--batch-classes", "--batch-class",
                "--batch-eager", "--cli", "--large", "--grad-reference", "--calibrate-copy"):
        assert opt in r.stdout

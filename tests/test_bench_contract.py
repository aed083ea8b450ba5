"""bench.py command-line contract that needs no GPU: a request for more than
one GPU ends with an explanatory JSON line and a non-zero status (this build
runs on one GPU), and the flag set the harness uses parses."""
import json
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_more_than_one_gpu_is_refused_with_a_message():
    p = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2",
                        "--steps", "5", "--warmup", "1"],
                       capture_output=True, text=True, timeout=120)
    assert p.returncode != 0
    line = json.loads(p.stdout.strip().splitlines()[-1])
    assert "one GPU" in line["error"]


def test_bench_flags_parse():
    sys.path.insert(0, ROOT)
    import importlib
    bench = importlib.import_module("bench")
    argv = sys.argv
    try:
        sys.argv = ["bench.py", "--gpus", "1", "--steps", "7", "--warmup", "3", "--neigh", "CSR",
                    "--no-cpu", "--no-batch", "--no-forcematch", "--no-alt-layout"]
        a = bench.parse()
    finally:
        sys.argv = argv
    assert (a.gpus, a.steps, a.warmup, a.neigh) == (1, 7, 3, "CSR")
    assert a.no_cpu and a.no_batch and a.no_forcematch and a.no_alt_layout

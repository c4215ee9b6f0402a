"""The bench driver's command-line contract, checked without a GPU: the file
compiles, and a request for more than one GPU is refused with a message
before any device work (the library drives one GPU per process)."""
import os
import py_compile
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BENCH = os.path.join(ROOT, "bench.py")


def test_bench_compiles():
    py_compile.compile(BENCH, doraise=True)


def test_bench_refuses_several_gpus():
    r = subprocess.run([sys.executable, BENCH, "--gpus", "2"], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode != 0
    assert "--gpus 2" in r.stderr

"""The reference's own Level-3 conformance tester on the GPU path.

oracle/_ref/dblat3 is tests/netlib/dblat3.f of the reference compiled from the
reference tree (oracle/Makefile, target ref) and linked against
libgpublas.so.  With cutoff=0 every DGEMM/DSYMM/DTRSM/DSYRK call it makes
(N in {0,1,2,3,5,9}, alpha in {0,1,0.7}, beta in {0,1,1.3}, every trans /
side / uplo / diag, ld padding filled with ROGUE) runs on the gfx950 kernels;
dblat3 checks each result with its DMMCH gauge (threshold 16) and the
untouched regions bit-for-bit, and checks every error exit through XERBLA.
DTRMM and DSYR2K are not on the hot path and are switched off in the input.
"""
import os
import subprocess
import tempfile

import pytest

import libgpublas_amd as gb

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_dblat3_all_hot_path_routines_on_gpu(gpu):
    exe = os.path.join(ROOT, "oracle", "_ref", "dblat3")
    if not os.path.exists(exe):
        pytest.skip("oracle/_ref/dblat3 was not built (needs the reference tree at build time)")
    inp = open(os.path.join(ROOT, "tests", "golden", "input.dblat3.gpu")).read()
    with tempfile.TemporaryDirectory() as td:
        env = dict(os.environ)
        env["GPUBLAS_OPTIONS"] = "cutoff=0;stats=" + os.path.join(td, "stats.csv")
        out = subprocess.run([exe], input=inp, cwd=td, env=env, capture_output=True, text=True,
                             timeout=600)
        summ = open(os.path.join(td, "DBLAT3.SUMM")).read()
        stats = open(os.path.join(td, "stats.csv")).read().splitlines()
    assert out.returncode == 0, out.stderr[-3000:]
    for r in ("DGEMM ", "DSYMM ", "DTRSM ", "DSYRK "):
        assert f"{r} PASSED THE TESTS OF ERROR-EXITS" in summ, summ
        assert f"{r} PASSED THE COMPUTATIONAL TESTS" in summ, summ
    assert "FAIL" not in summ and "**" not in summ
    hdr = [h.strip() for h in stats[0].split(",")]
    vals = dict(zip(hdr, (int(v) for v in stats[1].split(","))))
    # every call that is not an error exit or a quick / alpha == 0 return ran on the GPU
    assert vals["DeviceCalls"] > 5000 and vals["Fallthrough"] == 0

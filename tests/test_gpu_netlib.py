"""The reference's own Level-3 conformance testers on the GPU path.

oracle/_ref/{d,s,c,z}blat3 are tests/netlib/{d,s,c,z}blat3.f of the
reference compiled from the reference tree (oracle/Makefile, target ref) and
linked against libgpublas.so; they run with the reference's own input files
(tests/netlib/input.{d,s,c,z}blat3, committed unchanged under tests/golden/), as
the reference's tests/netlib/meson.build:1-21 runs them.  With cutoff=0 every
GEMM / SYMM / TRMM / TRSM / SYRK / SYR2K call they make (N in {0,1,2,3,5,9},
alpha in {0,1,0.7}, beta in {0,1,1.3}, every trans / side / uplo / diag, ld
padding filled with ROGUE) runs on the gfx950 kernels -- for the complex
testers also HEMM, HERK and HER2K, alpha and beta complex; the testers check
each result with their gauge (threshold 16) and the untouched regions
bit-for-bit, and check every error exit through XERBLA.
"""
import os
import subprocess
import tempfile

import pytest

import libgpublas_amd as gb

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


ROUTINES = {"d": ("GEMM ", "SYMM ", "TRMM ", "TRSM ", "SYRK ", "SYR2K"),
            "c": ("GEMM ", "HEMM ", "SYMM ", "TRMM ", "TRSM ", "HERK ", "SYRK ", "HER2K", "SYR2K")}
ROUTINES["s"], ROUTINES["z"] = ROUTINES["d"], ROUTINES["c"]


@pytest.mark.parametrize("prec", ["d", "s", "c", "z"])
def test_netlib_level3_tester_on_gpu(gpu, prec):
    exe = os.path.join(ROOT, "oracle", "_ref", f"{prec}blat3")
    if not os.path.exists(exe):
        pytest.skip(f"oracle/_ref/{prec}blat3 was not built (needs the reference tree at build time)")
    inp = open(os.path.join(ROOT, "tests", "golden", f"input.{prec}blat3")).read()
    with tempfile.TemporaryDirectory() as td:
        env = dict(os.environ)
        env["GPUBLAS_OPTIONS"] = "cutoff=0;stats=" + os.path.join(td, "stats.csv")
        out = subprocess.run([exe], input=inp, cwd=td, env=env, capture_output=True, text=True,
                             timeout=600)
        summ = open(os.path.join(td, f"{prec.upper()}BLAT3.SUMM")).read()
        stats = open(os.path.join(td, "stats.csv")).read().splitlines()
    assert out.returncode == 0, out.stderr[-3000:]
    p = prec.upper()
    for r in ROUTINES[prec]:
        assert f"{p}{r} PASSED THE TESTS OF ERROR-EXITS" in summ, summ
        assert f"{p}{r} PASSED THE COMPUTATIONAL TESTS" in summ, summ
    assert "FAIL" not in summ and "**" not in summ
    hdr = [h.strip() for h in stats[0].split(",")]
    vals = dict(zip(hdr, (int(v) for v in stats[1].split(","))))
    # every call that is not an error exit or a quick / alpha == 0 return ran on the GPU
    assert vals["DeviceCalls"] > 5000 and vals["Fallthrough"] == 0

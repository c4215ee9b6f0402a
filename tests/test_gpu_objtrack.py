"""Capturing an object-tracking file from a run (objtrack_out=<file>) and
replaying it -- the onboarding path the reference splits over an LD_PRELOAD
tracer and scripts (scripts/objtrace.sh:22-31, scripts/analyze_trace.py).

The chain driver allocates its operands A, x, y, z with four malloc calls and
runs dgemv / daxpy / ddot / dnrm2 / idamax on them.  A capture run tracks
every allocation (heuristic=true); the file it writes must name exactly those
four call sites and ordinals, and a replay with heuristic=objtrack:<file>
must track exactly those four buffers -- through the call-site rules alone as
well as through the ordinals alone."""
import os
import re
import subprocess
import tempfile

import pytest

import libgpublas_amd as gb

pytestmark = pytest.mark.gpu


def _chain(opts, n=256, iters=1):
    env = dict(os.environ)
    env["LD_PRELOAD"] = gb.PRELOAD_PATH
    env["GPUBLAS_OPTIONS"] = opts
    r = subprocess.run([gb.CHAIN_BENCH, "--n", str(n), "--iters", str(iters), "--warmup", "0"],
                       env=env, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr[-2000:]
    return r.stdout


def test_capture_then_replay():
    with tempfile.TemporaryDirectory() as td:
        cap = os.path.join(td, "chain.objtrack")
        _chain(f"heuristic=true;min_track=0;objtrack_out={cap}")
        text = open(cap).read()
        rules = [ln for ln in text.splitlines() if ln and not ln.startswith("#")]
        sites = [ln for ln in rules if ln.startswith("D @")]
        ords = [ln for ln in rules if ln.startswith("D #")]
        assert len(sites) == 4 and all(re.match(r"D @chain_bench\+0x[0-9a-f]+ ", s) for s in sites)
        assert len(ords) == 4
        sizes = sorted(int(re.search(r"reqsize=\[(\d+)\]", o).group(1)) for o in ords)
        assert sizes == [256 * 8] * 3 + [256 * 256 * 8]  # x, y, z and A
        for keep, name in ((sites, "sites"), (ords, "ordinals"), (rules, "both")):
            f = os.path.join(td, name)
            open(f, "w").write("\n".join(keep) + "\n")
            out = _chain(f"heuristic=objtrack:{f};min_track=0")
            assert '"tracked": 4' in out, (name, out)

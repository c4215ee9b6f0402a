"""Staged host-link copies of operands that are not resident (stage.hip).

Every untracked operand crosses the host link in pieces of at most 4 MiB
through pinned slots.  These tests cover the three ways a block is cut --
contiguous byte ranges, groups of whole columns with a leading dimension
larger than the row count (the gaps between columns must never be written),
and columns longer than a slot -- and that a process exits normally after the
copy pool has run."""
import os
import subprocess
import sys

import numpy as np
import pytest

import _oracle
from _oracle import uniform

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _padded(rows, cols, ld, seed, fill):
    buf = np.full((ld, cols), fill, order="F")
    buf[:rows, :] = uniform(seed, (rows, cols))
    return buf


@pytest.mark.parametrize("m,n,k,pad", [(1000, 3000, 64, 3),     # column groups, many pieces
                                       (600000, 3, 2, 5),       # a column > one slot
                                       (2048, 2048, 32, 0)])    # contiguous byte ranges
def test_dgemm_staged_layouts(gpu, ora, m, n, k, pad):
    lda, ldc = m + pad, m + pad
    A = _padded(m, k, lda, 1, -3.0)
    B = uniform(2, (k, n))
    C = _padded(m, n, ldc, 3, 7.25)
    R = C.copy(order="F")
    gpu.dgemm("N", "N", m, n, k, 0.5, A, lda, B, k, 1.5, C, ldc)
    ora.ora_dgemm(b"N", b"N", m, n, k, 0.5, _oracle.p(A), lda, _oracle.p(B), k, 1.5, _oracle.p(R),
                  ldc)
    if pad:
        assert np.all(C[m:, :] == 7.25), "gap rows between columns were written"
        assert np.all(A[m:, :] == -3.0)
    err = np.abs(C[:m] - R[:m]).max()
    bound = 4 * k * np.finfo(float).eps * (0.5 * np.abs(A[:m]).max() * np.abs(B).max() * k + 1.5 * 7.25)
    assert err <= bound, (err, bound)


def test_daxpy_long_vector_staged(gpu):
    n = 3_000_001  # 24 MB per vector: 6 pieces, the last one ragged
    x, y = uniform(4, (n,)), uniform(5, (n,))
    r = y + 0.75 * x
    gpu.daxpy(n, 0.75, x, 1, y, 1)
    assert np.max(np.abs(y - r)) <= 2 * np.finfo(float).eps * np.max(np.abs(r))


def test_process_exits_after_staged_copies():
    """The copy pool's threads must not keep a finished process alive."""
    code = ("import sys; sys.path.insert(0, %r); import numpy as np, libgpublas_amd as gb; "
            "gb.set_cutoff(0); n = 1024; a = np.ones((n, n), order='F'); "
            "c = np.zeros((n, n), order='F'); "
            "gb.dgemm('N', 'N', n, n, n, 1.0, a, n, a, n, 0.0, c, n); "
            "assert c[0, 0] == n; print('done', flush=True)") % ROOT
    r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=90)
    assert r.returncode == 0 and r.stdout.strip() == "done", (r.returncode, r.stderr[-2000:])

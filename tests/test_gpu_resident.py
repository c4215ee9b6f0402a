"""The resident reduction server (GPUBLAS_OPTIONS resident): DDOT / DNRM2 /
IDAMAX / IDAMIN / DASUM on tracked operands answered by a kernel that stays
on the device, instead of one launch per call.

Bars: the served results are bit-identical to the launched kernels' (the
same single-workgroup reduction runs in both), stream order is kept (a
reduction right after an asynchronous DAXPY sees its result: it is launched
behind it, the server taking only calls whose stream has drained), the
server restarts after its idle exit, and concurrent callers (which take the
launched path while another request is in flight) all get correct results.
"""
import threading
import time

import numpy as np
import pytest

from _oracle import p, uniform

pytestmark = pytest.mark.gpu


@pytest.fixture()
def resident(gpu):
    assert gpu.set_options("resident") == 0
    yield gpu
    gpu.set_options("resident=0")


def _chain(gb, n, iters, seed):
    A = gb.tracked_copy(uniform(seed, (n, n)))
    x = gb.tracked_copy(uniform(seed + 1, (n,)))
    y = gb.tracked_empty((n,))
    z = gb.tracked_copy(uniform(seed + 3, (n,)))
    out = []
    for _ in range(iters):
        gb.dgemv("N", n, n, 1.0, A, n, x, 1, 0.0, y, 1)
        gb.daxpy(n, 0.3, y, 1, z, 1)  # asynchronous: the reductions must wait for it
        out.append((gb.ddot(n, y, 1, z, 1), gb.dnrm2(n, z, 1), gb.idamax(n, z, 1),
                    gb.idamin(n, z, 1), gb.dasum(n, z, 1)))
    return out


def test_resident_chain_matches_launched(gpu, ora):
    gb = gpu
    n, iters = 8192, 4
    launched = _chain(gb, n, iters, 700)
    assert gb.set_options("resident") == 0
    try:
        gb.reset_stats()
        served = _chain(gb, n, iters, 700)
        st = gb.stats()
    finally:
        gb.set_options("resident=0")
    # DDOT follows the asynchronous DAXPY (queued behind it, it is launched);
    # the reductions after it find the stream drained and are served
    assert 4 * iters <= st["resident"] <= 5 * iters
    assert served == launched  # bit for bit, ties and all
    # and against the oracle on the same inputs
    A, x, z = uniform(700, (n, n)), uniform(701, (n,)), uniform(703, (n,))
    y = np.zeros(n)
    for _ in range(iters):
        ora.ora_dgemv(b"N", n, n, 1.0, p(A), n, p(x), 1, 0.0, p(y), 1)
        ora.ora_daxpy(n, 0.3, p(y), 1, p(z), 1)
    assert served[-1][2] == ora.ora_idamax(n, p(z), 1)
    assert abs(served[-1][1] - ora.ora_dnrm2(n, p(z), 1)) <= 1e-12 * served[-1][1]


def test_resident_restarts_after_idle_exit(resident):
    gb = resident
    n = 5000
    v = gb.tracked_copy(uniform(710, (n,)))
    w = gb.tracked_copy(uniform(711, (n,)))
    want = float(np.array(v) @ np.array(w))
    gb.ddot(n, v, 1, w, 1)  # uploads the operands (a call with an upload is launched)
    gb.reset_stats()
    r1 = gb.ddot(n, v, 1, w, 1)
    time.sleep(0.02)  # longer than the server's idle time: it exits, the next call restarts it
    r2 = gb.ddot(n, v, 1, w, 1)
    assert r1 == r2
    assert abs(r1 - want) <= 1e-12 * float(np.abs(np.array(v)) @ np.abs(np.array(w)))
    assert gb.stats()["resident"] == 2


def test_resident_strided_and_untracked(resident, ora):
    gb = resident
    n = 3001
    big = gb.tracked_copy(uniform(720, (3 * n,)))
    h = np.array(big)
    gb.dasum(3 * n, big, 1)  # uploads it (a call with an upload is launched)
    gb.reset_stats()
    assert gb.idamax(n, big, 3) == ora.ora_idamax(n, p(h), 3)
    r = gb.dnrm2(n, big, -2)
    assert abs(r - ora.ora_dnrm2(n, p(h), -2)) <= 1e-13 * r
    assert gb.stats()["resident"] == 2
    plain = uniform(721, (n,))  # not tracked: the launched path
    assert gb.idamax(n, plain, 1) == ora.ora_idamax(n, p(plain), 1)
    assert gb.stats()["resident"] == 2


def test_resident_concurrent_callers(resident):
    gb = resident
    n = 20000
    vs = [gb.tracked_copy(uniform(730 + t, (n,))) for t in range(8)]
    want = [float(np.sum(np.abs(np.array(v)))) for v in vs]
    errors = []

    def work(t):
        try:
            for _ in range(50):
                r = gb.dasum(n, vs[t], 1)
                if abs(r - want[t]) > 1e-12 * want[t]:
                    errors.append((t, r, want[t]))
        except Exception as e:  # noqa: BLE001 -- reported below
            errors.append(e)

    th = [threading.Thread(target=work, args=(t,)) for t in range(8)]
    for t in th:
        t.start()
    for t in th:
        t.join()
    assert not errors, errors[:3]

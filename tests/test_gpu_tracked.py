"""Residency of tracked operands on the GPU: results, coherence and the
hit/miss/upload accounting (the reference's gpuptr semantics,
runtime-mem.hpp:37-165, and its object tracker, lib/obj_tracker.c:602-637)."""
import numpy as np
import pytest

from _oracle import p, uniform

pytestmark = pytest.mark.gpu


def test_tracked_gemm_stays_resident(gpu):
    gb = gpu
    n = 256
    A = gb.tracked_copy(uniform(1, (n, n)))
    B = gb.tracked_copy(uniform(2, (n, n)))
    C = gb.tracked_empty((n, n))
    ref = np.array(A) @ np.array(B)
    gb.reset_stats()
    for _ in range(3):
        gb.dgemm("N", "N", n, n, n, 1.0, A, n, B, n, 0.0, C, n)
    s = gb.stats()
    assert s["hits"] == 9 and s["misses"] == 0
    assert s["uploads"] == 3  # each object crossed the host link once
    np.testing.assert_allclose(np.array(C), ref, rtol=0, atol=1e-12)
    assert gb.stats()["faults"] >= 1  # reading C pulled it back


def test_host_write_invalidates_device_copy(gpu):
    gb = gpu
    n = 128
    A = gb.tracked_copy(uniform(3, (n, n)))
    x = gb.tracked_copy(uniform(4, (n,)))
    y = gb.tracked_empty((n,))
    gb.dgemv("N", n, n, 1.0, A, n, x, 1, 0.0, y, 1)
    y1 = np.array(y)
    np.testing.assert_allclose(y1, np.array(A) @ np.array(x), atol=1e-12)
    x[5] += 1.0  # host write after upload: next call must see it
    gb.dgemv("N", n, n, 1.0, A, n, x, 1, 0.0, y, 1)
    np.testing.assert_allclose(np.array(y) - y1, np.array(A)[:, 5], atol=1e-12)


def test_interior_pointer_is_a_hit(gpu):
    gb = gpu
    big = gb.tracked_copy(uniform(5, (300, 300)))
    sub = big[50:, 40:]  # interior pointer, ld 300
    C = np.zeros((250, 250), order="F")
    B = np.asfortranarray(np.eye(260, 250))
    gb.reset_stats()
    gb.dgemm("N", "N", 250, 250, 260, 1.0, sub, 300, B, 260, 0.0, C, 250)
    s = gb.stats()
    assert s["hits"] == 1 and s["misses"] == 2
    np.testing.assert_allclose(C, np.array(big)[50:300, 40:300] @ np.eye(260, 250), atol=1e-12)


def test_chain_resident_small(gpu, ora):
    """The Level-1/2 chain on tracked buffers (n = 2048): after the first
    iteration nothing crosses the host link; idamax is bit-exact."""
    gb = gpu
    n = 2048
    A = gb.tracked_copy(uniform(7, (n, n)))
    x = gb.tracked_copy(uniform(8, (n,)))
    y = gb.tracked_empty((n,))
    z = gb.tracked_copy(uniform(9, (n,)))
    An, xn, zn = np.array(A), np.array(x), np.array(z)
    yr = np.zeros(n)
    for it in range(3):
        if it == 1:
            gb.reset_stats()
        gb.dgemv("N", n, n, 1.0, A, n, x, 1, 0.0, y, 1)
        gb.daxpy(n, 0.3, y, 1, z, 1)
        d = gb.ddot(n, y, 1, z, 1)
        r = gb.dnrm2(n, z, 1)
        im = gb.idamax(n, z, 1)
        ora.ora_dgemv(b"N", n, n, 1.0, p(An), n, p(xn), 1, 0.0, p(yr), 1)
        ora.ora_daxpy(n, 0.3, p(yr), 1, p(zn), 1)
    s = gb.stats()
    assert s["uploads"] == 0 and s["misses"] == 0 and s["faults"] == 0
    assert s["device_calls"] == 10
    assert im == ora.ora_idamax(n, p(zn), 1)
    assert d == pytest.approx(ora.ora_ddot(n, p(yr), 1, p(zn), 1), rel=1e-12)
    assert r == pytest.approx(ora.ora_dnrm2(n, p(zn), 1), rel=1e-13)
    np.testing.assert_allclose(np.array(z), zn, rtol=1e-12, atol=1e-12)

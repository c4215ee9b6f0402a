#!/usr/bin/env python3
"""bench.py -- headline measurement of libgpublas on one MI355X.

Workload (one "step"): the memory-bound Level-1/2 chain on device-resident
(tracked) FP64 buffers, n = 8192, entries uniform[-1,1]:
    y = dgemv('N', A[8192x8192], x); daxpy(0.3, y, z); ddot(y, z); dnrm2(z); idamax(z)
It runs as an ordinary BLAS client (libgpublas_amd/bin/chain_bench) under
LD_PRELOAD=libgpublas_preload.so with an object-tracking file naming its four
operand allocations, so the measured path is the drop-in path: Fortran-ABI
entry points, interposed malloc, HBM-resident operands, gfx950 kernels.

value = algorithmic bytes of the chain (537,460,736 B per step) / wall time of
the K timed steps, inputs already resident in HBM, with the resident
reduction server answering DDOT/DNRM2/IDAMAX (option `resident`; the line also
gives the chain with those launched as kernels).  The roofline object is
for the dominant kernel (dgemv 'N'): 8*(m*n + m + n) bytes per launch over its
average duration measured with HIP events on its launch stream.  The line also
carries the DGEMM figures the composite metric names (1024^3 and 8192^3 NN,
kernel TFLOP/s vs the 78.6 TFLOP/s FP64 peak), the serving trace (20,000 calls
from 16 threads, closed loop and three Poisson rates, batching loop vs one
launch per call) and the CPU baselines: the host
BLAS on all cores (cpu_baseline) and the oracle's C loops on one
(cpu_baseline_port).
"""
import argparse
import ctypes
import json
import os
import re
import subprocess
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "DGEMM TFLOP/s + L1/L2 GB/s vs roofline; batched-call p50/p99 & calls/s vs plain"
N = 8192
BYTES_PER_STEP = 8 * (N * N + 9 * N)
DGEMV_BYTES = 8 * (N * N + N + N)
HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6
SCIPY_OPENBLAS = "/usr/local/lib/python3.10/dist-packages/scipy.libs/libscipy_openblas-68440149.so"
SERVE_CALLS = 20000
SERVE_RATES = (0, 10000, 20000, 30000)  # offered calls/s; 0 = closed loop


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def run_chain(steps, warmup, profile, resident=True, seed=20240601):
    import libgpublas_amd as gb
    with tempfile.TemporaryDirectory() as td:
        objtrack = os.path.join(td, "chain.objtrack")
        with open(objtrack, "w") as f:
            # the driver's first four allocations: A, x, y, z
            f.write("D #0 A\nD #1 x\nD #2 y\nD #3 z\n")
        env = dict(os.environ)
        pre = env.get("LD_PRELOAD", "").strip()
        env["LD_PRELOAD"] = (pre + " " if pre else "") + gb.PRELOAD_PATH
        env["GPUBLAS_OPTIONS"] = chain_options(objtrack, resident) + (";profile" if profile else "")
        cmd = [gb.CHAIN_BENCH, "--n", str(N), "--iters", str(steps), "--warmup", str(warmup),
               "--seed", str(seed)]
        log("chain: " + " ".join(cmd))
        out = subprocess.run(cmd, env=env, capture_output=True, text=True, timeout=600)
    if out.returncode != 0:
        raise RuntimeError(f"chain_bench failed ({out.returncode}): {out.stderr[-2000:]}")
    res = json.loads(out.stdout.strip().splitlines()[-1])
    if res.get("tracked") != 4:
        raise RuntimeError(f"chain operands not all tracked: {res}")
    return res


def chain_options(objtrack, resident):
    """GPUBLAS_OPTIONS of the chain: the tracking file, no size cutoff (every
    call of the chain goes to the device) and, with `resident`, the resident
    reduction server for DDOT/DNRM2/IDAMAX."""
    return f"heuristic=objtrack:{objtrack};cutoff=0" + (";resident" if resident else "")


def run_host_blas_chain(steps, threads):
    if not os.path.exists(SCIPY_OPENBLAS):
        return None
    import libgpublas_amd as gb
    env = dict(os.environ)
    env["OPENBLAS_NUM_THREADS"] = str(threads)
    cmd = [gb.CHAIN_BENCH, "--n", str(N), "--iters", str(steps), "--warmup", "1",
           "--blas", SCIPY_OPENBLAS, "--prefix", "scipy_", "--threads", str(threads)]
    out = subprocess.run(cmd, env=env, capture_output=True, text=True, timeout=600)
    if out.returncode != 0:
        return None
    return json.loads(out.stdout.strip().splitlines()[-1])


def host_blas_chain_baseline(budget_s=12.0):
    """The reference's CPU path on the same chain: libgpublas forwards every
    call it does not run on the device to the host BLAS (dlsym RTLD_NEXT), so
    the CPU baseline is that host BLAS (scipy's OpenBLAS 0.3.29) on every core
    this process may run on, over a bounded sample of about `budget_s`."""
    cores = len(os.sched_getaffinity(0))
    probe = run_host_blas_chain(8, cores)
    if not probe:
        return None
    iters = int(min(3000, max(20, budget_s * 1e3 / max(probe["ms_per_iter"], 1e-3))))
    hb = run_host_blas_chain(iters, cores)
    if not hb:
        return None
    return {"value": round(hb["GBps"], 3), "unit": "GB/s", "cores": hb.get("blas_threads", cores),
            "kind": "reference", "nproc": os.cpu_count(), "affinity_cpus": cores,
            "library": "OpenBLAS 0.3.29 (scipy), the host BLAS the reference forwards to",
            "sample": f"{hb['iters']} chain steps at n={N} ({hb['seconds']:.1f} s), "
                      f"p50 {hb.get('iter_ms_p50', 0):.3f} ms/step"}


def host_dgemm_1024(cores):
    """The host BLAS (scipy's OpenBLAS) on the 1024^3 DGEMM (alpha 1, beta 0.5),
    best of 5 calls, on `cores` threads and on one."""
    try:
        import numpy as np
        import scipy.linalg.blas as sb
        from threadpoolctl import threadpool_limits
    except ImportError:
        return None
    n = 1024
    rng = np.random.default_rng(101)
    a, b = (np.asfortranarray(rng.uniform(-1, 1, (n, n))) for _ in range(2))
    c = np.asfortranarray(rng.uniform(-1, 1, (n, n)))
    out = {"cores": cores, "library": "OpenBLAS 0.3.29 (scipy)"}
    for key, th in (("tflops", cores), ("tflops_1_thread", 1)):
        with threadpool_limits(limits=th, user_api="blas"):
            best = None
            for _ in range(5 if th > 1 else 2):
                t0 = time.perf_counter()
                sb.dgemm(1.0, a, b, 0.5, c, overwrite_c=0)
                dt = time.perf_counter() - t0
                best = dt if best is None else min(best, dt)
        out[key] = round(2.0 * n ** 3 / best / 1e12, 3)
    return out


def cpu_oracle_chain(budget_s=12.0):
    """The oracle's C loops (one core) on a bounded sample of the same chain."""
    import numpy as np
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import _oracle
    from _oracle import p, uniform
    ora = _oracle.load()
    A = uniform(20240601, (N, N))
    x = uniform(20240602, (N,))
    z = uniform(20240604, (N,))
    y = np.zeros(N)
    iters = 0
    t0 = time.perf_counter()
    while True:
        ora.ora_dgemv(b"N", N, N, 1.0, p(A), N, p(x), 1, 0.0, p(y), 1)
        ora.ora_daxpy(N, 0.3, p(y), 1, p(z), 1)
        ora.ora_ddot(N, p(y), 1, p(z), 1)
        ora.ora_dnrm2(N, p(z), 1)
        ora.ora_idamax(N, p(z), 1)
        iters += 1
        el = time.perf_counter() - t0
        if el >= budget_s or iters >= 400:
            break
    return {"value": round(BYTES_PER_STEP * iters / el / 1e9, 3), "unit": "GB/s", "cores": 1,
            "kind": "port",
            "sample": f"{iters} chain steps at n={N} through oracle/refblas.c ({el:.1f} s)"}


def gemm_figures():
    """DGEMM NN kernel TFLOP/s at 1024^3 (alpha 1, beta 0.5) and 8192^3 on
    device buffers through the device C ABI, timed with HIP events."""
    import libgpublas_amd as gb
    lib = gb.library()
    lib.gpublas_dev_dgemm.argtypes = [ctypes.c_void_p, ctypes.c_char, ctypes.c_char, ctypes.c_int,
                                      ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_void_p,
                                      ctypes.c_int, ctypes.c_void_p, ctypes.c_int, ctypes.c_double,
                                      ctypes.c_void_p, ctypes.c_int]
    lib.gpublas_dev_prof_enable(1)
    out = {}
    for n, reps in ((1024, 50), (8192, 5)):
        bufs = [lib.gpublas_dev_alloc(n * n * 8) for _ in range(3)]
        for i, b in enumerate(bufs):
            lib.gpublas_dev_fill_uniform(None, 101 + i, n * n, b)
        # 1024^3 runs as split-K over 64 x 64 tiles (its 256 tiles fill half
        # the workgroup slots); 8192^3 on 128 x 128 tiles
        kerns = ["dgemm_fast_128x128"] if n >= 2048 else ["dgemm_splitk", "dgemm_fast_64x64"]
        for _ in range(2):
            lib.gpublas_dev_dgemm(None, b"N", b"N", n, n, n, 1.0, bufs[0], n, bufs[1], n, 0.5, bufs[2], n)
        lib.gpublas_dev_sync(None)
        lib.gpublas_dev_prof_reset()
        t0 = time.perf_counter()
        for _ in range(reps):
            lib.gpublas_dev_dgemm(None, b"N", b"N", n, n, n, 1.0, bufs[0], n, bufs[1], n, 0.5, bufs[2], n)
        lib.gpublas_dev_sync(None)
        wall = time.perf_counter() - t0
        tot, mn, mx = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
        for kern in kerns:
            cnt = lib.gpublas_dev_prof_query(kern.encode(), ctypes.byref(tot), ctypes.byref(mn),
                                             ctypes.byref(mx))
            if cnt:
                break
        flops = 2.0 * n ** 3
        kavg = tot.value / max(cnt, 1) / 1e3
        out[f"dgemm_nn_{n}"] = {
            "kernel": kern, "kernel_tflops": round(flops / kavg / 1e12, 3) if cnt else None,
            "wall_tflops": round(flops * reps / wall / 1e12, 3),
            "frac_of_fp64_peak": round(flops / kavg / 1e12 / FP64_PEAK_TFLOPS, 4) if cnt else None}
        for b in bufs:
            lib.gpublas_dev_release(b)
    # ZGEMM 8192^3 (C := A^H B + 0.5 C): one real product of m x 2n x 2k plus
    # the pack and combine passes; wall rate over the whole call, 8mnk flops
    lib.gpublas_dev_zgemm.argtypes = [ctypes.c_void_p, ctypes.c_char, ctypes.c_char, ctypes.c_int,
                                      ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_double,
                                      ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_int,
                                      ctypes.c_double, ctypes.c_double, ctypes.c_void_p, ctypes.c_int]
    n, reps = 8192, 3
    bufs = [lib.gpublas_dev_alloc(2 * n * n * 8) for _ in range(3)]
    for i, b in enumerate(bufs):
        lib.gpublas_dev_fill_uniform(None, 201 + i, 2 * n * n, b)
    call = lambda: lib.gpublas_dev_zgemm(None, b"C", b"N", n, n, n, 1.0, 0.0, bufs[0], n, bufs[1],
                                         n, 0.5, 0.0, bufs[2], n)
    call()
    lib.gpublas_dev_sync(None)
    t0 = time.perf_counter()
    for _ in range(reps):
        call()
    lib.gpublas_dev_sync(None)
    wall = time.perf_counter() - t0
    flops = 8.0 * n ** 3
    out[f"zgemm_cn_{n}"] = {"wall_tflops": round(flops * reps / wall / 1e12, 3),
                            "frac_of_fp64_peak": round(flops * reps / wall / 1e12 / FP64_PEAK_TFLOPS, 4)}
    for b in bufs:
        lib.gpublas_dev_release(b)
    return out


def level3_suite(reps):
    """The Level-3 sweep: GEMM at the four trans combinations with M, N, K in
    {4096, 8192} in FP64 and FP32; DSYRK lower n=8192 k=2048; DTRSM
    left/lower/no-trans/non-unit m=8192 nrhs=4096 (diagonally dominant A);
    DSYMM left/lower m=n=8192; DSYR2K lower n=8192 k=2048; DTRMM left/lower
    m=n=8192.  One JSON line per shape, kernel TFLOP/s from HIP events."""
    import itertools
    import libgpublas_amd as gb
    exe = os.path.join(gb.BIN_DIR, "blas_bench")
    env = dict(os.environ)
    env["GPUBLAS_OPTIONS"] = "profile;cutoff=0"
    runs = []
    for prec in ("d", "s"):
        for ta, tb in itertools.product("NT", "NT"):
            for m, n, k in itertools.product((4096, 8192), repeat=3):
                runs.append(["--op", "gemm", "--prec", prec, "--ta", ta, "--tb", tb, "--m", str(m),
                             "--n", str(n), "--k", str(k)])
        runs.append(["--op", "syrk", "--prec", prec, "--uplo", "L", "--trans", "N", "--n", "8192",
                     "--k", "2048"])
        runs.append(["--op", "trsm", "--prec", prec, "--side", "L", "--uplo", "L", "--ta", "N",
                     "--diag", "N", "--m", "8192", "--n", "4096"])
        runs.append(["--op", "symm", "--prec", prec, "--side", "L", "--uplo", "L", "--m", "8192",
                     "--n", "8192"])
        runs.append(["--op", "syr2k", "--prec", prec, "--uplo", "L", "--trans", "N", "--n", "8192",
                     "--k", "2048"])
        runs.append(["--op", "trmm", "--prec", prec, "--side", "L", "--uplo", "L", "--ta", "N",
                     "--diag", "N", "--m", "8192", "--n", "8192"])
    peak = {"d": FP64_PEAK_TFLOPS, "s": 157.3}
    results = []
    for r in runs:
        cmd = [exe, "--lib", gb.LIB_PATH, "--reps", str(reps), "--warmup", "1"] + r
        out = subprocess.run(cmd, env=env, capture_output=True, text=True, timeout=300)
        if out.returncode != 0:
            raise RuntimeError(f"{' '.join(cmd)} failed: {out.stderr[-1500:]}")
        res = json.loads(out.stdout.strip().splitlines()[-1])
        kms = sum(v[1] for v in res["kernels"].values()) / reps
        res["kernel_tflops"] = round(res["flops"] / (kms / 1e3) / 1e12, 3) if kms else None
        res["frac_of_peak"] = round(res["kernel_tflops"] / peak[res["prec"]], 4) if kms else None
        print(json.dumps(res), flush=True)
        results.append(res)
    gemm64 = [r["kernel_tflops"] for r in results if r["op"] == "gemm" and r["prec"] == "d"]
    gemm32 = [r["kernel_tflops"] for r in results if r["op"] == "gemm" and r["prec"] == "s"]
    print(json.dumps({"suite": "level3", "dgemm_tflops_min": min(gemm64), "dgemm_tflops_max": max(gemm64),
                      "sgemm_tflops_min": min(gemm32), "sgemm_tflops_max": max(gemm32)}), flush=True)


SPLIT_VARIANTS = [("bf16", s) for s in (3, 4, 5, 6)] + [("fp16", 6)] + \
    [("int8", s) for s in (4, 5, 6, 7, 8)]
SLICE_FMT = {"bf16": 0, "fp16": 1, "int8": 2}


def split_suite(n=8192, variants=SPLIT_VARIANTS):
    """Split-precision DGEMM M=N=K=n against plain FP64 DGEMM: (a) uniform[-1,1];
    (b) uniform entries with every row of A and column of B scaled by
    2^U[-20,20].  Errors are measured against the host BLAS (OpenBLAS, all
    cores) on the same inputs; times are kernel times from HIP events (split:
    exponent/splitting pass + slice products)."""
    import numpy as np
    import scipy.linalg.blas as sb
    import libgpublas_amd as gb
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from _oracle import uniform
    lib = gb.library()
    vp, ci, cd = ctypes.c_void_p, ctypes.c_int, ctypes.c_double
    lib.gpublas_dev_dgemm.argtypes = [vp, ctypes.c_char, ctypes.c_char, ci, ci, ci, cd, vp, ci, vp,
                                      ci, cd, vp, ci]
    lib.gpublas_dev_dgemm_split_fmt.argtypes = [vp, ci, ci, ctypes.c_char, ctypes.c_char, ci, ci,
                                                ci, cd, vp, ci, vp, ci, cd, vp, ci]
    lib.gpublas_dev_prof_enable(1)
    bytes_ = n * n * 8
    dA, dB, dC = (lib.gpublas_dev_alloc(bytes_) for _ in range(3))
    host = np.empty((n, n), order="F")

    def ktime(names):
        tot = 0.0
        for nm in names:
            t = ctypes.c_double()
            lib.gpublas_dev_prof_query(nm.encode(), ctypes.byref(t), None, None)
            tot += t.value
        return tot

    for dist in ("uniform", "rowcol_scaled"):
        A = uniform(501, (n, n))
        B = uniform(502, (n, n))
        if dist == "rowcol_scaled":
            A *= (2.0 ** np.floor(uniform(503, (n,)) * 20.0))[:, None]
            B *= (2.0 ** np.floor(uniform(504, (n,)) * 20.0))[None, :]
        lib.gpublas_dev_h2d(dA, A.ctypes.data, bytes_, None)
        lib.gpublas_dev_h2d(dB, B.ctypes.data, bytes_, None)
        lib.gpublas_dev_sync(None)
        t0 = time.perf_counter()
        R = sb.dgemm(1.0, A, B)
        host_s = time.perf_counter() - t0
        normR = np.linalg.norm(R)
        maxR = np.max(np.abs(R))
        runs = [("fp64", 0, None)] + [("split", s, f) for f, s in variants]
        for kind, s, fmt in runs:
            for rep in range(3):
                if rep == 1:
                    lib.gpublas_dev_prof_reset()
                if kind == "fp64":
                    lib.gpublas_dev_dgemm(None, b"N", b"N", n, n, n, 1.0, dA, n, dB, n, 0.0, dC, n)
                else:
                    lib.gpublas_dev_dgemm_split_fmt(None, s, SLICE_FMT[fmt], b"N", b"N", n, n, n,
                                                    1.0, dA, n, dB, n, 0.0, dC, n)
            lib.gpublas_dev_sync(None)
            ms = ktime(["dgemm_fast_128x128"] if kind == "fp64" else
                       ["ozaki_split", "ozaki_i8" if fmt == "int8" else "ozaki_gemm"]) / 2
            lib.gpublas_dev_d2h(host.ctypes.data, dC, bytes_, None)
            lib.gpublas_dev_sync(None)
            err_norm = float(np.linalg.norm(host - R) / normR)
            err_max = float(np.max(np.abs(host - R)) / maxR)
            npairs = s * (s + 1) // 2 if s else 0
            rec = {"suite": "split", "dist": dist, "n": n, "kind": kind, "slices": s, "slice_format": fmt,
                   "slice_products": npairs, "kernel_ms": round(ms, 3),
                   "effective_tflops": round(2.0 * n ** 3 / (ms / 1e3) / 1e12, 2),
                   ("int8_mfma_tops" if fmt == "int8" else "bf16_mfma_tflops"):
                       round(npairs * 2.0 * n ** 3 / (ms / 1e3) / 1e12, 1) if s else None,
                   "normwise_rel_err": err_norm, "max_rel_err": err_max,
                   "reference": f"OpenBLAS dgemm on the host ({host_s:.1f} s)"}
            print(json.dumps(rec), flush=True)
    for b in (dA, dB, dC):
        lib.gpublas_dev_release(b)


def serve_run(mode, rate, calls, threads=16, options="serve=64G"):
    """One serving trace (csrc/bench/serve_bench.c) -> its JSON record."""
    import libgpublas_amd as gb
    cmd = [gb.SERVE_BENCH, "--lib", gb.LIB_PATH, "--calls", str(calls), "--threads", str(threads),
           "--rate", str(rate)]
    cmd += ["--plain"] if mode == "plain" else ["--options", options]
    out = subprocess.run(cmd, capture_output=True, text=True, timeout=900)
    if out.returncode != 0:
        raise RuntimeError(f"serve_bench failed ({out.returncode}): {out.stderr[-2000:]}")
    return json.loads(out.stdout.strip().splitlines()[-1])


def serve_suite(calls=20000, rates=(0,)):
    """Serving trace: 16 threads, Poisson arrivals (rate 0 = closed loop,
    i.e. saturation), one-launch-per-call vs the batching loop (eager
    launches, the default; graph replay; batches of one)."""
    variants = [("plain", ""), ("serve", "serve=64G"), ("serve", "serve=64G;serve_graph=1"),
                ("serve", "serve=64G;serve_batch=1")]
    for rate in rates:
        base = None
        for mode, opts in variants:
            rec = serve_run(mode, rate, calls, options=opts or "serve=64G")
            rec["suite"] = "serve"
            if mode == "plain":
                base = rec
            elif base:
                rec["speedup_vs_plain"] = round(rec["calls_per_s"] / base["calls_per_s"], 3)
                rec["p99_vs_plain"] = round(rec["p99_us"] / base["p99_us"], 3)
            print(json.dumps(rec), flush=True)


def traffic_from_profile():
    path = os.path.join(ROOT, "profiles", "dgemv_n_chain8192_pmc.txt")
    if not os.path.exists(path):
        return None
    m = re.search(r"traffic_bytes_per_launch:\s*([0-9.eE+]+)", open(path).read())
    return float(m.group(1)) if m else None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baselines")
    ap.add_argument("--suite", default="chain", choices=["chain", "level3", "split", "serve"],
                    help="chain: the headline line (default); level3: the GEMM/SYRK/TRSM sweep; "
                         "split: split-precision DGEMM vs FP64; serve: the serving trace")
    ap.add_argument("--calls", type=int, default=20000, help="serve: calls per trace")
    ap.add_argument("--rates", default="0", help="serve: comma-separated offered rates "
                                                "(calls/s; 0 = closed loop)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one MI355X: the library drives one GPU per process "
              f"(asked for --gpus {args.gpus}).", file=sys.stderr)
        sys.exit(2)
    if args.suite == "level3":
        level3_suite(reps=3)
        return
    if args.suite == "split":
        split_suite()
        return
    if args.suite == "serve":
        serve_suite(args.calls, tuple(float(r) for r in args.rates.split(",")))
        return

    # value: the K timed steps with no instrumentation; roofline: a second
    # run with HIP events around every kernel on its launch stream
    chain = run_chain(args.steps, args.warmup, profile=False)
    prof = run_chain(args.steps, args.warmup, profile=True)
    # the same chain with every reduction launched as a kernel (no server)
    plain_chain = run_chain(args.steps, args.warmup, profile=False, resident=False)
    secs = chain["seconds"]
    value = BYTES_PER_STEP * args.steps / secs / 1e9
    kcount = prof.get("k_dgemv_n_count", 0)
    kavg_s = prof["k_dgemv_n_ms"] / kcount / 1e3 if kcount else None
    achieved = DGEMV_BYTES / kavg_s / 1e9 if kavg_s else None
    roofline = {"bound": "hbm", "achieved": round(achieved, 2) if achieved else None,
                "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4) if achieved else None,
                "traffic": traffic_from_profile(), "kernel": "dgemv_n",
                "kernel_avg_us": round(kavg_s * 1e6, 3) if kavg_s else None,
                "algorithmic_bytes_per_launch": DGEMV_BYTES}
    kernels = {k[2:-3]: round(prof[k] / max(prof[k[:-3] + "_count"], 1) * 1e3, 3)
               for k in prof if k.startswith("k_") and k.endswith("_ms")}

    gemm = gemm_figures()
    # the serving trace itself: 20,000 calls from 16 threads, closed loop and
    # Poisson arrivals at 10k / 20k / 30k calls/s, the batching loop against
    # one launch per call
    srv = {}
    try:
        rates = []
        for rate in SERVE_RATES:
            plain = serve_run("plain", rate, SERVE_CALLS)
            batched = serve_run("serve", rate, SERVE_CALLS)
            rates.append({"offered_calls_per_s": rate or "closed loop",
                          "calls_per_s": round(batched["calls_per_s"], 1),
                          "plain_calls_per_s": round(plain["calls_per_s"], 1),
                          "p50_us": batched["p50_us"], "p99_us": batched["p99_us"],
                          "plain_p50_us": plain["p50_us"], "plain_p99_us": plain["p99_us"],
                          "mean_batch": batched["mean_batch"]})
        closed = rates[0]
        srv = {"workload": f"16 threads, {SERVE_CALLS} calls: 60% D/SGEMM logU[32,2048]^3, "
                           "30% DGEMV logU[256,16384]^2, 10% DTRSM m logU[64,1024] x 256; "
                           "tracked operands",
               "calls_per_s": closed["calls_per_s"],
               "plain_calls_per_s": closed["plain_calls_per_s"],
               "speedup_vs_plain": round(closed["calls_per_s"] / closed["plain_calls_per_s"], 3),
               "p50_us": closed["p50_us"], "p99_us": closed["p99_us"],
               "plain_p50_us": closed["plain_p50_us"], "plain_p99_us": closed["plain_p99_us"],
               "mean_batch": closed["mean_batch"], "rates": rates}
    except Exception as e:  # reported, not fatal for the headline line
        srv = {"error": str(e)[:300]}
    line = {
        "metric": METRIC, "value": round(value, 3), "unit": "GB/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(1e3 * secs / args.steps, 4),
        "ms_per_step_p10_p50_p90": [round(chain.get("iter_ms_p10", 0), 4),
                                    round(chain.get("iter_ms_p50", 0), 4),
                                    round(chain.get("iter_ms_p90", 0), 4)],
        "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic (SplitMix64 uniform[-1,1], seed 20240601)",
        "config": {"workload": "level-1/2 chain on tracked buffers: dgemv N 8192x8192 + daxpy + "
                               "ddot + dnrm2 + idamax, FP64, LD_PRELOAD drop-in",
                   "n": N, "bytes_per_step": BYTES_PER_STEP,
                   "options": chain_options("<file>", True)},
        "value_without_resident": round(BYTES_PER_STEP * args.steps / plain_chain["seconds"] / 1e9, 3),
        "ms_per_step_p50_without_resident": round(plain_chain.get("iter_ms_p50", 0), 4),
        "roofline": roofline,
        "kernel_avg_us": kernels,
        "chain_results": {"ddot": chain["ddot"], "dnrm2": chain["dnrm2"], "idamax": chain["idamax"]},
        "dgemm": gemm,
        "serving": srv,
    }
    if not args.no_cpu:
        cores = len(os.sched_getaffinity(0))
        port = cpu_oracle_chain()
        hb = host_blas_chain_baseline()
        if hb:
            hb1 = run_host_blas_chain(10, 1)
            if hb1:
                hb["value_1_thread"] = round(hb1["GBps"], 3)
            line["cpu_baseline"] = hb
            line["cpu_baseline_port"] = port
        else:  # no host BLAS on this machine: the oracle port is the baseline
            line["cpu_baseline"] = port
        hg = host_dgemm_1024(cores)
        if hg:
            line["dgemm"]["host_blas_1024"] = hg
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()

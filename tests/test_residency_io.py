"""Tracked memory is ordinary memory to system calls, stdio, signal handlers
and fork (the reference's managed memory, blas2cuda.c:127-148, gives that by
page migration; here the preload library's interposers and fault path do).

tests/residency_probe.c is a plain C program run under
LD_PRELOAD=libgpublas_preload.so; each case prints "ok <case>":
  - write/writev/pwrite/fwrite of a DEVICE-state object write the device's
    data (no EFAULT), and leave it SHARED;
  - read/pread/preadv/fread into a SHARED (read-only) view succeed and demote
    the object to HOST, so the next BLAS call uploads what was read;
  - a SIGSEGV handler the application installs after the library's (with
    sigaction or signal) receives every fault that is not a tracked object's,
    tracked faults stay the library's, and sigaction reports the
    application's own handler back;
  - faults from 8 threads at once are served by the helper thread;
  - fork: the child sees a snapshot (device data downloaded at the fork),
    its writes do not reach the parent and vice versa, the child's objects
    are no longer tracked, the parent keeps its objects (and, on the GPU,
    runs BLAS on them afterwards).
CPU: states set with the test hook gpublas_debug_set_residency (no device
copy).  GPU: objects made device-newer by dscal_ on the GPU.
"""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PRELOAD = os.path.join(ROOT, "libgpublas_amd", "lib", "libgpublas_preload.so")
CASES = ["write_device_object", "write_device_object_contents", "write_leaves_shared",
         "pread_into_shared_object", "pread_into_shared_object_contents", "read_demotes_to_host",
         "writev_device_objects", "writev_contents", "preadv_into_shared",
         "fwrite_device_object", "fread_into_shared", "app_sigaction_installs",
         "app_sees_default_before", "tracked_fault_not_chained", "foreign_fault_chained",
         "app_sees_its_handler", "signal_handler_chained", "concurrent_faults",
         "fork_child_sees_snapshot", "fork_parent_isolated", "fork_parent_still_tracks"]


def _run_probe(tmp_path, mode, options):
    exe = tmp_path / "residency_probe"
    subprocess.run(["gcc", "-O1", "-std=gnu11", "-Wall", "-o", str(exe),
                    os.path.join(ROOT, "tests", "residency_probe.c"), "-ldl", "-lpthread"],
                   check=True)
    env = dict(os.environ, LD_PRELOAD=PRELOAD, GPUBLAS_OPTIONS=options)
    log = tmp_path / "probe.log"
    with open(log, "w") as f:
        try:
            rc = subprocess.run([str(exe), mode, str(tmp_path)], env=env, stdout=f,
                                stderr=subprocess.STDOUT, timeout=60).returncode
        except subprocess.TimeoutExpired:
            rc = None
    text = log.read_text()
    print(text, flush=True)  # the cases passed before a stall are named
    assert rc is not None, "residency probe did not finish in 60 s; last lines:\n" + text[-2000:]
    return subprocess.CompletedProcess([], rc, text, "")


def _check(out, cases):
    lines = out.stdout.splitlines()
    assert out.returncode == 0, out.stdout[-2000:]
    for c in cases:
        assert f"ok {c}" in lines, c


def test_residency_io_signals_fork_cpu(tmp_path):
    if not os.path.exists(PRELOAD):
        pytest.fail(f"{PRELOAD} is missing: build it with __graft_entry__.build()")
    _check(_run_probe(tmp_path, "cpu", "heuristic=false"), CASES)


@pytest.mark.gpu
def test_residency_io_signals_fork_gpu(tmp_path):
    out = _run_probe(tmp_path, "gpu", "heuristic=false;cutoff=0")
    _check(out, CASES + ["blas_after_read", "blas_after_fork"])

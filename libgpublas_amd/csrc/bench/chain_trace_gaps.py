#!/usr/bin/env python3
"""chain_trace_gaps.py <rocprofv3 output dir> -- the per-step gap table of
chain_bench from a `rocprofv3 --runtime-trace --output-format csv` run:
per kernel of a step, its duration and the time from the start of its
hipLaunchKernel call to its start on the device; per step, the time from the
last kernel's end to the next step's DGEMV launch call (the host sees the
step's scalars and runs the rest of the step: reductions answered by the
resident server run in this window) and the device-idle share.  Steps run
from one gemv_n_stream start to the next; the first five are skipped.
"""
import csv
import glob
import os
import statistics as st
import sys


def short(name):
    for key in ("res_server", "gemv_n_stream", "gemv_n_fold", "axpy", "dot", "nrm2", "iamax"):
        if key in name:
            return key
    return name[:30]


def main(d):
    kt = list(csv.DictReader(open(glob.glob(os.path.join(d, "*kernel_trace.csv"))[0])))
    api = list(csv.DictReader(open(glob.glob(os.path.join(d, "*hip_api_trace.csv"))[0])))
    launch = {r["Correlation_Id"]: int(r["Start_Timestamp"]) for r in api
              if "LaunchKernel" in r["Function"]}
    ks = sorted((int(r["Start_Timestamp"]), int(r["End_Timestamp"]), short(r["Kernel_Name"]),
                 launch.get(r["Correlation_Id"])) for r in kt)
    server = [k for k in ks if k[2] == "res_server"]
    ks = [k for k in ks if k[2] != "res_server"]
    idx = [i for i, k in enumerate(ks) if k[2] == "gemv_n_stream"]
    rows = []
    for a, b in zip(idx[5:], idx[6:]):
        seq = ks[a:b]
        rows.append({"step": (ks[b][0] - ks[a][0]) / 1e3,
                     "kern": {k[2]: (k[1] - k[0]) / 1e3 for k in seq},
                     "l2s": {k[2]: (k[0] - k[3]) / 1e3 for k in seq if k[3]},
                     "tail": (ks[b][3] - seq[-1][1]) / 1e3 if ks[b][3] else None})
    print(f"dispatches: {len(ks)} + {len(server)} resident-server dispatch(es); steps: {len(rows)}")
    sv = sorted(r["step"] for r in rows)
    print(f"step (dgemv start to next dgemv start), us: p10 {sv[len(sv) // 10]:.1f}  "
          f"p50 {st.median(sv):.1f}  p90 {sv[len(sv) * 9 // 10]:.1f}")
    print("| kernel | duration p50 (us) | launch call -> kernel start p50 (us) |")
    print("|---|---|---|")
    for nm in rows[0]["kern"]:
        lat = [r["l2s"][nm] for r in rows if nm in r["l2s"]]
        print(f"| {nm} | {st.median(r['kern'][nm] for r in rows):.2f} | "
              f"{st.median(lat):.2f} |" if lat else "- |")
    tail = [r["tail"] for r in rows if r["tail"] is not None]
    print(f"last kernel end -> next dgemv launch call, p50: {st.median(tail):.2f} us")
    busy = [sum(r["kern"].values()) for r in rows]
    print(f"device busy per step p50 {st.median(busy):.1f} us; "
          f"idle p50 {st.median([r['step'] - b for r, b in zip(rows, busy)]):.1f} us")


if __name__ == "__main__":
    main(sys.argv[1])

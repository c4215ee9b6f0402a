#!/usr/bin/env python3
"""trace_busy.py <kernel_trace.csv> <wall_s> -- share of the last `wall_s`
seconds of a rocprofv3 kernel trace during which at least one kernel ran (the
union of the kernels' intervals), the summed kernel time (its ratio to the
window is the mean number of kernels running at once), and the kernels with
the most time in the window."""
import collections
import csv
import sys


def main(path, wall):
    rows = list(csv.DictReader(open(path)))
    iv = sorted((int(r["Start_Timestamp"]), int(r["End_Timestamp"]), r["Kernel_Name"]) for r in rows)
    t1 = max(e for _, e, _ in iv)
    cut = t1 - wall * 1e9
    sel = [(max(s, cut), e, n) for s, e, n in iv if e > cut]
    busy, cs, ce, total = 0, None, None, 0
    for s, e, _ in sel:
        total += e - s
        if cs is None:
            cs, ce = s, e
        elif s <= ce:
            ce = max(ce, e)
        else:
            busy += ce - cs
            cs, ce = s, e
    if cs is not None:
        busy += ce - cs
    span = t1 - cut
    print(f"window {span / 1e6:.1f} ms: {len(sel)} kernels, device busy {busy / 1e6:.2f} ms "
          f"({100 * busy / span:.1f}%), kernel time {total / 1e6:.1f} ms "
          f"({total / span:.2f} kernels at once on average)")
    agg = collections.defaultdict(lambda: [0, 0])
    for s, e, n in sel:
        k = n.replace("(anonymous namespace)::", "").replace("void ", "")
        k = k.split("(")[0][:48]
        agg[k][0] += 1
        agg[k][1] += e - s
    for k, (c, d) in sorted(agg.items(), key=lambda x: -x[1][1])[:6]:
        print(f"  {k:48s} {c:6d} launches {d / 1e6:8.2f} ms")


if __name__ == "__main__":
    main(sys.argv[1], float(sys.argv[2]))

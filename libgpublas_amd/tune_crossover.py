"""Measure this machine's host/GPU crossover table for cutoff=auto.

    python -m libgpublas_amd.tune_crossover OUT_FILE [--max-n 2048]

Times square problems of every routine on the host BLAS and on the GPU with
operands that are not resident, writes the resulting table (and the timings
behind it, as comments) to OUT_FILE and prints it.  Programs use it with
GPUBLAS_OPTIONS="cutoff=auto;xover=OUT_FILE".  Needs the GPU."""
import argparse
import sys

import libgpublas_amd as gb


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("out")
    ap.add_argument("--max-n", type=int, default=2048,
                    help="largest Level-3 size (Level 2: twice that; Level 1: 4096 times)")
    a = ap.parse_args(argv)
    if not gb.device_ready():
        print("tune_crossover: no usable gfx950 GPU", file=sys.stderr)
        return 1
    gb.tune_crossover(a.out, a.max_n)
    sys.stdout.write(open(a.out).read())
    return 0


if __name__ == "__main__":
    sys.exit(main())

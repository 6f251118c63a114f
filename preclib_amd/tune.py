"""Plan autotuner: times every NTT plan variant per transform length on this
GPU and writes preclib_amd/ntt_plans.txt, which libpgpu.so reads at pg_init.

Usage: python -m preclib_amd.tune [--lo 12] [--hi 26] [--reps 5]
"""
import argparse
import os

from preclib_amd import pgpu

HERE = os.path.dirname(os.path.abspath(__file__))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--lo", type=int, default=12)
    ap.add_argument("--hi", type=int, default=25)
    ap.add_argument("--reps", type=int, default=5)
    ap.add_argument("--out", default=os.path.join(HERE, "ntt_plans.txt"))
    args = ap.parse_args()
    pgpu.init(0)
    res = pgpu.ntt_autotune(args.lo, args.hi, args.reps)
    lines = ["# NTT plan variant per transform length (wide mode), from preclib_amd.tune",
             "# logn variant   ms per variant 0..3 (-1: same as the default plan)"]
    for logn in sorted(res):
        v, ms = res[logn]
        lines.append("%d %d   %s" % (logn, v, " ".join("%.4f" % m for m in ms)))
        print(lines[-1], flush=True)
    with open(args.out, "w") as f:
        f.write("\n".join(lines) + "\n")
    print("wrote", args.out)


if __name__ == "__main__":
    main()

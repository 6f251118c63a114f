"""Plan autotuner: times the NTT plan variants per transform length on this
GPU and writes preclib_amd/ntt_plans.txt, which libpgpu.so reads at pg_init.

Usage: python -m preclib_amd.tune [--lo 12] [--hi 25] [--reps 5]
       python -m preclib_amd.tune --batch [--lo 12] [--hi 18]

The table has two parts.  "logn variant" lines are the single-product plans:
every variant of 2^lo .. 2^hi points timed as one product at a time
(pg_ntt_autotune).  "batch logn variant" lines are the plans of batched
launches of 2^21 points or more (pg_batch_plan), where throughput counts:
--batch tunes them on the 4096-product uneven batch itself (the default
sizes are its leaf sizes), by coordinate descent, two sweeps over the sizes,
each candidate timed as whole graph replays of the planned batch; a size
keeps its variant unless another is faster by 0.5% on a second timing, and
the chosen table must reproduce the batch's golden digest of digests
(tests/golden/products.json) before it is written.  Either mode rewrites
only the lines of the sizes it tunes and keeps the others (those of --out,
or of the library's own table where --out does not exist yet).
"""
import argparse
import hashlib
import json
import os
import time

from preclib_amd import pgpu

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = os.path.join(os.path.dirname(HERE), "tests", "golden", "products.json")
HEADER = [
    "# NTT plan variants per transform length (wide mode), written by preclib_amd.tune",
    "# 'logn variant  ms..': single products (python -m preclib_amd.tune), ms per product",
    "#   for variants 0..15 (-1: the same plan as the default)",
    "# 'batch logn variant  ms..': batched launches of 2^21 points or more (python -m",
    "#   preclib_amd.tune --batch), ms per replay of the 4096-product uneven batch with",
    "#   each variant at that size (last sweep)",
]


def _read_table(path):
    """({logn: line}, {logn: batch line}) of a plan file."""
    rows, brows = {}, {}
    if not os.path.exists(path):
        path = os.path.join(HERE, "ntt_plans.txt")
    for line in open(path):
        f = line.split()
        if line.startswith("#") or len(f) < 2:
            continue
        if f[0] == "batch" and len(f) >= 3:
            brows[int(f[1])] = line.rstrip("\n")
        elif f[0].isdigit():
            rows[int(f[0])] = line.rstrip("\n")
    return rows, brows


def _write_table(path, rows, brows):
    with open(path, "w") as fh:
        fh.write("\n".join(HEADER + [rows[k] for k in sorted(rows)] + [brows[k] for k in sorted(brows)]) + "\n")
    print("wrote", path)


def _batch_bufs():
    from preclib_amd import workloads
    dev = []
    for k, (an, bn) in enumerate(workloads.batch_shapes()):
        a, b = workloads.pair(an, bn, "uniform", k)
        dev.append((pgpu.DevBuf(an + bn), pgpu.DevBuf.of(a), an, pgpu.DevBuf.of(b), bn))
    return dev


def _batch_ms(dev, table, reps=12):
    """Mean ms of one graph replay of the planned batch under batch plan table."""
    pgpu.set_ntt_batch_plans(table)
    b = pgpu.DevBatch(dev)
    try:
        for _ in range(3):
            b.run()
        pgpu.sync()
        t0 = time.perf_counter()
        for _ in range(reps):
            b.run()
        pgpu.sync()
        return (time.perf_counter() - t0) / reps * 1e3
    finally:
        b.free()


def _batch_golden(dev, table):
    pgpu.set_ntt_batch_plans(table)
    b = pgpu.DevBatch(dev)
    try:
        b.run()
        pgpu.sync()
        per = []
        for (r, _, an, _, bn) in dev:
            x = r.download(an + bn)
            n = len(x)
            while n and x[n - 1] == 0:
                n -= 1
            per.append(hashlib.sha256(x[:n].astype("<u8").tobytes()).hexdigest())
    finally:
        b.free()
    return hashlib.sha256("".join(per).encode()).hexdigest() == json.load(open(GOLDEN))["batch"]["digest_of_digests"]


def tune_batch(lo, hi, out):
    pgpu.init(0)
    dev = _batch_bufs()
    single, bt = pgpu.ntt_plans(), pgpu.ntt_batch_plans()
    base = [b if b >= 0 else s for s, b in zip(single, bt)]
    cur = list(base)
    last = {}
    for sweep in range(2):
        for logn in range(lo, hi + 1):
            res = {}
            for v in range(pgpu.NTT_VARIANTS):
                t = list(cur)
                t[logn] = v
                res[v] = _batch_ms(dev, t)
            vb = min(res, key=res.get)
            t = list(cur)
            t[logn] = vb
            tb, tc = min(res[vb], _batch_ms(dev, t)), min(res[cur[logn]], _batch_ms(dev, cur))
            if tb < tc * 0.995:
                cur[logn] = vb
            last[logn] = res
            print("sweep %d  2^%d: variant %d (%.3f ms per batch; best %d %.3f)" % (sweep, logn, cur[logn], tc, vb, tb),
                  flush=True)
    t_new = min(_batch_ms(dev, cur) for _ in range(3))
    t_old = min(_batch_ms(dev, base) for _ in range(3))
    print("batch %.3f ms (table before: %.3f ms)" % (t_new, t_old), flush=True)
    if not _batch_golden(dev, cur):
        raise SystemExit("tune: the chosen table does not reproduce the batch's golden digest")
    rows, brows = _read_table(out)
    for logn in range(lo, hi + 1):
        brows[logn] = "batch %d %d   %s" % (logn, cur[logn],
                                             " ".join("%.4f" % last[logn][v] for v in range(pgpu.NTT_VARIANTS)))
    _write_table(out, rows, brows)


def tune_single(lo, hi, reps, out):
    pgpu.init(0)
    res = pgpu.ntt_autotune(lo, hi, reps)
    rows, brows = _read_table(out)
    for logn in sorted(res):
        v, ms = res[logn]
        rows[logn] = "%d %d   %s" % (logn, v, " ".join("%.4f" % m for m in ms))
        print(rows[logn], flush=True)
    _write_table(out, rows, brows)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--lo", type=int, default=12)
    ap.add_argument("--hi", type=int, default=None)
    ap.add_argument("--reps", type=int, default=5)
    ap.add_argument("--batch", action="store_true", help="tune the batch plans of 2^lo..2^hi (default 12..18)")
    ap.add_argument("--out", default=os.path.join(HERE, "ntt_plans.txt"))
    args = ap.parse_args()
    if args.batch:
        tune_batch(args.lo, 18 if args.hi is None else args.hi, args.out)
    else:
        tune_single(args.lo, 25 if args.hi is None else args.hi, args.reps, args.out)


if __name__ == "__main__":
    main()

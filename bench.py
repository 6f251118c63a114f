#!/usr/bin/env python3
"""Throughput of the plain GPU product path: product limbs/s at 2^22 x 2^22.

One step = one bit-exact 2^22 x 2^22-limb product (preclib's mul_ntt path:
digits, forward NTTs, pointwise product, inverse NTT, CRT, carry) with both
operands already resident in HBM and the product left in HBM.  Operands are
synthetic (splitmix64 uniform limbs, top limb nonzero).

Prints one JSON line.  Besides the contract fields it carries
  roofline     -- the dominant kernel: algorithmic bytes per launch (its share
                  of the byte model B = 16(Na+Nb) + 48 P n) over its average
                  launch time from HIP events on the library's stream;
  cpu_baseline -- the reference's own compiled product (oracle/_ref, built
                  from the reference sources) on the same operands, 1 core;
  product_roofline -- the whole product against the same byte model.
"""
import argparse
import hashlib
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK = 8.0e12          # MI355X published HBM3E bandwidth, bytes/s
# integer VALU issue peak: 256 CUs, one wave64 instruction per cycle per CU
# (4 SIMDs of 16 lanes, 4 cycles each), 2.4 GHz peak engine clock
VALU_PEAK = 256 * 2.4e9    # wave-instructions/s (= 39.3e12 32-bit lane ops/s)
NPRIMES = 2
RAMP_S = 0.06              # untimed load before the warm-up steps (clock ramp)


def host_cpu():
    """Model name of the host CPU and the number of cores this process may use."""
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        ncpu = len(os.sched_getaffinity(0))
    except Exception:
        ncpu = os.cpu_count()
    return {"host_cpu": model, "host_nproc": ncpu, "host_cpu_count": os.cpu_count()}


def model_bytes(an, bn, n):
    """Per-kernel share of the reference byte model B = 16(Na+Nb) + 48 P n with
    the reference's P = 2 and its transform length n (base-2^16 digits)."""
    P = NPRIMES
    return {
        "ntt_col_fwd": 8.0 * (an + bn) + 16.0 * P * n,
        "ntt_row": 24.0 * P * n,
        "ntt_col_inv": 8.0 * P * n,
        "carry": 8.0 * (an + bn),
    }


def transform_len(a, b):
    def nd(top):
        d = 0
        top = int(top)
        while top:
            d += 1
            top >>= 16
        return d
    da = 4 * (len(a) - 1) + nd(a[-1])
    db = 4 * (len(b) - 1) + nd(b[-1])
    n = 1
    while n < da + db:
        n <<= 1
    return n


def cpu_baseline(a, b, budget_s=25.0):
    """The reference's mul_into on the same operands, one core (bounded)."""
    import ctypes
    host = host_cpu()   # before the process pins itself to one core below
    ref = os.path.join(ROOT, "oracle", "_ref", "libpreclib_ref.so")
    P64 = ctypes.POINTER(ctypes.c_uint64)
    r = np.zeros(len(a) + len(b), np.uint64)
    if os.path.exists(ref):
        L = ctypes.CDLL(ref)
        f = L.ref_mul_into
        f.restype = ctypes.c_size_t
        f.argtypes = [P64, ctypes.c_size_t, P64, ctypes.c_size_t, P64, ctypes.c_size_t, ctypes.c_int]
        kind = "reference"
        call = lambda: f(r.ctypes.data_as(P64), len(r), a.ctypes.data_as(P64), len(a),
                         b.ctypes.data_as(P64), len(b), 0)
    else:
        from tests.oracle_py import Oracle
        o = Oracle()
        kind = "port"
        call = lambda: o.mul_into(a, b)
    try:
        os.sched_setaffinity(0, {sorted(os.sched_getaffinity(0))[0]})
    except Exception:
        pass
    times = []
    t_start = time.perf_counter()
    while True:
        t0 = time.perf_counter()
        call()
        times.append(time.perf_counter() - t0)
        if time.perf_counter() - t_start + times[-1] > budget_s or len(times) >= 3:
            break
    med = sorted(times)[len(times) // 2]
    out = {"value": (len(a) + len(b)) / med, "unit": "product limbs/s", "cores": 1, "kind": kind,
           "sample": "%d full %dx%d-limb products (median %.2f s), pinned to one host core"
                     % (len(times), len(a), len(b), med)}
    out.update(host)
    return out


def ref_leaf_bytes(an, bn, sqr=False):
    """Byte model of one product under the reference's own decomposition
    (mul_basic.cpp:238-276): NTT leaf 16(Na+Nb) + 48*2*n (base-2^16 n), FFT
    leaf 16(Na+Nb) + 128 n, blocked = sum of its blocks + 16 (block + lo) per
    block accumulate, schoolbook leaf 16(Na+Nb)."""
    lo, hi = min(an, bn), max(an, bn)
    if lo <= 24 or hi <= 192:
        return 16.0 * (an + bn)
    if hi > 2 * lo and not sqr:
        tot, B = 0.0, 2 * lo
        for off in range(0, hi, B):
            pn = min(B, hi - off)
            tot += ref_leaf_bytes(pn, lo) + 16.0 * (pn + lo)
        return tot
    def np2(x):
        n = 1
        while n < x:
            n <<= 1
        return n
    if hi > 6144 and not sqr:
        return 16.0 * (an + bn) + 96.0 * np2(4 * (an + bn))
    dpl = 4 if hi <= 1024 else 8
    return 16.0 * (an + bn) + 128.0 * np2(dpl * (an + bn))


def bench_batch(args, pgpu, workloads):
    """Uneven batch: plain per-product launches vs one planned, graph-replayed
    batch; both device-resident."""
    shapes = workloads.batch_shapes()
    pairs = [workloads.pair(an, bn, "uniform", k) for k, (an, bn) in enumerate(shapes)]
    limbs = sum(an + bn for an, bn in shapes)
    dev = []
    for (a, b) in pairs:
        dev.append((pgpu.DevBuf(len(a) + len(b)), pgpu.DevBuf.of(a), len(a), pgpu.DevBuf.of(b), len(b)))
    # plain: one product at a time through the dispatcher
    for t in dev[:64]:
        pgpu.mul_dev(t[0], t[1], t[2], t[3], t[4], want_len=False)
    pgpu.sync()
    t0 = time.perf_counter()
    for t in dev:
        pgpu.mul_dev(t[0], t[1], t[2], t[3], t[4], want_len=False)
    pgpu.sync()
    plain_s = time.perf_counter() - t0
    # batched: plan once, replay the captured graph
    batch = pgpu.DevBatch(dev)
    for _ in range(args.warmup):
        batch.run()
    pgpu.sync()
    per = []
    for (r, _, an, _, bn) in dev:
        x = r.download(an + bn)
        n = len(x)
        while n and x[n - 1] == 0:
            n -= 1
        per.append(hashlib.sha256(x[:n].astype("<u8").tobytes()).hexdigest())
    check = None
    gold = os.path.join(ROOT, "tests", "golden", "products.json")
    if os.path.exists(gold):
        check = hashlib.sha256("".join(per).encode()).hexdigest() == json.load(open(gold))["batch"]["digest_of_digests"]
    if check is False:
        print("bench.py: batch differs from the golden digests", file=sys.stderr)
        sys.exit(1)
    pgpu.timing(True)
    pgpu.timing_reset()
    pgpu.sync()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        batch.run()
    pgpu.sync()
    dt = (time.perf_counter() - t0) / args.steps
    stages = pgpu.timing_read()
    pgpu.timing(False)
    # untimed-events rate (graph replay; events disable the graph path)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        batch.run()
    pgpu.sync()
    dt_graph = (time.perf_counter() - t0) / args.steps
    model = sum(ref_leaf_bytes(an, bn) for an, bn in shapes)
    best = min(dt, dt_graph)
    line = {
        "metric": "product limbs/s, 4096 uneven products in one submission (batched vs plain)",
        "value": limbs / best, "unit": "product limbs/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": best * 1e3, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "u32 modular (NTT) / f64 (FFT), u64 limbs",
        "data": "synthetic (splitmix64 uniform limbs)",
        "config": {"workload": "4096 products, Nhi log-uniform [193, 65536], 30% unbalanced up to 4:1",
                   "limbs": limbs, "bit_exact_vs_golden": check},
        "plain": {"value": limbs / plain_s, "ms": plain_s * 1e3},
        "batched_over_plain": plain_s / best,
        "graph_replay_ms": dt_graph * 1e3, "eager_with_events_ms": dt * 1e3,
        "product_roofline": {"bytes": model, "achieved_GBps": model / best / 1e9, "frac": model / best / HBM_PEAK},
        "stages_ms_per_step": {k: v[1] / args.steps for k, v in stages.items()},
    }
    if not args.no_cpu:
        import ctypes
        ref = os.path.join(ROOT, "oracle", "_ref", "libpreclib_ref.so")
        if os.path.exists(ref):
            L = ctypes.CDLL(ref)
            f = L.ref_mul_into
            P64 = ctypes.POINTER(ctypes.c_uint64)
            f.restype = ctypes.c_size_t
            f.argtypes = [P64, ctypes.c_size_t, P64, ctypes.c_size_t, P64, ctypes.c_size_t, ctypes.c_int]
            host = host_cpu()   # before pinning to one core
            try:
                os.sched_setaffinity(0, {sorted(os.sched_getaffinity(0))[0]})
            except Exception:
                pass
            done, t0 = 0, time.perf_counter()
            for (a, b) in pairs:
                r = np.zeros(len(a) + len(b), np.uint64)
                f(r.ctypes.data_as(P64), len(r), a.ctypes.data_as(P64), len(a), b.ctypes.data_as(P64), len(b), 0)
                done += len(a) + len(b)
                if time.perf_counter() - t0 > 20.0:
                    break
            el = time.perf_counter() - t0
            line["cpu_baseline"] = {"value": done / el, "unit": "product limbs/s", "cores": 1, "kind": "reference",
                                    "sample": "first %d limbs of the batch in order (%.1f s), one host core" % (done, el)}
            line["cpu_baseline"].update(host)
    print(json.dumps(line), flush=True)
    batch.free()


def bench_pi(args, pgpu):
    """The pi program (pi_chudnovsky DIGITS) end to end on the GPU: binary
    splitting, sqrt, final division, decimal conversion; string checked
    against the reference program's output digest."""
    d = args.digits
    gold = {}
    gp = os.path.join(ROOT, "tests", "golden", "pi.json")
    if os.path.exists(gp):
        gold = {e["digits"]: e for e in json.load(open(gp))["pi"]}
    for _ in range(args.warmup):
        pgpu.pi(d)
    times, phases, s = [], [], ""
    for _ in range(args.steps):
        t0 = time.perf_counter()
        s, ph = pgpu.pi(d, want_phases=True)
        times.append(time.perf_counter() - t0)
        phases.append(ph)
    check = None
    if d in gold:
        check = hashlib.sha256((s + "\n").encode()).hexdigest() == gold[d]["sha256"]
        if not check:
            print("bench.py: pi digits differ from the reference program's output", file=sys.stderr)
            sys.exit(1)
    best = min(times)
    ph = phases[times.index(best)]
    line = {
        "metric": "pi digits/s (Chudnovsky, decimal string, 1 GPU)", "value": d / best, "unit": "digits/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": best * 1e3,
        "higher_is_better": True, "scaling": "strong", "vs_baseline": None,
        "dtype": "u32 modular (NTT) / f64 (FFT), u64 limbs", "data": "deterministic (pi)",
        "config": {"workload": "pi_chudnovsky %d digits" % d, "digits": d, "bit_exact_vs_reference": check},
        "phases_ms": {"binary_split": ph[0], "sqrt_scale": ph[1], "final_div": ph[2], "to_decimal": ph[3]},
        "median_ms": float(np.median(times)) * 1e3,
    }
    if not args.no_cpu:
        import subprocess
        exe = os.path.join(ROOT, "oracle", "_ref", "pi_chudnovsky")
        if os.path.exists(exe):
            sd = min(d, 1000000)
            core = sorted(os.sched_getaffinity(0))[0]
            t0 = time.perf_counter()
            subprocess.run(["taskset", "-c", str(core), exe, str(sd)], check=True, capture_output=True)
            el = time.perf_counter() - t0
            line["cpu_baseline"] = {"value": sd / el, "unit": "digits/s", "cores": 1, "kind": "reference",
                                    "sample": "reference pi program, %d digits, one host core (%.1f s)" % (sd, el)}
            line["cpu_baseline"].update(host_cpu())
    print(json.dumps(line), flush=True)


def bench_serve(args, pgpu, workloads):
    """Poisson request stream through the continuous-batching server: the
    serving stream (workloads.serve_stream: 80% multiplies, sizes log-uniform
    in [256, 2^21]; 20% divisions, divisor >= 32768 limbs, dividend above
    twice the divisor).  Every result of every run is compared with the same
    request run alone through pg_mul / pg_div."""
    import threading
    reqs = workloads.serve_stream(args.requests)
    limbs = [len(a) + len(b) for _, a, b in reqs]
    alone = [pgpu.mul(a, b) if op == 0 else pgpu.divmod_(a, b) for op, a, b in reqs]

    def same(x, y):
        if isinstance(x, tuple):
            return all(np.array_equal(u, v) for u, v in zip(x, y))
        return np.array_equal(x, y)

    def verify(hs, what):
        for i, h in enumerate(hs):
            if not same(h.result(), alone[i]):
                print("bench.py: %s: request %d differs from the same request alone" % (what, i), file=sys.stderr)
                sys.exit(1)

    # saturated throughput: every request offered at once, three runs (one
    # run's estimate has varied by a third between runs on one box); the
    # median run sets the offered rates below
    sat = []
    for k in range(3):
        srv = pgpu.Server()
        warm = [srv.submit(op, a, b) for op, a, b in reqs[:8]]
        srv.drain()
        del warm
        t0 = time.perf_counter()
        hs = [srv.submit(op, a, b) for op, a, b in reqs]
        srv.drain()
        sat.append((time.perf_counter() - t0, srv.stats()))
        srv.stop()
        verify(hs, "saturated")
        del hs
    sat_s, st = sorted(sat, key=lambda x: x[0])[1]
    sat_rate = len(reqs) / sat_s
    levels = []
    for frac in (0.5, 0.8, 1.1):
        rate = frac * sat_rate
        rng = np.random.default_rng(int(frac * 1000))
        gaps = rng.exponential(1.0 / rate, len(reqs))
        srv = pgpu.Server()
        hs = [None] * len(reqs)

        def feed():
            t = time.perf_counter()
            for i, (op, a, b) in enumerate(reqs):
                t += gaps[i]
                dt = t - time.perf_counter()
                if dt > 0:
                    time.sleep(dt)
                hs[i] = srv.submit(op, a, b)
        t0 = time.perf_counter()
        th = threading.Thread(target=feed)
        th.start()
        th.join()
        srv.drain()
        wall = time.perf_counter() - t0
        lat = np.array([h.latency for h in hs]) * 1e3
        verify(hs, "%.1fx load" % frac)
        mul_lat = np.array([lat[i] for i, r in enumerate(reqs) if r[0] == 0])
        div_lat = np.array([lat[i] for i, r in enumerate(reqs) if r[0] == 1])
        levels.append({"offered_frac_of_saturation": frac, "offered_req_s": rate, "achieved_req_s": len(reqs) / wall,
                       "product_limbs_s": sum(limbs) / wall, "p50_ms": float(np.percentile(lat, 50)),
                       "p99_ms": float(np.percentile(lat, 99)),
                       "mul_p50_ms": float(np.percentile(mul_lat, 50)), "mul_p99_ms": float(np.percentile(mul_lat, 99)),
                       "div_p50_ms": float(np.percentile(div_lat, 50)), "div_p99_ms": float(np.percentile(div_lat, 99)),
                       "max_active": srv.stats()["max_active"], "rounds": srv.stats()["rounds"],
                       "all_equal_alone": True})
        srv.stop()
        del hs
    line = {
        "metric": "served product limbs/s (80% mul / 20% Newton division, Poisson arrivals)",
        "value": sum(limbs) / sat_s, "unit": "product limbs/s", "n_gpus": 1, "steps": len(reqs), "warmup": 8,
        "ms_per_step": sat_s / len(reqs) * 1e3, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "u32 modular (NTT) / f64 (FFT), u64 limbs", "data": "synthetic (splitmix64 uniform limbs)",
        "config": {"workload": "serving stream, sizes log-uniform [256, 2^21], divisor >= 32768",
                   "requests": len(reqs), "divisions": sum(1 for r in reqs if r[0] == 1),
                   "host_to_host": True, "checked": "every result of every run equals the request alone"},
        "saturated": {"req_s": sat_rate, "runs_req_s": [len(reqs) / x[0] for x in sat], "rounds": st["rounds"],
                      "products": st["products"], "max_active": st["max_active"]},
        "levels": levels,
    }
    print(json.dumps(line), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None,
                    help="timed steps (default 100 for the single product, 20 otherwise)")
    ap.add_argument("--warmup", type=int, default=30,
                    help="untimed steps before the timed ones (the GPU clocks need ~30 ms to ramp up)")
    ap.add_argument("--logn", type=int, default=22, help="operand size 2^logn limbs")
    ap.add_argument("--dist", default="uniform")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--workload", default="single", choices=["single", "batch", "pi", "serve"],
                    help="single: one 2^22 x 2^22 product (default); batch: the 4096-product uneven batch; "
                         "pi: the pi program at --digits; serve: Poisson request stream")
    ap.add_argument("--digits", type=int, default=10000000, help="pi workload: decimal digits")
    ap.add_argument("--requests", type=int, default=240, help="serve workload: requests per rate level")
    ap.add_argument("--copy-probe", action="store_true",
                    help="first run a 256 MiB device-to-device copy (counter calibration)")
    args = ap.parse_args()
    if args.steps is None:
        args.steps = 100 if args.workload == "single" else 20
    if args.gpus != 1:
        print("bench.py: measures one GPU only; --gpus %d is not supported (a product's "
              "transform columns and rows would be split across GPUs, but no multi-GPU "
              "path is built)" % args.gpus)
        sys.exit(2)

    from preclib_amd import pgpu, workloads
    if args.copy_probe:
        import torch
        x = torch.ones(1 << 26, dtype=torch.int32, device="cuda")
        y = torch.empty_like(x)
        y.copy_(x)
        torch.cuda.synchronize()
        del x, y
    pgpu.init(0)
    if args.workload == "batch":
        return bench_batch(args, pgpu, workloads)
    if args.workload == "pi":
        return bench_pi(args, pgpu)
    if args.workload == "serve":
        return bench_serve(args, pgpu, workloads)
    N = 1 << args.logn
    a, b = workloads.pair(N, N, args.dist)
    n = transform_len(a, b)
    nw = max(1, n // 2)   # base-2^32 digits: half as many points
    da, db = pgpu.DevBuf.of(a), pgpu.DevBuf.of(b)
    r = pgpu.DevBuf(2 * N)
    # correctness of the measured configuration (outside the timed region,
    # before the warm-up, so that the timed steps follow the warm-up directly)
    rl, _ = pgpu.mul_dev(r, da, N, db, N)
    got = r.download(rl)
    check = None
    gold = os.path.join(ROOT, "tests", "golden", "products.json")
    if os.path.exists(gold):
        for e in json.load(open(gold))["workload"]:
            if e["an"] == N and e["bn"] == N and e["dist"] == args.dist:
                check = hashlib.sha256(got.astype("<u8").tobytes()).hexdigest() == e["sha256"]
    if check is False:
        print("bench.py: product differs from the golden digest", file=sys.stderr)
        sys.exit(1)

    # clock ramp: the engine clock rises over the first ~30 ms of load, so the
    # same products keep the GPU busy for RAMP_S before the W warm-up steps
    t_ramp = time.perf_counter()
    while time.perf_counter() - t_ramp < RAMP_S:
        for _ in range(8):
            pgpu.mul_dev(r, da, N, db, N, want_len=False)
        pgpu.sync()
    for _ in range(args.warmup):
        pgpu.mul_dev(r, da, N, db, N, want_len=False)
    # timed region: K products back to back, no per-stage events between the
    # kernels
    pgpu.sync()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        pgpu.mul_dev(r, da, N, db, N, want_len=False)
    pgpu.sync()
    t1 = time.perf_counter()
    # per-kernel launch times (HIP events on the library's stream) from a
    # second pass of the same K products
    pgpu.timing(True)
    pgpu.timing_reset()
    for _ in range(args.steps):
        pgpu.mul_dev(r, da, N, db, N, want_len=False)
    pgpu.sync()
    stages = pgpu.timing_read()
    pgpu.timing(False)

    dt = t1 - t0
    ms = dt / args.steps * 1e3
    value = args.steps * 2 * N / dt
    mb = model_bytes(N, N, n)
    dom = max(stages, key=lambda k: stages[k][1])
    launches, tot_ms, _ = stages[dom]
    avg_s = tot_ms / launches * 1e-3
    achieved = mb[dom] / avg_s
    prod_bytes = 16.0 * 2 * N + 48.0 * NPRIMES * n
    traffic = None
    tpath = os.path.join(ROOT, "profiles", "traffic_2p22.json")
    if args.logn == 22 and os.path.exists(tpath):
        st = json.load(open(tpath))["stages"].get(dom)
        # only a profile of the plan variant this run used (kernel<24, 1, v>)
        v = pgpu.ntt_plans()[24]
        if st and st["kernel"].replace(" ", "").find("<24,1,%d>" % v) >= 0:
            traffic = st["traffic_bytes"]
    # VALU roofline: SQ_INSTS_VALU per dispatch (rocprofv3 --pmc, a count
    # fixed by the plan) over each kernel's measured launch time
    valu = None
    vpath = os.path.join(ROOT, "profiles", "valu_2p22.json")
    if args.logn == 22 and os.path.exists(vpath):
        vst = json.load(open(vpath))["stages"]
        v = pgpu.ntt_plans()[24]
        ok = all(st["kernel"].replace(" ", "").find("<24,1,%d>" % v) >= 0
                 for k, st in vst.items() if k.startswith("ntt_"))
        if ok and dom in vst:
            ins = vst[dom]["valu_wave_instructions"]
            valu = {"bound": "valu", "kernel": dom, "achieved": ins / avg_s / 1e9, "peak": VALU_PEAK / 1e9,
                    "unit": "G wave-instructions/s", "frac": ins / avg_s / VALU_PEAK,
                    "instructions": ins,
                    "per_kernel_frac": {k: vst[k]["valu_wave_instructions"] / (stages[k][1] / stages[k][0] * 1e-3)
                                        / VALU_PEAK for k in stages if k in vst}}
    line = {
        "metric": "product limbs/s at 2^22 limbs (plain GPU NTT product, device-resident)",
        "value": value,
        "unit": "product limbs/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32 modular (NTT), u64 limbs",
        "data": "synthetic (splitmix64 %s limbs, top limb nonzero)" % args.dist,
        "config": {"workload": "single product %dx%d limbs" % (N, N), "an": N, "bn": N, "dist": args.dist,
                   "executed": {"digits": "base 2^32", "primes": 3, "transform_len": "2^%d" % (nw.bit_length() - 1),
                                "plan_variant": pgpu.ntt_plans()[nw.bit_length() - 1]},
                   "byte_model": {"digits": "base 2^16", "primes": NPRIMES,
                                  "transform_len": "2^%d" % (n.bit_length() - 1),
                                  "bytes": 16.0 * 2 * N + 48.0 * NPRIMES * n},
                   "clock_ramp_s": RAMP_S, "bit_exact_vs_golden": check},
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": achieved / 1e9, "peak": HBM_PEAK / 1e9,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK, "traffic": traffic,
                     "avg_launch_ms": avg_s * 1e3},
        "product_roofline": {"bytes": prod_bytes, "achieved_GBps": prod_bytes / (ms * 1e-3) / 1e9,
                             "frac": prod_bytes / (ms * 1e-3) / HBM_PEAK},
        "valu_roofline": valu,
        "stages_ms": {k: v[1] / v[0] for k, v in stages.items()},
    }
    if not args.no_cpu:
        cb = cpu_baseline(a, b)
        line["cpu_baseline"] = cb
        line["gpu_over_cpu"] = value / cb["value"]
    else:
        line["cpu_baseline"] = None
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()

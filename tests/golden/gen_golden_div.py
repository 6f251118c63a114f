"""Writes tests/golden/division.json: SHA-256 digests of exact quotients,
remainders, reciprocals and square roots computed by GMP.

Run in a container that has GMP (libgmp.so); the GPU box does not need it.
Operands come from preclib_amd.workloads (splitmix64, top limb nonzero) and
are imported with mpz_import(z, n, -1, 8, 0, 0, limbs) as bench_gmp.cpp:30-36
does.  Digest = sha256 of the trimmed limbs as little-endian uint64 bytes.

Cases:
  * div: floor(a / b) and a mod b for divisors of >= 32768 limbs with
    dividends above twice the divisor (the division requests a multiply
    server sees), and the 2^21 / 2^20 shape of pi's final division;
  * recip: floor(2^bits / b) (precn_reciprocal_newton);
  * sqrt: floor(sqrt(a)) (precn_sqrt).
"""
import ctypes
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
sys.path.insert(0, HERE)
from preclib_amd import workloads as W  # noqa: E402
from gen_golden import Mpz, digest, load_gmp  # noqa: E402

DIV_CASES = [(65536 + 17, 32768), (100000, 40000), (1 << 18, 50000), (1 << 21, 1 << 20)]
RECIP_CASES = [(32768, 32768 * 64 * 3), (40000, 64 * 100000 + 5)]
SQRT_CASES = [1 << 16, (1 << 20) + 3]

G = load_gmp()
G.__gmpz_import.argtypes = [ctypes.POINTER(Mpz), ctypes.c_size_t, ctypes.c_int, ctypes.c_size_t, ctypes.c_int,
                            ctypes.c_size_t, ctypes.c_void_p]
G.__gmpz_export.restype = ctypes.c_void_p
G.__gmpz_export.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_size_t), ctypes.c_int, ctypes.c_size_t,
                            ctypes.c_int, ctypes.c_size_t, ctypes.POINTER(Mpz)]
G.__gmpz_setbit.argtypes = [ctypes.POINTER(Mpz), ctypes.c_ulong]
G.__gmpz_sizeinbase.restype = ctypes.c_size_t


def z_of(a):
    z = Mpz()
    G.__gmpz_init(ctypes.byref(z))
    a = np.ascontiguousarray(a, dtype=np.uint64)
    G.__gmpz_import(ctypes.byref(z), len(a), -1, 8, 0, 0, a.ctypes.data)
    return z


def new():
    z = Mpz()
    G.__gmpz_init(ctypes.byref(z))
    return z


def limbs(z):
    n = (G.__gmpz_sizeinbase(ctypes.byref(z), 2) + 63) // 64 + 1
    out = np.zeros(n, np.uint64)
    cnt = ctypes.c_size_t(0)
    G.__gmpz_export(out.ctypes.data, ctypes.byref(cnt), -1, 8, 0, 0, ctypes.byref(z))
    return out[: cnt.value]


def main():
    out = {"format": "sha256 of trimmed limbs, little-endian uint64", "div": [], "recip": [], "sqrt": []}
    for k, (an, bn) in enumerate(DIV_CASES):
        a, b = W.pair(an, bn, "uniform", 100 + k)
        za, zb, q, r = z_of(a), z_of(b), new(), new()
        G.__gmpz_tdiv_qr(ctypes.byref(q), ctypes.byref(r), ctypes.byref(za), ctypes.byref(zb))
        hq, nq = digest(limbs(q))
        hr, nr = digest(limbs(r))
        out["div"].append({"an": an, "bn": bn, "k": 100 + k, "q_sha256": hq, "q_len": nq, "r_sha256": hr,
                           "r_len": nr})
        print("div", an, bn, nq, nr, flush=True)
    for k, (bn, bits) in enumerate(RECIP_CASES):
        b = W.operand(bn, W.SEED_B + 200 + k, "uniform")
        zb, p, q = z_of(b), new(), new()
        G.__gmpz_setbit(ctypes.byref(p), bits)
        G.__gmpz_tdiv_q(ctypes.byref(q), ctypes.byref(p), ctypes.byref(zb))
        h, n = digest(limbs(q))
        out["recip"].append({"bn": bn, "bits": bits, "k": 200 + k, "sha256": h, "len": n})
        print("recip", bn, bits, n, flush=True)
    for k, an in enumerate(SQRT_CASES):
        a = W.operand(an, W.SEED_A + 300 + k, "uniform")
        za, s = z_of(a), new()
        G.__gmpz_sqrt(ctypes.byref(s), ctypes.byref(za))
        h, n = digest(limbs(s))
        out["sqrt"].append({"an": an, "k": 300 + k, "sha256": h, "len": n})
        print("sqrt", an, n, flush=True)
    with open(os.path.join(HERE, "division.json"), "w") as f:
        json.dump(out, f, indent=1)
    print("wrote division.json")


if __name__ == "__main__":
    main()

"""Writes tests/golden/products.json: SHA-256 digests of exact products.

Run in a container that has GMP (libgmp.so); the GPU box does not need it.
The products are computed by GMP's mpz_mul with the limb import bench_gmp.cpp
uses (mpz_import(z, n, -1, 8, 0, 0, limbs), bench_gmp.cpp:30-36), so every
digest is independent of both the reference and this project's code.

Digest = sha256 of the trimmed product limbs as little-endian uint64 bytes.

Cases:
  * the workload shapes (16384^2, 2^20^2, 2^22^2, 4096^2) x (uniform, ones,
    sparse), operands from preclib_amd.workloads.pair;
  * the reference's fft_torture shapes x patterns (fft_torture.cpp:58-72);
  * the reference's expect_all_mul_eq grid (test_prec.cpp:429-478, 32-bit LCG
    pattern) and bench_gmp's 64-bit LCG rows (bench_gmp.cpp:104-129);
  * the uneven batch (4096 products): digest of the concatenated per-product
    digests, plus the first 64 per-product digests.
"""
import ctypes
import hashlib
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
from preclib_amd import workloads as W  # noqa: E402

GMP_PATHS = ["/opt/conda/lib/libgmp.so", "libgmp.so.10", "libgmp.so"]


class Mpz(ctypes.Structure):
    _fields_ = [("alloc", ctypes.c_int), ("size", ctypes.c_int), ("d", ctypes.c_void_p)]


def load_gmp():
    for p in GMP_PATHS:
        try:
            return ctypes.CDLL(p)
        except OSError:
            continue
    raise SystemExit("GMP not found")


G = None


def gmp_mul(a, b):
    global G
    if G is None:
        G = load_gmp()
        G.__gmpz_import.argtypes = [ctypes.POINTER(Mpz), ctypes.c_size_t, ctypes.c_int,
                                    ctypes.c_size_t, ctypes.c_int, ctypes.c_size_t, ctypes.c_void_p]
        G.__gmpz_export.restype = ctypes.c_void_p
        G.__gmpz_export.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_size_t), ctypes.c_int,
                                    ctypes.c_size_t, ctypes.c_int, ctypes.c_size_t, ctypes.POINTER(Mpz)]
    x, y, z = Mpz(), Mpz(), Mpz()
    for m in (x, y, z):
        G.__gmpz_init(ctypes.byref(m))
    a = np.ascontiguousarray(a, dtype=np.uint64)
    b = np.ascontiguousarray(b, dtype=np.uint64)
    G.__gmpz_import(ctypes.byref(x), len(a), -1, 8, 0, 0, a.ctypes.data)
    G.__gmpz_import(ctypes.byref(y), len(b), -1, 8, 0, 0, b.ctypes.data)
    G.__gmpz_mul(ctypes.byref(z), ctypes.byref(x), ctypes.byref(y))
    out = np.zeros(len(a) + len(b) + 1, np.uint64)
    cnt = ctypes.c_size_t(0)
    G.__gmpz_export(out.ctypes.data, ctypes.byref(cnt), -1, 8, 0, 0, ctypes.byref(z))
    for m in (x, y, z):
        G.__gmpz_clear(ctypes.byref(m))
    return out[: cnt.value]


def digest(limbs):
    limbs = np.asarray(limbs, dtype=np.uint64)
    n = len(limbs)
    while n and limbs[n - 1] == 0:
        n -= 1
    return hashlib.sha256(limbs[:n].astype("<u8").tobytes()).hexdigest(), n


def torture_operands(an, bn, pat):
    """fft_torture.cpp:22-49 operand construction."""
    name, ea, eb, alternate, parity = pat

    def filled(n, even, odd):
        x = np.array([odd if (alternate and (i & 1)) else even for i in range(n)], dtype=np.uint64)
        if n and x[-1] == 0:
            x[-1] = even if even else 0xFFFFFFFFFFFFFFFF
        return x

    def par(n):
        x = np.array([0xFFFF if bin(i).count("1") & 1 else 0 for i in range(n)], dtype=np.uint64)
        m = n
        while m and x[m - 1] == 0:
            m -= 1
        return x[:m]

    if parity:
        return par(an), par(bn)
    return filled(an, ea, eb), filled(bn, eb, ea)


TORTURE_SHAPES = [(193, 193), (255, 256), (511, 767), (1024, 1024), (1536, 2048), (2048, 2048),
                  (2049, 2049), (3072, 4096), (4096, 4096), (6144, 8192), (8192, 8192)]
TORTURE_PATTERNS = [
    ("all-ones", 0xFFFFFFFFFFFFFFFF, 0xFFFFFFFFFFFFFFFF, False, False),
    ("ffff0000", 0xFFFF0000FFFF0000, 0xFFFF0000FFFF0000, False, False),
    ("0000ffff", 0x0000FFFF0000FFFF, 0x0000FFFF0000FFFF, False, False),
    ("ff00x00ff", 0xFF00FF00FF00FF00, 0x00FF00FF00FF00FF, False, False),
    ("highxlow", 0xFFFFFFFF00000000, 0x00000000FFFFFFFF, False, False),
    ("carry-run", 0xFFFFFFFFFFFFFF00, 0x00FFFFFFFFFFFFFF, False, False),
    ("alternating", 0xFFFFFFFFFFFFFFFF, 0, True, False),
    ("popcount", 0, 0, False, True),
]


def main():
    out = {"format": "sha256 of trimmed product limbs, little-endian uint64",
           "workload": [], "torture": [], "all_mul_eq": [], "bench_gmp": [], "batch": {}}
    for (an, bn) in [(16384, 16384), (4096, 4096), (1 << 20, 1 << 20), (1 << 22, 1 << 22)]:
        for dist in ("uniform", "ones", "sparse"):
            a, b = W.pair(an, bn, dist)
            h, n = digest(gmp_mul(a, b))
            out["workload"].append({"an": an, "bn": bn, "dist": dist, "sha256": h, "len": n})
            print("workload", an, bn, dist, n, h[:16], flush=True)
    for (an, bn) in TORTURE_SHAPES:
        for pat in TORTURE_PATTERNS:
            a, b = torture_operands(an, bn, pat)
            h, n = digest(gmp_mul(a, b))
            out["torture"].append({"an": an, "bn": bn, "pattern": pat[0], "sha256": h, "len": n})
    # test_prec.cpp:445-478 grid (pattern(n, seed) / pattern(m, seed + 1009))
    for i, base in enumerate([32, 128, 512, 2048]):
        for k, bn in enumerate([base, base + base // 4, base + base // 3, base + base // 2,
                                base * 2, base * 3]):
            seed = [3000, 3100, 3150, 3200, 3300, 3400][k] + i
            a = W.reference_pattern32(base, seed)
            b = W.reference_pattern32(bn, seed + 1009)
            h, n = digest(gmp_mul(a, b))
            out["all_mul_eq"].append({"an": base, "bn": bn, "seed": seed, "sha256": h, "len": n})
    # bench_gmp.cpp:104-129 rows, 2^p limbs with seeds 11+p and 19+p
    for p in range(8, 17):
        n_ = 1 << p
        a = W.reference_pattern(n_, 11 + p)
        b = W.reference_pattern(n_, 19 + p)
        h, n = digest(gmp_mul(a, b))
        out["bench_gmp"].append({"p": p, "an": n_, "bn": n_, "sha256": h, "len": n})
    shapes = W.batch_shapes()
    per = []
    for k, (an, bn) in enumerate(shapes):
        a, b = W.pair(an, bn, "uniform", k)
        per.append(digest(gmp_mul(a, b))[0])
    out["batch"] = {"count": len(shapes), "shapes_sha256": hashlib.sha256(json.dumps(shapes).encode()).hexdigest(),
                    "digest_of_digests": hashlib.sha256("".join(per).encode()).hexdigest(),
                    "first": per[:64]}
    with open(os.path.join(HERE, "products.json"), "w") as f:
        json.dump(out, f, indent=1)
    print("wrote products.json")


if __name__ == "__main__":
    main()

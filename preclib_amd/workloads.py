"""Synthetic operands for the product path (host side, numpy).

Limbs come from the splitmix64 stream (Vigna's published generator): limb k of
an operand seeded with ``seed`` is the (k+1)-th output of a splitmix64 state
that starts at ``seed``.  Distributions:

* ``uniform`` -- every limb is a stream output;
* ``ones``    -- every limb is 0xFFFFFFFFFFFFFFFF (worst-case carries);
* ``sparse``  -- limb i is nonzero with probability ~1% (outputs 2i and 2i+1
  decide presence and value).

The top limb is forced nonzero in every distribution, so every shape is exact.
Seeds: operand a uses 0x243F6A8885A308D3 + k, operand b 0x13198A2E03707344 + k
for the k-th product of a workload.
"""
import numpy as np

SEED_A = 0x243F6A8885A308D3
SEED_B = 0x13198A2E03707344
GOLDEN = np.uint64(0x9E3779B97F4A7C15)
DISTS = {"uniform": 0, "ones": 1, "sparse": 2}
SPARSE_CUT = 42949673  # ~0.01 * 2**32


def _mix(z):
    z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
    z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
    return z ^ (z >> np.uint64(31))


def splitmix64(seed, k):
    """Outputs k (array of stream indices, 0-based) of the stream at ``seed``."""
    with np.errstate(over="ignore"):
        z = np.uint64(seed) + (np.asarray(k, dtype=np.uint64) + np.uint64(1)) * GOLDEN
        return _mix(z)


def operand(n, seed, dist="uniform"):
    """n limbs (uint64, little-endian) of the given distribution; top limb != 0."""
    d = DISTS[dist] if isinstance(dist, str) else int(dist)
    if n <= 0:
        return np.zeros(0, np.uint64)
    idx = np.arange(n, dtype=np.uint64)
    if d == 1:
        a = np.full(n, 0xFFFFFFFFFFFFFFFF, dtype=np.uint64)
    elif d == 2:
        u = splitmix64(seed, 2 * idx)
        v = splitmix64(seed, 2 * idx + np.uint64(1))
        v = np.where(v == 0, np.uint64(1), v)
        a = np.where((u >> np.uint64(32)) < np.uint64(SPARSE_CUT), v, np.uint64(0))
    else:
        a = splitmix64(seed, idx)
    if a[-1] == 0:
        v = splitmix64(seed, np.uint64(2 * (n - 1) + 1))
        a[-1] = v if v != 0 else np.uint64(1)
    return a


def pair(an, bn, dist="uniform", k=0):
    """Operands a (an limbs) and b (bn limbs) of the k-th product."""
    return operand(an, SEED_A + k, dist), operand(bn, SEED_B + k, dist)


def batch_shapes(count=4096, lo=193, hi=65536, unbalanced=0.30, seed=0x5EED5EED):
    """Shapes of the uneven batch: Nhi log-uniform in [lo, hi]; a fraction
    ``unbalanced`` of pairs gets Nlo = max(lo, floor(Nhi / r)), r ~ U(2, 4]."""
    rng = np.random.Generator(np.random.PCG64(seed))
    nhi = np.floor(np.exp(rng.uniform(np.log(lo), np.log(hi), count))).astype(np.int64)
    nhi = np.clip(nhi, lo, hi)
    unb = rng.uniform(0.0, 1.0, count) < unbalanced
    r = 4.0 - rng.uniform(0.0, 2.0, count)            # (2, 4]
    nlo = np.maximum(lo, np.floor(nhi / r)).astype(np.int64)
    nlo = np.where(unb, nlo, nhi)
    swap = rng.uniform(0.0, 1.0, count) < 0.5          # which side is long
    an = np.where(swap, nlo, nhi)
    bn = np.where(swap, nhi, nlo)
    return [(int(x), int(y)) for x, y in zip(an, bn)]


SERVE_SEED = 0x5E12E


def serve_shapes(count, seed=SERVE_SEED):
    """The serving stream's requests as (op, an, bn, k): 80% multiplies with
    both sizes log-uniform in [256, 2^21] limbs, 20% divisions with the
    divisor log-uniform in [32768, 2^20] and the dividend log-uniform in
    (2 * divisor, 2^21]; request k's operands are pair(an, bn, "uniform",
    5000 + k)."""
    rng = np.random.default_rng(seed)
    out = []
    lo, hi = np.log(256), np.log(1 << 21)
    for k in range(count):
        if rng.uniform() < 0.8:
            an, bn = (int(np.exp(rng.uniform(lo, hi))) for _ in range(2))
            op = 0
        else:
            bn = int(np.exp(rng.uniform(np.log(32768), np.log(1 << 20))))
            an = int(np.exp(rng.uniform(np.log(2 * bn + 1), hi)))
            op = 1
        out.append((op, an, bn, 5000 + k))
    return out


def serve_stream(count, seed=SERVE_SEED):
    """serve_shapes with the operands: a list of (op, a, b)."""
    return [(op,) + pair(an, bn, "uniform", k) for op, an, bn, k in serve_shapes(count, seed)]


def reference_pattern(n, seed):
    """bench_gmp.cpp / bench_ntt.cpp pattern(): 64-bit LCG, limb = x | 1."""
    out = np.empty(n, np.uint64)
    x = int(seed) & 0xFFFFFFFFFFFFFFFF
    for i in range(n):
        x = (x * 6364136223846793005 + 1442695040888963407) & 0xFFFFFFFFFFFFFFFF
        out[i] = x | 1
    return out


def reference_pattern32(n, seed):
    """test_prec.cpp pattern(): 32-bit LCG, limb = x | 1 (limbs < 2^32)."""
    out = np.empty(n, np.uint64)
    x = int(seed) & 0xFFFFFFFF
    for i in range(n):
        x = (x * 1664525 + 1013904223) & 0xFFFFFFFF
        out[i] = x | 1
    return out

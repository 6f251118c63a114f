"""GPU parity of the FP64 FFT product path (mul_fft semantics) against the
oracle, the GMP digests, the all-ones closed form and the NTT path."""
import numpy as np
import pytest

from preclib_amd import workloads as W
from tests.oracle_py import digest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def pg():
    from preclib_amd import pgpu
    pgpu.init(0)
    return pgpu


def ones_square(n):
    return np.concatenate([np.array([1], np.uint64), np.zeros(n - 1, np.uint64),
                           np.array([0xFFFFFFFFFFFFFFFE], np.uint64),
                           np.full(n - 1, 0xFFFFFFFFFFFFFFFF, np.uint64)])


@pytest.mark.parametrize("shape", [(193, 193), (255, 256), (511, 767), (1000, 1000), (1024, 1024),
                                   (1025, 1025), (2049, 2039), (3072, 4096), (6144, 6144),
                                   (6144, 8192), (8192, 8192)])
@pytest.mark.parametrize("dist", ["uniform", "ones", "sparse"])
def test_fft_vs_oracle(pg, oracle, shape, dist):
    an, bn = shape
    a, b = W.pair(an, bn, dist, k=an + 1)
    print("case", an, bn, dist, flush=True)
    got, err = pg.mul(a, b, path=pg.PATH_FFT, want_round=True)
    want, ref_err = oracle.mul_fft(a, b)
    assert np.array_equal(got, want)
    assert 0.0 <= err < 0.125          # far from the 1/2 that would misround


@pytest.mark.parametrize("dist", ["uniform", "ones", "sparse"])
def test_fft_4096_golden(pg, golden, dist):
    a, b = W.pair(4096, 4096, dist)
    got, err = pg.mul(a, b, path=pg.PATH_FFT, want_round=True)   # preclib's FFT-range shape
    e = next(x for x in golden["workload"] if x["an"] == 4096 and x["dist"] == dist)
    assert digest(got) == e["sha256"]
    assert 0.0 <= err < 0.125
    print("max rounding distance", dist, err)
    # the dispatcher sends the same shape to the NTT on the GPU (no rounding)
    auto, aerr = pg.mul(a, b, want_round=True)
    assert digest(auto) == e["sha256"] and aerr == 0.0


def test_fft_torture_golden(pg, golden):
    """fft_torture.cpp:58-103: 11 shapes x 8 adversarial patterns, exact."""
    from tests.golden.gen_golden import TORTURE_PATTERNS, torture_operands
    pats = {p[0]: p for p in TORTURE_PATTERNS}
    worst = 0.0
    for e in golden["torture"]:
        a, b = torture_operands(e["an"], e["bn"], pats[e["pattern"]])
        if max(len(a), len(b)) <= 192:      # trimmed below the transform threshold
            continue
        got, err = pg.mul(a, b, path=pg.PATH_FFT, want_round=True)
        assert digest(got) == e["sha256"], e
        worst = max(worst, err)
    assert worst < 0.125
    print("worst rounding distance over the torture set", worst)


@pytest.mark.parametrize("n", [1 << 14, 100000, (1 << 19) + 5])
def test_fft_large_squares_closed_form(pg, n):
    """All-ones squares through the FFT up to n = 2^24 complex points (the
    transform preclib's precn_sqr uses for pi's largest squarings)."""
    a = np.full(n, 0xFFFFFFFFFFFFFFFF, np.uint64)
    got, err = pg.mul(a, a, path=pg.PATH_FFT, want_round=True)
    assert np.array_equal(got, ones_square(n))
    assert err < 0.125


@pytest.mark.parametrize("n", [(1 << 19) + 5, 300000])
def test_fft_equals_ntt_large(pg, n):
    """fft_torture.cpp's check (mul_fft == mul_ntt) at sizes where the NTT
    path is independently pinned by GMP digests."""
    a, b = W.pair(n, n - 7, "uniform", k=9)
    f, err = pg.mul(a, b, path=pg.PATH_FFT, want_round=True)
    t = pg.mul(a, b, path=pg.PATH_NTT)
    assert np.array_equal(f, t)
    assert err < 0.125


def test_fft_square_alias(pg, oracle):
    a, _ = W.pair(3000, 1, "uniform", k=4)
    got = pg.mul(a, a, path=pg.PATH_FFT)
    assert np.array_equal(got, oracle.mul_fft(a, a)[0])

"""Pins the pi restatement (oracle/or_pi.py) against the digit strings of the
reference's own pi program (tests/golden/pi.json, whose entries were checked
byte for byte against that program's --file output where the flag says so)
and against closed forms of the series."""
import hashlib
import json
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
import or_pi  # noqa: E402


@pytest.fixture(scope="module")
def golden_pi():
    with open(os.path.join(ROOT, "tests", "golden", "pi.json")) as f:
        return {e["digits"]: e for e in json.load(f)["pi"]}


def test_golden_was_checked_against_the_reference_program(golden_pi):
    for d in (100, 1000, 100000, 1000000, 10000000):
        assert golden_pi[d]["reference_program_match"] is True
    assert golden_pi[10000000]["len"] == 10000002
    assert golden_pi[10000000]["head"].startswith("3.14159265358979323846")


@pytest.mark.parametrize("d", [0, 1, 2, 9, 10, 100, 1000, 12345])
def test_restatement_matches_golden(golden_pi, d):
    s = or_pi.pi_digits(d)
    assert hashlib.sha256((s + "\n").encode()).hexdigest() == golden_pi[d]["sha256"]


def test_split_independence():
    # P, Q, T of a range do not depend on the split point
    def alt(a, b):
        if b - a == 1:
            return or_pi.chud_bs(a, b)
        m = a + 1
        lp, lq, lt = alt(a, m)
        rp, rq, rt = alt(m, b)
        return lp * rp, lq * rq, lt * rq + rt * lp
    for a, b in [(0, 7), (3, 20), (0, 33)]:
        assert alt(a, b) == or_pi.chud_bs(a, b)


def test_series_closed_forms():
    # one term: P = p(0) = 1, Q = 1, T = A; Q(0, n) is the product of k^3 C^3/24
    assert or_pi.chud_bs(0, 1) == (1, 1, or_pi.CHUD_A)
    q = 1
    for k in range(1, 9):
        q *= k ** 3 * or_pi.CHUD_C3_OVER_24
    assert or_pi.chud_bs(0, 9)[1] == q
    # sign of T(a, b) is (-1)^a: the first term dominates
    for a in range(0, 12):
        t = or_pi.chud_bs(a, a + 5)[2]
        assert (t > 0) == (a % 2 == 0)

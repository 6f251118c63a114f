"""CPU restatement of preclib's pi program (pi/pi_chudnovsky.cpp) in Python
integers.  TEST INFRASTRUCTURE ONLY: nothing in the product may import it;
only tests/ use it, as the checker, at small digit counts.
"""
import math
import sys

if hasattr(sys, "set_int_max_str_digits"):   # exact conversion of long results
    sys.set_int_max_str_digits(0)

CHUD_C3_OVER_24 = 10939058860032000   # pi_chudnovsky.cpp:11
CHUD_A = 13591409                     # 12
CHUD_B = 545140134                    # 13


def chud_bs(a, b, need_p=True):
    """(P, Q, T) of terms [a, b) (pi_chudnovsky.cpp:117-148); T is signed."""
    if b - a == 1:
        if a == 0:
            return 1, 1, CHUD_A                                    # 119-121
        p = (6 * a - 5) * (2 * a - 1) * (6 * a - 1)                # 124-126
        q = a * a * a * CHUD_C3_OVER_24                            # 128-131
        t = p * (CHUD_B * a + CHUD_A)                              # 133
        return p, q, (-t if a & 1 else t)                          # 134
    m = a + (b - a) // 2                                           # 137
    lp, lq, lt = chud_bs(a, m, True)
    rp, rq, rt = chud_bs(m, b, need_p)
    return (lp * rp if need_p else None), lq * rq, lt * rq + rt * lp   # 141-147


def pi_digits(digits):
    """The text pi_digits returns (pi_chudnovsky.cpp:162-193)."""
    work = digits + 10
    terms = work // 14 + 2
    _, q, t = chud_bs(0, terms, False)
    s = math.isqrt(10 ** (2 * work) * 10005)                       # 109-113, 173-174
    x = q * 426880 * s // t // 10 ** 10                            # 176-178
    text = str(x) if x else "0"                                    # 150-160
    if digits == 0:
        return text
    if len(text) <= digits:
        text = "0" * (digits + 1 - len(text)) + text               # 190
    return text[:-digits] + "." + text[-digits:]                   # 191

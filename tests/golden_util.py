"""Helpers to read the mpmath golden fixtures and measure errors exactly."""
import json
import os
from fractions import Fraction

import numpy as np

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load(name):
    with open(os.path.join(GOLDEN, name + ".json")) as f:
        return json.load(f)


def fx(h):
    return float.fromhex(h)


def exact_sum(*limbs):
    """Exact rational value of a floating-point expansion."""
    return sum((Fraction(float(v)) for v in limbs), Fraction(0))


def frac(s):
    """mpmath decimal string -> Fraction (exact value of that string)."""
    if s == "0":
        return Fraction(0)
    mant, _, ex = s.partition("e")
    f = Fraction(mant)
    if ex:
        e = int(ex)
        f = f * (Fraction(10) ** e)
    return f


def rel_err(limbs, exact_str):
    ex = frac(exact_str)
    got = exact_sum(*limbs)
    if ex == 0:
        return float(abs(got))
    return float(abs(got - ex) / abs(ex))


def dd_arr(pairs):
    return np.array([[fx(a), fx(b)] for a, b in pairs], dtype=np.float64)

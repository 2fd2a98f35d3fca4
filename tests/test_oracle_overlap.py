"""Pins the oracle's overlap restatements (crates/overlap/src) and the dd overlap entry.

* reference tests: normalised self overlap (integral.rs:367-383), f64 vs f128 primitive
  agreement (integral.rs:385-408), H2 STO-3G symmetric S with S12 ~ 0.66 and all-f64 tier
  (assemble.rs:193-216; single-primitive form), diffuse pairs trigger f128 (screen.rs);
* mpmath closed form at 400 bits (tests/golden/overlap.json) for the dd entry built from
  dd.rs operations: within 8 dd ulps (2^-101 relative) -- a wrong exponent (3/2 vs 3/4),
  a dropped lo in the centre difference or a mis-scaled exp all fail it by many orders.
"""
import math

import numpy as np

import oracle_lib as O
from golden_util import frac, fx, load, rel_err


def test_self_overlap_normalized():  # integral.rs:367-383
    alpha = 1.0
    norm = (2.0 * alpha / math.pi) ** 0.75
    c = np.zeros(3)
    s = norm * norm * O.overlap_primitive_f64(alpha, c, alpha, c)
    assert abs(s - 1.0) < 1e-14


def test_f64_f128_primitive_agree():  # integral.rs:385-408 (per primitive pair)
    for ai, aj in ((10.0, 8.0), (5.0, 3.0), (10.0, 3.0)):
        a = O.overlap_primitive_f64(ai, np.zeros(3), aj, np.array([1.0, 0, 0]))
        b = O.overlap_primitive_f128(ai, np.zeros(3), aj, np.array([1.0, 0, 0]))
        assert abs(a - (b[0] + b[1])) < 1e-14


def test_assemble_h2_like_single_primitive():  # assemble.rs:193-216 shape
    alpha = np.array([0.62391, 0.62391])
    centers = np.array([[0.0, 0, 0], [1.4, 0, 0]])
    coeff = (2 * alpha / math.pi) ** 0.75
    s, counts = O.assemble_overlap_s(centers, alpha, coeff)
    assert s[1] == s[2]
    assert abs(s[0] - 1) < 1e-14 and abs(s[3] - 1) < 1e-14
    assert 0.5 < s[1] < 0.8
    assert counts[1] == 0 and counts[0] == 3


def test_diffuse_triggers_f128_and_far_skips():  # screen.rs:57-101
    alpha = np.array([0.04, 0.04, 5.0])
    centers = np.array([[0.0, 0, 0], [2.0, 0, 0], [40.0, 0, 0]])
    coeff = (2 * alpha / math.pi) ** 0.75
    s, counts = O.assemble_overlap_s(centers, alpha, coeff)
    assert counts[1] > 0 and counts[2] > 0
    assert s[0 * 3 + 2] == 0.0


G = load("overlap")


def test_dd_entry_vs_mpmath():
    worst = 0.0
    for c in G:
        ai, aj = fx(c["ai"]), fx(c["aj"])
        ci = np.array([fx(v) for v in c["ci"]])
        cj = np.array([fx(v) for v in c["cj"]])
        s = O.overlap_s_dd_entry(ai, ci, aj, cj)
        ex = float(frac(c["s"]))
        if ex < 1e-290:
            assert s[0] == 0.0 or abs(s[0] - ex) < 1e-300
            continue
        e = rel_err(s, c["s"])
        # exp's argument reduction adds ~|k| 2^-107 (dd.rs:110-111)
        k = abs(math.log(ex)) / math.log(2)
        assert e <= (16 + k) * 2.0 ** -104, (c, e)
        worst = max(worst, e)
    assert worst > 0


def test_dd_diagonal_is_exactly_one_and_symmetric():
    rng = np.random.default_rng(5)
    n = 17
    centers = rng.uniform(-3, 3, (n, 3))
    alpha = 10 ** rng.uniform(-2, 1, n)
    s = O.overlap_s_dd(centers, alpha).reshape(n, n, 2)
    for i in range(n):
        assert list(s[i, i]) == [1.0, 0.0]
    assert np.array_equal(s, s.transpose(1, 0, 2))


def test_primitive_f128_vs_mpmath_to_f64_level():
    # the reference "f128" primitive uses f64 pi, f64 differences and libm exp
    # (integral.rs:38-53), so it is only ~1e-16 accurate
    for c in G[:60]:
        ai, aj = fx(c["ai"]), fx(c["aj"])
        ci = np.array([fx(v) for v in c["ci"]])
        cj = np.array([fx(v) for v in c["cj"]])
        p = O.overlap_primitive_f128(ai, ci, aj, cj)
        ex = float(frac(c["prim"]))
        if ex > 1e-200:
            assert abs((p[0] + p[1]) - ex) <= 1e-12 * ex

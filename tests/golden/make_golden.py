"""Writes the golden fixtures under tests/golden/ from mpmath at high precision.

These vectors are independent of both the oracle and the GPU library: every expected
value is the exact (or 300-bit) mathematical result of the operation on the stored
binary64 inputs.  The tests check the oracle against them with error bounds tight
enough that a dropped term, a wrong sign or a swapped operand in a restated reference
operation fails.

Run:  python tests/golden/make_golden.py      (takes a few seconds)
"""
import json
import os
import random

import mpmath as mp

HERE = os.path.dirname(os.path.abspath(__file__))
mp.mp.prec = 400


def hx(x):
    return float(x).hex()


def mpd(hi, lo=0.0):
    return mp.mpf(hi) + mp.mpf(lo)


def mp_str(v):
    return mp.nstr(v, 110, min_fixed=-1, max_fixed=-1) if v != 0 else "0"


def rand_dd(rng, scale=1.0):
    hi = rng.gauss(0.0, 1.0) * scale
    lo = hi * rng.uniform(-2.0 ** -53, 2.0 ** -53)
    s = hi + lo
    e = lo - (s - hi)
    return s, e


def rand_mf4(rng, scale=1.0):
    l0 = rng.gauss(0.0, 1.0) * scale
    l1 = l0 * rng.uniform(-2.0 ** -53, 2.0 ** -53)
    l2 = l1 * rng.uniform(-2.0 ** -53, 2.0 ** -53)
    l3 = l2 * rng.uniform(-2.0 ** -53, 2.0 ** -53)
    return [l0, l1, l2, l3]


def dd_ops():
    rng = random.Random(20240601)
    cases = []
    for i in range(400):
        sa = 10.0 ** rng.uniform(-8, 8)
        sb = 10.0 ** rng.uniform(-8, 8) if i % 3 else sa  # some same-scale pairs: cancellation
        a = rand_dd(rng, sa)
        b = rand_dd(rng, sb)
        if i % 7 == 0:  # near-cancellation in add
            b = (-a[0] * (1 + rng.uniform(-1e-12, 1e-12)), rng.uniform(-1, 1) * abs(a[0]) * 2.0 ** -60)
        A, B = mpd(*a), mpd(*b)
        rec = {"a": [hx(a[0]), hx(a[1])], "b": [hx(b[0]), hx(b[1])],
               "add": mp_str(A + B), "mul": mp_str(A * B), "div": mp_str(A / B)}
        if A > 0:
            rec["sqrt"] = mp_str(mp.sqrt(A))
        rec["recip"] = mp_str(1 / A)
        cases.append(rec)
    exps = []
    for i in range(200):
        x = rng.uniform(-700, 700) if i % 2 else rng.uniform(-3, 3)
        a = rand_dd(rng, 1.0)
        a = (x, a[1] * x) if x != 0 else (x, 0.0)
        s = a[0] + a[1]
        a = (s, a[1] - (s - a[0]))
        exps.append({"a": [hx(a[0]), hx(a[1])], "exp": mp_str(mp.exp(mpd(*a)))})
    return {"binary": cases, "exp": exps}


def mf4_ops():
    rng = random.Random(777)
    cases = []
    for i in range(200):
        a = rand_mf4(rng, 10.0 ** rng.uniform(-6, 6))
        b = rand_mf4(rng, 10.0 ** rng.uniform(-6, 6))
        A = sum(mp.mpf(v) for v in a)
        B = sum(mp.mpf(v) for v in b)
        rec = {"a": [hx(v) for v in a], "b": [hx(v) for v in b],
               "add": mp_str(A + B), "mul": mp_str(A * B), "div": mp_str(A / B), "recip": mp_str(1 / A)}
        if A > 0:
            rec["sqrt"] = mp_str(mp.sqrt(A))
        cases.append(rec)
    exps = []
    for i in range(60):
        x = rng.uniform(-600, 600) if i % 2 else rng.uniform(-2, 2)
        a = rand_mf4(rng, 1.0)
        a = [x] + [v * x for v in a[1:]]
        exps.append({"a": [hx(v) for v in a], "exp": mp_str(mp.exp(sum(mp.mpf(v) for v in a)))})
    return {"binary": cases, "exp": exps}


def overlap_entries():
    rng = random.Random(4242)
    out = []
    for i in range(120):
        ai = 10.0 ** rng.uniform(-2, 1)
        aj = 10.0 ** rng.uniform(-2, 1)
        ci = [rng.uniform(-10, 10) for _ in range(3)]
        cj = [c + rng.uniform(-6, 6) for c in ci] if i % 5 else list(ci)
        A, B = mp.mpf(ai), mp.mpf(aj)
        r2 = sum((mp.mpf(x) - mp.mpf(y)) ** 2 for x, y in zip(ci, cj))
        s = (4 * A * B / (A + B) ** 2) ** mp.mpf(0.75) * mp.exp(-(A * B / (A + B)) * r2)
        prim = (mp.pi / (A + B)) ** mp.mpf(1.5) * mp.exp(-(A * B / (A + B)) * r2)
        out.append({"ai": hx(ai), "aj": hx(aj), "ci": [hx(v) for v in ci], "cj": [hx(v) for v in cj],
                    "s": mp_str(s), "prim": mp_str(prim)})
    return out


def small_linalg():
    """Exact results (mpmath) for small dd GEMM, Cholesky and SPD solves."""
    rng = random.Random(99)
    out = {}
    m, n, k = 5, 7, 6
    A = [rand_dd(rng) for _ in range(m * k)]
    B = [rand_dd(rng) for _ in range(k * n)]
    C = [[sum(mpd(*A[i * k + p]) * mpd(*B[p * n + j]) for p in range(k)) for j in range(n)] for i in range(m)]
    out["gemm"] = {"m": m, "n": n, "k": k, "a": [[hx(x), hx(y)] for x, y in A],
                   "b": [[hx(x), hx(y)] for x, y in B], "c": [mp_str(v) for row in C for v in row]}
    # SPD: G G^T + n I with dd entries; exact solve of S x = b in mpmath
    nn = 8
    G = [[mp.mpf(rng.gauss(0, 1)) for _ in range(nn)] for _ in range(nn)]
    S = []
    for i in range(nn):
        for j in range(nn):
            v = sum(G[i][p] * G[j][p] for p in range(nn)) + (nn if i == j else 0)
            hi = float(v)
            lo = float(v - mp.mpf(hi))
            S.append((hi, lo))
    Smp = mp.matrix(nn, nn)
    for i in range(nn):
        for j in range(nn):
            Smp[i, j] = mpd(*S[i * nn + j])
    bvec = [rand_dd(rng) for _ in range(nn)]
    bmp = mp.matrix([mpd(*v) for v in bvec])
    x = mp.lu_solve(Smp, bmp)
    L = mp.cholesky(Smp)
    out["spd"] = {"n": nn, "a": [[hx(x_), hx(y_)] for x_, y_ in S], "b": [[hx(x_), hx(y_)] for x_, y_ in bvec],
                  "x": [mp_str(x[i]) for i in range(nn)],
                  "l": [mp_str(L[i, j]) for i in range(nn) for j in range(nn)]}
    return out


def hilbert_table():
    """README.md:189-194 (demo/src/main.rs:54-102): f128 Hilbert solves with b = A*1."""
    return [{"n": 12, "kappa": 1.1e13, "err": 9.89e-18, "res": 1.18e-31},
            {"n": 13, "kappa": 1.8e14, "err": 1.05e-16, "res": 5.70e-32},
            {"n": 14, "kappa": 2.9e15, "err": 5.10e-14, "res": 5.59e-32},
            {"n": 15, "kappa": 4.7e16, "err": 5.22e-12, "res": 9.14e-32},
            {"n": 18, "kappa": 1.9e20, "err": 5.46e-8, "res": 1.09e-31},
            {"n": 20, "kappa": 4.9e22, "err": 9.52e-5, "res": 1.07e-31}]


def main():
    data = {"dd_ops": dd_ops(), "mf4_ops": mf4_ops(), "overlap": overlap_entries(),
            "small_linalg": small_linalg(), "hilbert_readme": hilbert_table()}
    for name, obj in data.items():
        with open(os.path.join(HERE, name + ".json"), "w") as f:
            json.dump(obj, f, indent=0, separators=(",", ":"))
    print("wrote", ", ".join(data))


if __name__ == "__main__":
    main()

"""Writes tests/golden/adjoint_n2048.json: reference values for the gradients of
log det S and b^T S^-1 b of the N = 2048 overlap matrix (lattice Gaussians, seed 4,
exponents log-uniform in [0.01, 10], 1.4 bohr spacing, +-0.05 jitter; b = dd N(0,1) from
seed 44) at 64 coordinates:

  fd_ld, fd_q   f256 central differences at steps h = 1e-15 and 2h (MultiFloat<4>
                values), each difference F(theta +- h) - F(theta) evaluated exactly
                through a rank-2 update of the f256 factor (determinant lemma, Woodbury)
  ref_ld, ref_q the gradient assembled from the CPU dd operations (S^-1 by solves
                against unit vectors), at the same and 192 further coordinates
  logdet, quad  dd values from the same factor
  kappa         ||S||_2 ||S^-1||_2 by power iteration (dd matvec / dd solves)

Inputs come from the library's deterministic host generators (no device needed).  The
run takes several minutes on 8 cores (the f256 factorisation of S dominates).

Run:  python tests/golden/make_adjoint_golden.py
"""
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

import oracle_lib as O  # noqa: E402
import quadprec_amd as qp  # noqa: E402

N, SEED, ALO, AHI, B_SEED, H = 2048, 4, 0.01, 10.0, 44, 1e-15


def hx(v):
    return [float(x).hex() for x in np.ravel(v)]


def coords(n):
    rng = np.random.default_rng(2048)
    cen = rng.choice(3 * n, 40, replace=False)
    alp = 3 * n + rng.choice(n, 24, replace=False)
    fd = np.sort(np.concatenate([cen, alp]))
    rest = np.setdiff1d(np.arange(4 * n), fd)
    extra = np.sort(rng.choice(rest, 192, replace=False))
    return fd, extra


def kappa(s, l, n, iters=40):
    rng = np.random.default_rng(7)
    v = np.zeros((n, 2))
    v[:, 0] = rng.standard_normal(n)
    w = v.copy()
    lmax = lmin_inv = 0.0
    for _ in range(iters):
        v[:, 0] /= np.linalg.norm(v[:, 0])
        v[:, 1] = 0.0
        y = O.matvec(s, v, n)
        lmax = float(np.dot(y[:, 0], v[:, 0]))
        v = y
        w[:, 0] /= np.linalg.norm(w[:, 0])
        w[:, 1] = 0.0
        z = O.solve_cholesky(l, w, n)
        lmin_inv = float(np.dot(z[:, 0], w[:, 0]))
        w = z
    return lmax * lmin_inv


def main():
    t0 = time.time()
    c, a = qp.gen_overlap_basis(SEED, N, ALO, AHI)
    b = qp.gen_dd_normal(B_SEED, N)
    fd_idx, extra = coords(N)
    print("dd reference gradient ...", flush=True)
    ld, q, gl, gq, _ = O.overlap_adjoint_dd(c, a, b)
    s = O.overlap_s_dd(c, a)
    l = O.cholesky_blocked(s, N, 64)
    kap = kappa(s, l, N)
    print(f"  {time.time() - t0:.0f} s, kappa {kap:.3e}; f256 differences ...", flush=True)
    fl, fq = O.overlap_fd_f256(c, a, b, fd_idx, H)
    print(f"  {time.time() - t0:.0f} s", flush=True)
    ref_idx = np.concatenate([fd_idx, extra])
    out = {
        "n": N, "seed": SEED, "alpha_lo": ALO, "alpha_hi": AHI, "spacing": 1.4, "jitter": 0.05, "b_seed": B_SEED,
        "h": H, "kappa": kap,
        "logdet": hx(ld), "quad": hx(q),
        "fd_coords": fd_idx.tolist(),
        "fd_ld": [hx(fl[i]) for i in range(fd_idx.size)],
        "fd_q": [hx(fq[i]) for i in range(fd_idx.size)],
        "ref_coords": ref_idx.tolist(),
        "ref_ld": [hx(gl[k]) for k in ref_idx],
        "ref_q": [hx(gq[k]) for k in ref_idx],
    }
    with open(os.path.join(HERE, "adjoint_n2048.json"), "w") as f:
        json.dump(out, f, indent=0)
    print(f"wrote adjoint_n2048.json in {time.time() - t0:.0f} s")


if __name__ == "__main__":
    main()

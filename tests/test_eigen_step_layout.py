"""Host-side check of the GPU Jacobi step's indexing (quadprec_amd/csrc/k_eigen.hip,
k_jac_step), emulated in float64 with the same structure: pair layout, the circle
permutation sigma into the next step's layout, every 2 x 2 block (I, J) updated by its own
thread, upper blocks columns-then-rows and lower blocks rows-then-columns.  Checked: every
element of the next layout is written exactly once per step, A stays exactly (bitwise)
symmetric with no mirror writes, each sweep meets every pair once, and the sweeps converge
to numpy's eigenvalues with orthogonal eigenvectors."""
import numpy as np
import pytest


def sigma(pos, m):
    if m == 1 or pos == 0:
        return pos
    i = pos >> 1
    if pos % 2 == 0:
        return 2 * i + 2 if i <= m - 2 else 2 * m - 1
    return 2 if i == 0 else 2 * i - 1


def rotation(A, i):
    p, q = 2 * i, 2 * i + 1
    app, aqq, apq = A[p, p], A[q, q], A[p, q]
    if apq != 0.0:
        tau = (aqq - app) / (2.0 * apq)
        if abs(tau) <= 1e150:
            t = (1.0 if tau >= 0 else -1.0) / (abs(tau) + np.sqrt(1.0 + tau * tau))
            c = 1.0 / np.sqrt(1.0 + t * t)
            return c, t * c, app - t * apq, aqq + t * apq, True
    return 1.0, 0.0, app, aqq, False


def step(A, V):
    N2 = A.shape[0]
    m = N2 // 2
    rots = [rotation(A, i) for i in range(m)]
    B = np.full_like(A, np.nan)
    W = np.full_like(V, np.nan)
    written = np.zeros(A.shape, dtype=int)
    for I in range(m):
        cI, sI, ph, qh, act = rots[I]
        for J in range(m):
            cJ, sJ = rots[J][0], rots[J][1]
            p0, p1, q0, q1 = 2 * I, 2 * I + 1, 2 * J, 2 * J + 1
            x00, x01, x10, x11 = A[p0, q0], A[p0, q1], A[p1, q0], A[p1, q1]
            if I == J:
                if act:
                    x00, x11, x01, x10 = ph, qh, 0.0, 0.0
                y00, y01, y10, y11 = x00, x01, x10, x11
            elif I < J:  # columns by J, then rows by I
                c00, c01 = cJ * x00 - sJ * x01, sJ * x00 + cJ * x01
                c10, c11 = cJ * x10 - sJ * x11, sJ * x10 + cJ * x11
                y00, y10 = cI * c00 - sI * c10, sI * c00 + cI * c10
                y01, y11 = cI * c01 - sI * c11, sI * c01 + cI * c11
            else:  # rows by I, then columns by J
                r00, r10 = cI * x00 - sI * x10, sI * x00 + cI * x10
                r01, r11 = cI * x01 - sI * x11, sI * x01 + cI * x11
                y00, y01 = cJ * r00 - sJ * r01, sJ * r00 + cJ * r01
                y10, y11 = cJ * r10 - sJ * r11, sJ * r10 + cJ * r11
            for (a, b, y) in ((p0, q0, y00), (p0, q1, y01), (p1, q0, y10), (p1, q1, y11)):
                B[sigma(a, m), sigma(b, m)] = y
                written[sigma(a, m), sigma(b, m)] += 1
    for r in range(N2):
        for i in range(m):
            c, s = rots[i][0], rots[i][1]
            v0, v1 = V[r, 2 * i], V[r, 2 * i + 1]
            W[r, sigma(2 * i, m)] = c * v0 - s * v1
            W[r, sigma(2 * i + 1, m)] = s * v0 + c * v1
    return B, W, written


@pytest.mark.parametrize("n", [1, 2, 3, 4, 5, 8, 11, 16, 20])
def test_step_layout_symmetry_and_convergence(n):
    rng = np.random.default_rng(n)
    b = rng.standard_normal((n, n))
    a = (b + b.T) / 2
    N2 = n + (n & 1)
    m = N2 // 2
    A = np.zeros((N2, N2))
    A[:n, :n] = a
    V = np.eye(N2)
    layout = list(range(N2))
    met = set()
    for sweep in range(40):
        if np.abs(A - np.diag(np.diag(A))).max() < 1e-14 * max(1.0, np.abs(a).max()):
            break
        for _ in range(N2 - 1):
            for i in range(m):
                met.add(tuple(sorted((layout[2 * i], layout[2 * i + 1]))))
            A, V, written = step(A, V)
            assert (written == 1).all()
            assert np.array_equal(A, A.T)  # bitwise: no mirror writes needed
            nxt = [0] * N2
            for p in range(N2):
                nxt[sigma(p, m)] = layout[p]
            layout = nxt
        if sweep == 0:
            assert len(met) == N2 * (N2 - 1) // 2  # every pair met in one sweep
    ev = np.empty(N2)
    vec = np.empty((N2, N2))
    for pos, orig in enumerate(layout):
        ev[orig] = A[pos, pos]
        vec[:, orig] = V[:, pos]
    ev, vec = ev[:n], vec[:n, :n]
    assert np.allclose(np.sort(ev), np.linalg.eigvalsh(a), atol=1e-12 * max(1.0, np.abs(a).max()))
    assert np.allclose(vec.T @ vec, np.eye(n), atol=1e-12)
    assert np.allclose(a @ vec, vec * ev, atol=1e-10 * max(1.0, np.abs(a).max()))

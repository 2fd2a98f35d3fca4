"""assemble_overlap_mixed over contracted s/p/d/f shells on the GPU (hp_assemble_overlap_mixed)
against the oracle's restatement of crates/overlap/src (screen.rs, integral.rs,
assemble.rs).  The tier counts must be identical; the matrices agree to the last bits of
libm exp (glibc on the host, the device library on the GPU), bounded per entry by 16 ulps
of its envelope (the sum of the magnitudes of the terms the entry sums: |S_uv| itself
where the contraction does not cancel) and typically exactly equal.  The reference's
assembly tests
(assemble.rs:116-164) run against the GPU path with their bars."""
import math

import numpy as np
import pytest

import oracle_lib as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def qp():
    import quadprec_amd
    quadprec_amd.lib()
    return quadprec_amd


def norm(alpha):
    return (2.0 * alpha / math.pi) ** 0.75


def h2():
    prim = [(3.42525, 0.15433 * norm(3.42525)), (0.62391, 0.53533 * norm(0.62391)), (0.16886, 0.44463 * norm(0.16886))]
    return [((0, 0, 0), 0, prim), ((1.4, 0, 0), 0, prim)]


def test_h2_overlap_symmetric(qp):  # assemble.rs:116-129
    s, _ = qp.assemble_overlap_mixed(h2())
    assert s.shape == (2, 2)
    assert abs(s[0, 1] - s[1, 0]) < 1e-15
    assert abs(s[0, 0] - 1.0) < 0.02 and abs(s[1, 1] - 1.0) < 0.02
    assert 0.5 < s[0, 1] < 0.8


def test_all_tight_uses_f64(qp):  # assemble.rs:131-137
    _, c = qp.assemble_overlap_mixed(h2())
    assert c.f128_pairs == 0 and c.f64_pairs > 0


def test_diffuse_triggers_f128(qp):  # assemble.rs:139-163
    prim = [(50.0, 0.3 * norm(50.0)), (0.04, 0.7 * norm(0.04))]
    _, c = qp.assemble_overlap_mixed([((0, 0, 0), 0, prim), ((2.0, 0, 0), 0, prim)])
    assert c.f128_pairs > 0


def molecule(seed, n_atoms, diffuse):
    """atoms on a jittered line/plane, each with s, p, d (and some f) shells of 1-3 primitives"""
    rng = np.random.default_rng(seed)
    shells = []
    for a in range(n_atoms):
        c = (1.6 * (a % 5) + rng.uniform(-0.2, 0.2), 1.6 * (a // 5) + rng.uniform(-0.2, 0.2), rng.uniform(-0.3, 0.3))
        for l in (0, 0, 1, 2, 3 if a % 3 == 0 else 1):
            npr = int(rng.integers(1, 4))
            lo = -1.6 if diffuse else -0.5
            alphas = np.sort(10 ** rng.uniform(lo, 1.8, npr))[::-1]
            shells.append((c, l, [(float(x), float(rng.uniform(0.1, 1.0) * norm(x))) for x in alphas]))
    return shells


@pytest.mark.parametrize("seed,atoms,diffuse,params", [
    (1, 6, False, O.DEFAULT_SCREEN), (2, 10, True, O.DEFAULT_SCREEN), (3, 14, True, (0.3, 50.0, 1e-10, 6.0)),
    (4, 8, True, (1e300, 0.0, 1e-300, 1e300))])   # the last: every pair in the dd tier
def test_matches_oracle(qp, seed, atoms, diffuse, params):
    shells = molecule(seed, atoms, diffuse)
    s, c = qp.assemble_overlap_mixed(shells, params)
    ref, cref = O.assemble_overlap_mixed(shells, params)
    env = O.assemble_overlap_envelope(shells, params)
    assert tuple(c) == tuple(cref)
    d = np.abs(s - ref)
    bar = 16 * 2.0 ** -53 * env
    exact = float((d == 0).mean())
    print(f"seed {seed}: n_ao {s.shape[0]}, tiers {tuple(c)}, max |diff| / entry bar "
          f"{(d / (bar + 1e-300)).max():.3f}, {exact:.1%} bit-identical", flush=True)
    # blocks of distinct shells are written with their mirror (exactly symmetric); inside a
    # shell's own block (u, v) and (v, u) come from different 1D table entries, as in the
    # reference, so they agree only to rounding
    shell_of = np.repeat(np.arange(len(shells)), [(l + 1) * (l + 2) // 2 for _, l, _ in shells])
    cross = shell_of[:, None] != shell_of[None, :]
    assert np.array_equal(s[cross], s.T[cross])
    assert np.all(np.abs(s - s.T) <= bar)
    assert np.all(d <= bar)
    assert np.array_equal(ref == 0, s == 0)  # skipped pairs and exact zeros in the same places


def test_empty_and_shape_errors(qp):
    s, c = qp.assemble_overlap_mixed([((0, 0, 0), 2, [(1.0, 1.0)])])
    assert s.shape == (6, 6) and tuple(c) == (1, 0, 0)
    with pytest.raises(qp.HyperprecError):
        qp.assemble_overlap_mixed([((0, 0, 0), 4, [(1.0, 1.0)])])  # l = 4 is not a ShellKind

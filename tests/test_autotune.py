"""The split GEMM's accuracy-driven moduli choice (host logic, no GPU).

W(R, k) = floor((floor(log2 prod_{r<R} m_r) - ceil(log2 k) - 4) / 2) is the integer width
of R moduli at inner dimension k; the split error is 2^(1-W) of the row / column norm, so
the fewest moduli meeting `bits` is the least R with W >= bits + 1.  The moduli table is
restated here as data (the library's kModuli) and checked for pairwise coprimality."""
import math
import os

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "quadprec_amd", "libhyperprec_hip.so")


def moduli(count=32):
    return [256, 255, 253, 251, 247, 241, 239, 233, 229, 227, 223, 217, 211, 199, 197, 193, 191, 181, 179,
            173, 167, 163, 157, 151, 149, 139, 137, 131, 127, 113, 109, 79][:count]


def width(R, k):
    lk = math.ceil(math.log2(max(k, 1)))
    return (math.floor(sum(math.log2(m) for m in moduli()[:R])) - lk - 4) // 2


def test_moduli_are_pairwise_coprime():
    ms = moduli()
    assert len(set(ms)) == 32 and max(ms) == 256
    assert all(math.gcd(x, y) == 1 for i, x in enumerate(ms) for y in ms[i + 1:])


@pytest.mark.skipif(not os.path.exists(LIB), reason="library not built")
@pytest.mark.parametrize("k", [1, 2, 64, 300, 1000, 4096, 4097, 65536])
def test_fewest_moduli_for_the_requested_bits(k):
    import quadprec_amd as qp
    for bits in range(10, 116):
        r = qp.split_slices_for_bits(k, bits)
        feasible = [R for R in range(8, 33) if width(R, k) <= 117 and width(R, k) >= bits + 1]
        if not feasible:
            assert r == 0, (k, bits, r)
            continue
        assert r == feasible[0], (k, bits, r, feasible[0])
        assert width(r, k) >= bits + 1 and (r == 8 or width(r - 1, k) < bits + 1)


@pytest.mark.skipif(not os.path.exists(LIB), reason="library not built")
def test_dd_accuracy_at_4096_takes_30_moduli():
    import quadprec_amd as qp
    assert width(30, 4096) == 105 and width(29, 4096) == 102
    assert qp.split_slices_for_bits(4096, 104) == 30   # W = 105 (29 moduli give 102)
    assert qp.split_slices_for_bits(4096, 53) == 16    # W = 54 (15 moduli give 50)


def test_no_coprime_modulus_left_out():
    """The first 31 moduli are a pairwise-coprime set grown greedily from 256 down (every value
    in (109, 256] coprime to all of them is in it); the 32nd is the largest value coprime to
    those 31 that keeps the product of all 32 below 2^240, the width of the CRT's limb and
    digit formats."""
    ms = moduli()
    for v in range(110, 257):
        if v not in ms[:31]:
            assert any(math.gcd(v, m) > 1 for m in ms[:31]), v
    p31 = math.prod(ms[:31])
    fits = [v for v in range(2, 110) if all(math.gcd(v, m) == 1 for m in ms[:31]) and p31 * v < 2 ** 240]
    assert ms[31] == max(fits) and math.prod(ms) < 2 ** 240

"""The C-ABI library loads without a GPU and exports every function that
include/hyperprec_hip.h declares (no compute calls are made here)."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "hyperprec_hip.h")
LIB = os.path.join(ROOT, "quadprec_amd", "libhyperprec_hip.so")


def declared():
    src = open(HEADER).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(hp_[a-z0-9_]+)\s*\(", src)))


def test_header_declares_functions():
    names = declared()
    assert "hp_gemm" in names and "hp_last_error" in names


@pytest.mark.skipif(not os.path.exists(LIB), reason="library not built")
def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(LIB)
    missing = [n for n in declared() if not hasattr(lib, n)]
    assert not missing, missing


@pytest.mark.skipif(not os.path.exists(LIB), reason="library not built")
def test_generators_are_deterministic_and_normalised():
    import math
    import numpy as np
    import quadprec_amd as qp
    a = qp.gen_dd_normal(1, 1000)
    b = qp.gen_dd_normal(1, 1000)
    assert np.array_equal(a, b)
    assert not np.array_equal(a, qp.gen_dd_normal(2, 1000))
    assert abs(a[:, 0].mean()) < 0.1 and abs(a[:, 0].std() - 1) < 0.1
    for hi, lo in a:
        assert abs(lo) <= math.ulp(hi) / 2
    m = qp.gen_mf4_normal(3, 200)
    for l in m:
        for i in range(3):
            assert l[i] == 0 or abs(l[i + 1]) <= math.ulp(l[i]) / 2


@pytest.mark.skipif(not os.path.exists(LIB), reason="library not built")
def test_gemm_rejects_a_wrong_out_buffer_before_any_call():
    """out= is checked on the host (size, dtype, layout) before the library could write past it."""
    import numpy as np
    import quadprec_amd as qp
    a = np.zeros((6, 2))
    for bad in (np.zeros((3, 2)), np.zeros((4, 2), dtype=np.float32), np.zeros((4, 4))[:, ::2], np.zeros((5, 2))):
        with pytest.raises(qp.ShapeError):
            qp.gemm(a, a[:6], 2, 2, 3, out=bad)
    with pytest.raises(qp.ShapeError):
        qp.mf_gemm(np.zeros((6, 4)), np.zeros((6, 4)), 2, 2, 3, out=np.zeros((4, 2)))

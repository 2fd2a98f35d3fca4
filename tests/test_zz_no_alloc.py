"""hp_dev_dot_noalloc (the allocation-free form of hp_dev_dot) through the C ABI: two calls
in a row on device planes, the second on a smaller input, each bit-identical to hp_dev_dot
(and to the oracle's dot); sizes above the scratch's maximum are rejected before any work."""
import ctypes as C

import numpy as np
import pytest

import oracle_lib as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def qp():
    import quadprec_amd
    quadprec_amd.lib()
    return quadprec_amd


def run(qp, fn, a, b):
    import torch
    n = len(a)
    da, db = qp.dev.to_planes(a, 2), qp.dev.to_planes(b, 2)
    out = torch.full((2,), float("nan"), dtype=torch.float64, device="cuda")
    stream = torch.cuda.current_stream().cuda_stream
    rc = fn(C.c_void_p(da[0].data_ptr()), C.c_void_p(da[1].data_ptr()), C.c_void_p(db[0].data_ptr()),
            C.c_void_p(db[1].data_ptr()), n, C.c_void_p(out.data_ptr()), C.c_void_p(out.data_ptr() + 8),
            C.c_void_p(stream))
    torch.cuda.synchronize()
    assert rc == 0, qp.lib().hp_last_error()
    return out.cpu().numpy()


def test_noalloc_dot_matches_hp_dev_dot_on_two_calls(qp):
    lib = qp.lib()
    rng = np.random.default_rng(7)
    for n in ((3 << 20) + 123, 5000):  # the second call smaller than the first
        a = np.zeros((n, 2))
        b = np.zeros((n, 2))
        a[:, 0] = rng.standard_normal(n) * 10 ** rng.uniform(-3, 3, n)
        b[:, 0] = rng.standard_normal(n)
        a[:, 1] = a[:, 0] * 2.0 ** -54 * rng.uniform(-1, 1, n)
        got = run(qp, lib.hp_dev_dot_noalloc, a, b)
        ref = run(qp, lib.hp_dev_dot, a, b)
        print(f"n={n}: noalloc {got}, hp_dev_dot {ref}", flush=True)
        assert np.array_equal(got.view(np.uint64), ref.view(np.uint64))
        assert np.array_equal(got.view(np.uint64), O.dot(a, b).view(np.uint64))


def test_noalloc_dot_rejects_sizes_above_its_scratch(qp):
    lib = qp.lib()
    rc = lib.hp_dev_dot_noalloc(None, None, None, None, (1 << 26) + 1, None, None, None)
    assert rc == qp.HP_EINVAL
    assert b"maximum" in lib.hp_last_error()

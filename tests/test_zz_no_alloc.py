"""mdpd_tea_gaussian_noalloc through the C ABI: the same pair Gaussians as
mdpd_tea_gaussian, bit for bit, from the fixed device array (no allocation),
on two calls in a row (the second smaller, so stale contents of the first
call must not leak into it), and its size limit."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

MAX = 1 << 20                          # MDPD_TEA_SCRATCH_MAX
E_ARG, E_OVERFLOW = 1, 4               # MDPD_E_ARG, MDPD_E_OVERFLOW


def _lib(engine):
    lib = ctypes.CDLL(engine.library_path())
    P, I = ctypes.c_void_p, ctypes.c_int
    for name in ("mdpd_tea_gaussian", "mdpd_tea_gaussian_noalloc"):
        fn = getattr(lib, name)
        fn.restype = I
        fn.argtypes = [I, I, P, P, P]
    return lib


def _call(fn, a, b):
    out = np.full(a.shape[0], np.nan, dtype=np.float32)
    rc = fn(0, a.shape[0], a.ctypes.data_as(ctypes.c_void_p), b.ctypes.data_as(ctypes.c_void_p),
            out.ctypes.data_as(ctypes.c_void_p))
    return rc, out


def test_noalloc_gaussian_equals_the_allocating_one(engine):
    lib = _lib(engine)
    rng = np.random.default_rng(97)
    for n in (MAX, 1000):                       # the full array, then a smaller call
        a = rng.integers(0, 2**32, n, dtype=np.uint64).astype(np.uint32)
        b = rng.integers(0, 2**32, n, dtype=np.uint64).astype(np.uint32)
        rc0, ref = _call(lib.mdpd_tea_gaussian, a, b)
        rc1, got = _call(lib.mdpd_tea_gaussian_noalloc, a, b)
        assert rc0 == 0 and rc1 == 0
        assert np.array_equal(got.view(np.uint32), ref.view(np.uint32)), n
        assert np.all(np.isfinite(got)) and np.abs(got).max() <= 4.0


def test_noalloc_gaussian_limits(engine):
    lib = _lib(engine)
    fn = lib.mdpd_tea_gaussian_noalloc
    a = np.zeros(4, np.uint32)
    out = np.zeros(4, np.float32)
    p = lambda x: x.ctypes.data_as(ctypes.c_void_p)
    assert fn(0, MAX + 1, p(a), p(a), p(out)) == E_OVERFLOW     # rejected before any device work
    assert fn(0, -1, p(a), p(a), p(out)) == E_ARG
    assert fn(0, 0, None, None, None) == 0
    assert fn(0, 4, None, p(a), p(out)) == E_ARG

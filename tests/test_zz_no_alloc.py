"""The allocation-free histogram entry point (bd_gder_noalloc) against
bd_gder through the C ABI: two calls in a row, the second on a smaller
replica with fewer bins, both added onto nonzero counts; sizes beyond the
fixed array are refused."""
import ctypes as C

import numpy as np
import pytest

from bd_aging_amd import abi
from bd_aging_amd.engine import Engine, make_params, random_positions

pytestmark = pytest.mark.gpu


def _u(a):
    return a.ctypes.data_as(C.POINTER(C.c_uint64))


def test_gder_noalloc_equals_gder():
    params = [make_params(1000, 1000, phi=0.5, seed=1), make_params(320, 320, phi=0.4, seed=2)]
    e = Engine(params)
    for r, p in enumerate(params):
        e.set_positions(random_positions(p.n_big + p.n_sml, p.side, 40 + r), r)
    lib = e.lib
    for r, nbins, rng in ((0, 200, 0.5 * params[0].side), (1, 128, 5.0)):
        a = np.full(4 * nbins, 7, dtype=np.uint64)
        b = a.copy()
        abi.check(lib.bd_gder(e.h, r, nbins, rng, _u(a)))
        abi.check(lib.bd_gder_noalloc(e.h, r, nbins, rng, _u(b)))
        assert a.sum() > 7 * 4 * nbins  # pairs were counted
        assert np.array_equal(a, b), (r, nbins)
    big = np.zeros(4 * 513, dtype=np.uint64)
    assert lib.bd_gder_noalloc(e.h, 0, 513, 5.0, _u(big)) == abi.BD_E_ARG
    assert b"512" in lib.bd_last_error()
    e.close()

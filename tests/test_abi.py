"""The C ABI library loads and exports what include/bd_abi.h declares (CPU,
no compute calls that need a GPU)."""
import ctypes as C
import os
import re

import numpy as np
import pytest

import orc
from bd_aging_amd import abi
from bd_aging_amd.engine import make_params, geometry, random_positions, lane_seed

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    names = set()
    for h in os.listdir(os.path.join(ROOT, "include")):
        if h.endswith(".h"):
            txt = open(os.path.join(ROOT, "include", h)).read()
            txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
            names |= set(re.findall(r"\b(bd_\w+)\s*\(", txt))
    return names


def test_library_exports_every_declared_symbol():
    lib = C.CDLL(abi.LIB_PATH)
    missing = [n for n in sorted(declared_symbols()) if not hasattr(lib, n)]
    assert not missing, missing


def test_python_signatures_cover_header():
    assert declared_symbols() == set(abi.SIGNATURES)


def test_geometry_matches_oracle():
    for n, phi, pot in [(2000, 0.5, 0), (32768, 0.58, 0), (32768, 0.58, 1), (500, 0.2, 2),
                        (262144, 0.58, 0), (2000, 0.5, 3), (2000, 0.5, 4), (32768, 0.58, 4)]:
        p = make_params(n // 2, n - n // 2, phi=phi, potential=pot, sigma_pol=4.0, phi_pol=0.3)
        g = geometry(p)
        rc, go = orc.geometry(p)
        assert rc == 0
        for f, _ in abi.BDGeometry._fields_:
            assert getattr(g, f) == getattr(go, f), f


def test_geometry_rejects_bad_params():
    p = make_params(4, 4, side=3.0)
    with pytest.raises(abi.BDError):
        geometry(p)
    p = make_params(100, 100, phi=0.3, ntags_big=40)
    with pytest.raises(abi.BDError):
        geometry(p)
    p = make_params(1000, 1000, phi=0.5, potential=abi.BD_POT_WCA49_50_AO)  # no sigma_pol
    with pytest.raises(abi.BDError):
        geometry(p)
    p = make_params(1000, 1000, phi=0.5, potential=5)
    with pytest.raises(abi.BDError):
        geometry(p)


def test_lane_seed_matches_oracle():
    for s, l in [(0, 0), (1, 5), (12345, 999), (2 ** 63 + 17, 3)]:
        assert lane_seed(s, l) == orc.lib().orc_lane_seed(s, l)


def test_random_positions_stream_and_range():
    x = random_positions(1000, 7.5, 42)
    assert x.min() >= 0.0 and x.max() < 7.5
    st = orc.Stream(orc.lib().orc_lane_seed(42, 0))
    u = st.u64(3000)
    assert np.array_equal(x.reshape(-1), 7.5 * ((u >> np.uint64(11)).astype(np.float64) * 2.0 ** -53))


def test_error_strings():
    lib = abi.load()
    assert lib.bd_strerror(abi.BD_E_CELL_OVERFLOW) == b"cell capacity (ntags) reached"

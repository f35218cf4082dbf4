"""ctypes wrapper of the FP64 CPU oracle (oracle/liborc.so).

Test infrastructure: the oracle is the checker of the HIP engine; nothing in
bd_aging_amd/ imports this module.
"""
import ctypes as C
import os
import subprocess

import numpy as np

from bd_aging_amd.abi import BDParams, BDGeometry, BDObs

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORC_PATH = os.path.join(ROOT, "oracle", "liborc.so")

_D = C.POINTER(C.c_double)
_I = C.POINTER(C.c_int32)
_U = C.POINTER(C.c_uint64)
_P = C.c_void_p


class RN64(C.Structure):
    _fields_ = [("ring", C.c_uint64 * 250), ("lcg", C.c_uint64), ("np", C.c_int32),
                ("nq", C.c_int32)]


_SIG = {
    "orc_rn64_seed": (None, [C.POINTER(RN64), C.c_int64]),
    "orc_rn64_u64": (C.c_uint64, [C.POINTER(RN64)]),
    "orc_rn64_double": (C.c_double, [C.POINTER(RN64)]),
    "orc_lane_seed": (C.c_uint64, [C.c_uint64, C.c_uint64]),
    "orc_log": (C.c_double, [C.c_double]),
    "orc_sincos_turn": (None, [C.c_double, _D, _D]),
    "orc_gauss3": (None, [_U, _D]),
    "orc_geometry": (C.c_int, [C.POINTER(BDParams), C.POINTER(BDGeometry)]),
    "orc_temp_linear": (C.c_double, [C.c_double] * 5),
    "orc_temp_sine": (C.c_double, [C.c_double] * 6),
    "orc_temperature": (C.c_double, [C.POINTER(BDParams), C.c_int64]),
    "orc_sampling_indexes": (C.c_int32, [C.c_float, C.c_float, C.c_float, C.c_int32, _I,
                                         C.c_int32]),
    "orc_potential": (None, [C.POINTER(BDParams), C.c_double, C.c_double, _D, _D]),
    "orc_create": (_P, [C.POINTER(BDParams)]),
    "orc_destroy": (None, [_P]),
    "orc_get_geometry": (C.c_int, [_P, C.POINTER(BDGeometry)]),
    "orc_set_positions": (C.c_int, [_P, _D]),
    "orc_get_positions": (C.c_int, [_P, _D]),
    "orc_get_raw_positions": (C.c_int, [_P, _D]),
    "orc_set_forces": (C.c_int, [_P, _D]),
    "orc_get_forces": (C.c_int, [_P, _D, _D, _D]),
    "orc_mark_origin": (C.c_int, [_P]),
    "orc_get_step": (C.c_int64, [_P]),
    "orc_get_temperature": (C.c_double, [_P]),
    "orc_cell_locate": (C.c_int, [_P]),
    "orc_get_cells": (C.c_int, [_P, C.c_int32, _I, _I]),
    "orc_compute_forces": (C.c_int, [_P]),
    "orc_neighbours": (C.c_int, [_P, C.c_int32, _I, _I]),
    "orc_update": (C.c_int, [_P]),
    "orc_step": (C.c_int, [_P, C.c_int32]),
    "orc_get_rng_lane": (C.c_int, [_P, C.c_int32, _U, _U, _I, _I]),
    "orc_gder": (C.c_int, [_P, C.c_int32, C.c_double, _U]),
    "orc_observables": (C.c_int, [_P, C.c_double, C.POINTER(BDObs)]),
    "orc_forces_brute": (C.c_int, [C.POINTER(BDParams), _D, _D, _D, _D]),
    "orc_gder_normalize": (None, [C.POINTER(BDParams), C.c_int32, C.c_double, _U, _D]),
    "orc_cv3d_accumulate": (C.c_int, [C.c_int32, C.c_double, _D, _D, C.c_int32, C.c_double,
                                      _D, _D]),
    "orc_cv3d_solve": (C.c_int, [C.c_int32, _D, _D, _D]),
    "orc_jacobi_eig": (C.c_int, [C.c_int32, _D, _D, _D]),
}

_lib = None


def build():
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle")], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(ORC_PATH):
            build()
        L = C.CDLL(ORC_PATH)
        for name, (res, args) in _SIG.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def dp(a):
    return a.ctypes.data_as(_D)


def ip(a):
    return a.ctypes.data_as(_I)


def up(a):
    return a.ctypes.data_as(_U)


class Stream:
    """One reference-semantics rannew64 stream."""

    def __init__(self, idum):
        self.s = RN64()
        lib().orc_rn64_seed(C.byref(self.s), int(idum))

    def u64(self, n):
        f = lib().orc_rn64_u64
        return np.array([f(C.byref(self.s)) for _ in range(n)], dtype=np.uint64)

    def doubles(self, n):
        f = lib().orc_rn64_double
        return np.array([f(C.byref(self.s)) for _ in range(n)])


def geometry(p):
    g = BDGeometry()
    rc = lib().orc_geometry(C.byref(p), C.byref(g))
    return rc, g


class System:
    def __init__(self, p):
        self.p = p
        self.h = lib().orc_create(C.byref(p))
        if not self.h:
            raise ValueError("orc_create rejected the parameters")
        self.n = p.n_big + p.n_sml

    def close(self):
        if self.h:
            lib().orc_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def geometry(self):
        g = BDGeometry()
        lib().orc_get_geometry(self.h, C.byref(g))
        return g

    def set_positions(self, xyz):
        a = np.ascontiguousarray(xyz, dtype=np.float64)
        lib().orc_set_positions(self.h, dp(a))

    def positions(self):
        a = np.zeros((self.n, 3))
        lib().orc_get_positions(self.h, dp(a))
        return a

    def raw_positions(self):
        a = np.zeros((self.n, 3))
        lib().orc_get_raw_positions(self.h, dp(a))
        return a

    def set_forces(self, f):
        a = np.ascontiguousarray(f, dtype=np.float64)
        lib().orc_set_forces(self.h, dp(a))

    def forces(self):
        f = np.zeros((self.n, 3))
        v = np.zeros(self.n)
        u = np.zeros(self.n)
        lib().orc_get_forces(self.h, dp(f), dp(v), dp(u))
        return f, v, u

    def mark_origin(self):
        lib().orc_mark_origin(self.h)

    def step_index(self):
        return lib().orc_get_step(self.h)

    def temperature(self):
        return lib().orc_get_temperature(self.h)

    def cell_locate(self):
        return lib().orc_cell_locate(self.h)

    def cells(self, species):
        g = self.geometry()
        n = g.ncell_big if species == 0 else g.ncell_sml
        ntags = self.p.ntags_big if species == 0 else self.p.ntags_sml
        nocup = np.zeros(n ** 3, dtype=np.int32)
        cv = np.zeros(n ** 3 * ntags, dtype=np.int32)
        lib().orc_get_cells(self.h, species, ip(nocup), ip(cv))
        return nocup, cv

    def compute_forces(self):
        return lib().orc_compute_forces(self.h)

    def neighbours(self, maxn=64):
        counts = np.zeros(self.n, dtype=np.int32)
        lst = np.zeros(self.n * maxn, dtype=np.int32)
        rc = lib().orc_neighbours(self.h, maxn, ip(counts), ip(lst))
        assert rc == 0
        return counts, lst.reshape(self.n, maxn)

    def update(self):
        return lib().orc_update(self.h)

    def step(self, n=1):
        return lib().orc_step(self.h, int(n))

    def rng_lane(self, lane):
        ring = np.zeros(250, dtype=np.uint64)
        lcg = C.c_uint64()
        npp = C.c_int32()
        nq = C.c_int32()
        lib().orc_get_rng_lane(self.h, lane, up(ring), C.byref(lcg), C.byref(npp), C.byref(nq))
        return ring, lcg.value, npp.value, nq.value

    def gder_counts(self, nbins, rng):
        c = np.zeros(4 * nbins, dtype=np.uint64)
        assert lib().orc_gder(self.h, nbins, rng, up(c)) == 0
        return c.reshape(4, nbins)

    def observables(self, q):
        o = BDObs()
        lib().orc_observables(self.h, q, C.byref(o))
        return {k: getattr(o, k) for k, _ in BDObs._fields_}


def forces_brute(p, xyz):
    n = p.n_big + p.n_sml
    a = np.ascontiguousarray(xyz, dtype=np.float64)
    f = np.zeros((n, 3))
    v = np.zeros(n)
    u = np.zeros(n)
    rc = lib().orc_forces_brute(C.byref(p), dp(a), dp(f), dp(v), dp(u))
    assert rc == 0
    return f, v, u


def gauss3(d):
    d = np.ascontiguousarray(d, dtype=np.uint64)
    out = np.zeros(3)
    lib().orc_gauss3(up(d), dp(out))
    return out


def sincos_turn(u):
    s = C.c_double()
    c = C.c_double()
    lib().orc_sincos_turn(float(u), C.byref(s), C.byref(c))
    return s.value, c.value


def sampling_indexes(minimal_decade, run_time, dt, per_decade=10):
    out = np.zeros(10000, dtype=np.int32)
    n = lib().orc_sampling_indexes(minimal_decade, run_time, dt, per_decade, ip(out), 10000)
    return out[:n]


def cv3d_accumulate(bins, rng, xyz, f, side):
    n = len(xyz)
    ata = np.zeros((3, bins, bins))
    atb = np.zeros((3, bins))
    a = np.ascontiguousarray(xyz, dtype=np.float64)
    b = np.ascontiguousarray(f, dtype=np.float64)
    assert lib().orc_cv3d_accumulate(bins, rng, dp(a), dp(b), n, side, dp(ata), dp(atb)) == 0
    return ata, atb


def cv3d_solve(ata, atb):
    bins = ata.shape[1]
    out = np.zeros(bins)
    A = np.ascontiguousarray(ata)
    B = np.ascontiguousarray(atb)
    lib().orc_cv3d_solve(bins, dp(A), dp(B), dp(out))
    return out


def gder_normalize(p, nbins, rng, counts):
    c = np.ascontiguousarray(counts, dtype=np.uint64).reshape(-1)
    g = np.zeros(4 * nbins)
    lib().orc_gder_normalize(C.byref(p), nbins, rng, up(c), dp(g))
    return g.reshape(4, nbins)

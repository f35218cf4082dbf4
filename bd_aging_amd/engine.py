"""numpy-facing wrapper of the C ABI (include/bd_abi.h).

Engine mirrors the reference driver's use of its kernels (main.cu:417-586):
``cell_locate`` -> ``compute_forces`` -> ``step`` ...  Positions are N x 3 float64
arrays, big species first (main.cu:387-393).
"""
import ctypes as C
import math

import numpy as np

from . import abi
from .abi import BDParams, BDGeometry, BDObs, check, load

SPECIES_BIG = 0
SPECIES_SML = 1


def box_side(n_big, n_sml, sigma_big, sigma_sml, phi):
    """Edge of the cubic box holding the two species at packing fraction phi."""
    vol = math.pi / 6.0 * (n_big * sigma_big ** 3 + n_sml * sigma_sml ** 3)
    return (vol / phi) ** (1.0 / 3.0)


# The reference's commented-out 49-50 parameters (get_forces.cu:46-48):
# AA = 50 Apow(50/49) with Apow(x) = x^50 (get_forces.cu:7-17), invT = 0.678575.
PHS_AMP = 50.0 * (50.0 / 49.0) ** 50
PHS_INV_T = 0.678575


def make_params(n_big, n_sml, side=None, phi=None, sigma_big=1.4, sigma_sml=1.0, dt=1e-4,
                potential=abi.BD_POT_WCA, ntags_big=8, ntags_sml=8,
                temp_protocol=abi.BD_TEMP_LINEAR, T0=1.0, Tf=1.0, t0=0.0, tf=0.0, period=1.0,
                kappa=0.0, seed=12345, first_step=0, rng_lanes=0, pot_amp=None,
                pot_inv_t=None, sigma_pol=0.0, phi_pol=0.0):
    """bd_params of one system (encabezados.h:40-75).  Give the box edge `side`
    or the packing fraction `phi`.  kappa: stiffness of BD_POT_HARMONIC.
    pot_amp, pot_inv_t: A and 1/T of the 49-50 potentials (default: the
    reference's commented values PHS_AMP, PHS_INV_T); sigma_pol, phi_pol: the
    polymer diameter and volume fraction of BD_POT_WCA49_50_AO."""
    if side is None:
        if phi is None:
            raise ValueError("give side or phi")
        side = box_side(n_big, n_sml, sigma_big, sigma_sml, phi)
    p = BDParams()
    p.n_big, p.n_sml = int(n_big), int(n_sml)
    p.potential = int(potential)
    p.ntags_big, p.ntags_sml = int(ntags_big), int(ntags_sml)
    p.temp_protocol = int(temp_protocol)
    p.rng_lanes = int(rng_lanes)
    p.side = float(side)
    p.sigma_big, p.sigma_sml = float(sigma_big), float(sigma_sml)
    p.dt, p.kappa = float(dt), float(kappa)
    p.T0, p.Tf, p.t0, p.tf, p.period = float(T0), float(Tf), float(t0), float(tf), float(period)
    p.seed = int(seed) & ((1 << 64) - 1)
    p.first_step = int(first_step)
    phs = potential in (abi.BD_POT_WCA49_50, abi.BD_POT_WCA49_50_AO)
    p.pot_amp = float(PHS_AMP if pot_amp is None and phs else (pot_amp or 0.0))
    p.pot_inv_t = float(PHS_INV_T if pot_inv_t is None and phs else (pot_inv_t or 0.0))
    p.sigma_pol, p.phi_pol = float(sigma_pol), float(phi_pol)
    return p


def _dptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _iptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_int32))


def _uptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_uint64))


def geometry(params):
    lib = load()
    g = BDGeometry()
    check(lib.bd_compute_geometry(C.byref(params), C.byref(g)))
    return g


def random_positions(n, side, seed):
    """Uniform positions in [0, side)^3 from a rannew64 lane (host, C ABI)."""
    lib = load()
    out = np.zeros((n, 3), dtype=np.float64)
    check(lib.bd_random_positions(int(n), float(side), int(seed) & ((1 << 64) - 1), _dptr(out)))
    return out


class Engine:
    """One engine of R replicas on the GPU (R = 1: the plain single system)."""

    def __init__(self, params):
        if isinstance(params, BDParams):
            params = [params]
        self.lib = load()
        self.R = len(params)
        arr = (BDParams * self.R)(*params)
        self.params = list(params)
        h = C.c_void_p()
        check(self.lib.bd_create(arr, self.R, C.byref(h)))
        self.h = h
        self.n = [p.n_big + p.n_sml for p in params]

    # -- lifetime ----------------------------------------------------------
    def close(self):
        if self.h:
            self.lib.bd_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # -- state -------------------------------------------------------------
    def geometry(self, r=0):
        g = BDGeometry()
        check(self.lib.bd_get_geometry(self.h, r, C.byref(g)))
        return g

    def set_positions(self, xyz, r=0):
        a = np.ascontiguousarray(xyz, dtype=np.float64).reshape(self.n[r], 3)
        check(self.lib.bd_set_positions(self.h, r, _dptr(a)))

    def positions(self, r=0):
        a = np.zeros((self.n[r], 3))
        check(self.lib.bd_get_positions(self.h, r, _dptr(a)))
        return a

    def raw_positions(self, r=0):
        a = np.zeros((self.n[r], 3))
        check(self.lib.bd_get_raw_positions(self.h, r, _dptr(a)))
        return a

    def forces(self, r=0):
        f = np.zeros((self.n[r], 3))
        v = np.zeros(self.n[r])
        u = np.zeros(self.n[r])
        check(self.lib.bd_get_forces(self.h, r, _dptr(f), _dptr(v), _dptr(u)))
        return f, v, u

    def set_forces(self, f, r=0):
        a = np.ascontiguousarray(f, dtype=np.float64).reshape(self.n[r], 3)
        check(self.lib.bd_set_forces(self.h, r, _dptr(a)))

    def mark_origin(self):
        check(self.lib.bd_mark_origin(self.h))

    def step_index(self):
        v = C.c_int64()
        check(self.lib.bd_get_step(self.h, C.byref(v)))
        return v.value

    # -- hot path ----------------------------------------------------------
    def cell_locate(self):
        check(self.lib.bd_cell_locate(self.h))

    def max_occupancy(self, r=0):
        """(big, small) largest cell occupancy after the last binning."""
        b, s = C.c_int32(), C.c_int32()
        check(self.lib.bd_max_occupancy(self.h, r, C.byref(b), C.byref(s)))
        return b.value, s.value

    def cells(self, species, r=0):
        g = self.geometry(r)
        p = self.params[r]
        n = g.ncell_big if species == SPECIES_BIG else g.ncell_sml
        ntags = p.ntags_big if species == SPECIES_BIG else p.ntags_sml
        nocup = np.zeros(n ** 3, dtype=np.int32)
        cv = np.zeros(n ** 3 * ntags, dtype=np.int32)
        check(self.lib.bd_get_cells(self.h, r, species, _iptr(nocup), _iptr(cv)))
        return nocup, cv

    def compute_forces(self):
        check(self.lib.bd_compute_forces(self.h))

    def neighbours(self, maxn=64, r=0):
        counts = np.zeros(self.n[r], dtype=np.int32)
        lst = np.zeros(self.n[r] * maxn, dtype=np.int32)
        check(self.lib.bd_neighbours(self.h, r, maxn, _iptr(counts), _iptr(lst)))
        return counts, lst.reshape(self.n[r], maxn)

    def update(self):
        check(self.lib.bd_update(self.h))

    def step(self, nsteps=1):
        check(self.lib.bd_step(self.h, int(nsteps)))

    def prepare_step(self, nsteps):
        """Build the graphs step(nsteps) would replay, without stepping."""
        check(self.lib.bd_prepare_step(self.h, int(nsteps)))

    def autotune(self, trial_launches=10, cache_path=None):
        """Pick the fastest force-kernel shape (results are bit-identical)."""
        v = C.c_int32()
        path = cache_path.encode() if cache_path else None
        check(self.lib.bd_autotune(self.h, int(trial_launches), path, C.byref(v)))
        return v.value

    def set_force_order(self, order):
        check(self.lib.bd_set_force_order(self.h, int(order)))

    def force_order(self):
        v = C.c_int32()
        check(self.lib.bd_get_force_order(self.h, C.byref(v)))
        return v.value

    def set_persistent(self, steps):
        """Steps per persistent cooperative launch in step(); 0 switches it off."""
        check(self.lib.bd_set_persistent(self.h, int(steps)))

    def persistent(self):
        v = C.c_int32()
        check(self.lib.bd_get_persistent(self.h, C.byref(v)))
        return v.value

    def set_graph_steps(self, steps):
        """Steps per captured graph in step(); 0 enqueues every kernel eagerly."""
        check(self.lib.bd_set_graph_steps(self.h, int(steps)))

    def graph_steps(self):
        v = C.c_int32()
        check(self.lib.bd_get_graph_steps(self.h, C.byref(v)))
        return v.value

    def set_walk(self, walk):
        """Force walk: abi.BD_WALK_ROWS (0), _SLABS (1), _CHUNK2 (2), _CHUNK4 (3),
        _LIST (4, neighbour lists) or _FLAT (5, a wavefront's candidates
        tested together)."""
        check(self.lib.bd_set_walk(self.h, int(walk)))

    def set_flat_depth(self, depth):
        """Partner records per particle kept by the flat walk (1..32)."""
        check(self.lib.bd_set_flat_depth(self.h, int(depth)))

    def walk(self):
        v = C.c_int32()
        check(self.lib.bd_get_walk(self.h, C.byref(v)))
        return v.value

    def set_list_skin(self, skin):
        """Skin of the neighbour lists (BD_WALK_LIST), default 0.3."""
        check(self.lib.bd_set_list_skin(self.h, float(skin)))

    def list_skin(self):
        v = C.c_double()
        check(self.lib.bd_get_list_skin(self.h, C.byref(v)))
        return v.value

    def set_pass_split(self, split):
        """Wavefronts per particle block (slab walks with deferral): 0 one,
        1 two (each half of the slabs in visiting order), 2 four (the first
        the same-species slabs up to the centre, the others the remaining
        slabs in nearly equal runs)."""
        check(self.lib.bd_set_pass_split(self.h, int(split)))

    def pass_split(self):
        v = C.c_int32()
        check(self.lib.bd_get_pass_split(self.h, C.byref(v)))
        return v.value

    def set_fused_update(self, mode):
        """Update of each step after the first inside the force kernel of the
        step before: 0 never, 1 wherever the shape allows, 2 (default) with a
        pass split only."""
        check(self.lib.bd_set_fused_update(self.h, int(mode)))

    def fused_update(self):
        """Whether the fused update applies with the current settings."""
        v = C.c_int32()
        check(self.lib.bd_get_fused_update(self.h, C.byref(v)))
        return v.value

    def set_defer_depth(self, depth):
        """Deferred pair terms of the row walk: 0, 16 or 32 LDS entries."""
        check(self.lib.bd_set_defer_depth(self.h, int(depth)))

    def defer_depth(self):
        v = C.c_int32()
        check(self.lib.bd_get_defer_depth(self.h, C.byref(v)))
        return v.value

    def set_brick(self, brick):
        check(self.lib.bd_set_brick(self.h, int(brick)))

    def brick(self):
        v = C.c_int32()
        check(self.lib.bd_get_brick(self.h, C.byref(v)))
        return v.value

    def checkpoint_save(self, path):
        """Write the whole engine state (bit-exact resume) to `path`."""
        check(self.lib.bd_checkpoint_save(self.h, str(path).encode()))

    def checkpoint_load(self, path):
        """Restore a state written by checkpoint_save into an engine made with
        the same parameters."""
        check(self.lib.bd_checkpoint_load(self.h, str(path).encode()))

    def kernel_times(self, nsteps):
        """ms spent in (update, cells, forces) over nsteps eager steps."""
        ms = np.zeros(3)
        check(self.lib.bd_kernel_times(self.h, int(nsteps), _dptr(ms)))
        return ms

    def kernel_times_fused(self, nsteps):
        """ms spent in (cells, forces that also move the particles) over
        nsteps eager steps of the fused path (nsteps - 1 fused launches)."""
        ms = np.zeros(2)
        check(self.lib.bd_kernel_times_fused(self.h, int(nsteps), _dptr(ms)))
        return ms

    def rng_lane(self, lane, r=0):
        ring = np.zeros(250, dtype=np.uint64)
        lcg = C.c_uint64()
        draws = C.c_uint64()
        check(self.lib.bd_get_rng_lane(self.h, r, lane, _uptr(ring), C.byref(lcg),
                                       C.byref(draws)))
        return ring, lcg.value, draws.value

    # -- measurements ------------------------------------------------------
    def gder_counts(self, nbins, rng, r=0):
        c = np.zeros(4 * nbins, dtype=np.uint64)
        check(self.lib.bd_gder(self.h, r, int(nbins), float(rng), _uptr(c)))
        return c.reshape(4, nbins)

    def gder_counts_all(self, nbins, rng):
        c = np.zeros(self.R * 4 * nbins, dtype=np.uint64)
        check(self.lib.bd_gder_all(self.h, int(nbins), float(rng), _uptr(c)))
        return c.reshape(self.R, 4, nbins)

    def set_origin(self, slot):
        check(self.lib.bd_set_origin(self.h, int(slot)))

    def fself_origins(self, q, nslots):
        """(R, nslots, 3): F_s(q) big, MSD big, MSD small vs each origin."""
        out = np.zeros(self.R * nslots * 3)
        check(self.lib.bd_fself_origins(self.h, float(q), int(nslots), _dptr(out)))
        return out.reshape(self.R, nslots, 3)

    def observables(self, q, r=0):
        o = BDObs()
        check(self.lib.bd_observables(self.h, r, float(q), C.byref(o)))
        return {k: getattr(o, k) for k, _ in BDObs._fields_}


def rng_stream(seed, lanes, ndraws):
    lib = load()
    out = np.zeros(lanes * ndraws, dtype=np.uint64)
    check(lib.bd_rng_stream(int(seed) & ((1 << 64) - 1), int(lanes), int(ndraws), _uptr(out)))
    return out.reshape(lanes, ndraws)


def lane_seed(seed, lane):
    return load().bd_lane_seed(int(seed) & ((1 << 64) - 1), int(lane))


def stream_bandwidth(nbytes=1 << 30, reps=10):
    """Achievable HBM bandwidth (GB/s, read + write) of a device copy kernel."""
    v = C.c_double()
    check(load().bd_stream_bandwidth(int(nbytes), int(reps), C.byref(v)))
    return v.value


def grid_barrier_cost(workgroups, barriers=2000):
    """Microseconds per grid-wide barrier of a cooperative launch of
    `workgroups` 64-thread workgroups that does nothing else, and the
    co-resident maximum of such workgroups."""
    us, mx = C.c_double(), C.c_int32()
    check(load().bd_grid_barrier_cost(int(workgroups), int(barriers), C.byref(us), C.byref(mx)))
    return us.value, mx.value


def launch_floor(workgroups, threads=64, reps=200):
    """Mean microseconds of an (almost) empty kernel of `workgroups` x
    `threads`, timed as the engine's kernel times are."""
    v = C.c_double()
    check(load().bd_launch_floor(int(workgroups), int(threads), int(reps), C.byref(v)))
    return v.value


class Contraction:
    """Force contraction (Contracting_Vector3D): add samples, solve, read."""

    def __init__(self, bins, rng):
        self.lib = load()
        h = C.c_void_p()
        check(self.lib.bd_cv3d_create(int(bins), float(rng), C.byref(h)))
        self.h = h
        self.bins = int(bins)

    def close(self):
        if self.h:
            self.lib.bd_cv3d_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def add(self, engine, r=0, periodic=True):
        check(self.lib.bd_cv3d_add(self.h, engine.h, int(r), 1 if periodic else 0))

    def add_host(self, xyz, f, side=0.0):
        a = np.ascontiguousarray(xyz, dtype=np.float64)
        b = np.ascontiguousarray(f, dtype=np.float64)
        check(self.lib.bd_cv3d_add_host(self.h, _dptr(a), _dptr(b), len(a), float(side)))

    def solve(self):
        check(self.lib.bd_cv3d_solve(self.h))

    def vector(self):
        out = np.zeros(self.bins)
        check(self.lib.bd_cv3d_get(self.h, _dptr(out)))
        return out

    def normal(self):
        ata = np.zeros((3, self.bins, self.bins))
        atb = np.zeros((3, self.bins))
        check(self.lib.bd_cv3d_get_normal(self.h, _dptr(ata), _dptr(atb)))
        return ata, atb

"""GPU parity: the HIP engine (through the C ABI) against the FP64 oracle on
the same seeded inputs.

Bar: bit-exact cell lists, neighbour lists, rannew64 streams and g(r) counts;
forces within 1e-12 relative (the kernels use the oracle's summation order,
so they are expected to be bit-exact); trajectories within 1e-9 over 100
steps; reductions within 1e-12 relative.
"""
import math

import numpy as np
import pytest

import orc
from bd_aging_amd import abi
from bd_aging_amd.abi import (BD_POT_WCA, BD_POT_WCA_MODIFIED, BD_POT_HARMONIC, BD_POT_WCA49_50,
                              BD_POT_WCA49_50_AO, BD_TEMP_SINE, BDError)
from bd_aging_amd.engine import Engine, make_params, rng_stream, box_side, geometry
from bd_aging_amd.init import relax_overlaps

pytestmark = pytest.mark.gpu

FORCE_RTOL = 1e-12
TRAJ_TOL = 1e-9


@pytest.fixture(scope="module")
def config1_positions():
    """N=2000 (1000 big sigma=1.4 + 1000 small sigma=1.0) at phi=0.50, no overlaps."""
    side = box_side(1000, 1000, 1.4, 1.0, 0.50)
    x = relax_overlaps(1000, 1000, side, seed=2024)
    return side, x


def pair(p, x):
    e = Engine(p)
    e.set_positions(x)
    o = orc.System(p)
    o.set_positions(x)
    return e, o


def assert_forces_close(fe, fo):
    scale = max(np.abs(fo).max(), 1e-300)
    assert np.all(np.abs(fe - fo) <= FORCE_RTOL * np.maximum(np.abs(fo), scale * 1e-3))


# ------------------------------------------------------------------ rng
def test_rng_stream_bit_exact():
    seed, lanes, nd = 987654321, 96, 700
    got = rng_stream(seed, lanes, nd)
    for l in range(0, lanes, 7):
        ref = orc.Stream(orc.lib().orc_lane_seed(seed, l)).u64(nd)
        assert np.array_equal(got[l], ref), l


# ------------------------------------------------------------------ cells
@pytest.mark.parametrize("pot", [BD_POT_WCA, BD_POT_WCA_MODIFIED])
def test_cells_bit_exact(config1_positions, pot):
    side, x = config1_positions
    p = make_params(1000, 1000, side=side, potential=pot)
    e, o = pair(p, x)
    e.cell_locate()
    assert o.cell_locate() == 0
    for sp in (0, 1):
        ne, ce = e.cells(sp)
        no, co = o.cells(sp)
        assert np.array_equal(ne, no)
        assert np.array_equal(ce, co)
    e.close()


def test_max_occupancy_equals_cell_lists(config1_positions):
    side, x = config1_positions
    e = Engine(make_params(1000, 1000, side=side, seed=4))
    e.set_positions(x)
    with pytest.raises(BDError):
        e.max_occupancy()  # no binning yet
    e.step(3)
    assert e.max_occupancy() == (e.cells(0)[0].max(), e.cells(1)[0].max())
    e.close()


def test_cells_edge_positions():
    p = make_params(200, 200, phi=0.1)
    rng = np.random.default_rng(1)
    x = rng.random((400, 3)) * p.side
    g = orc.geometry(p)[1]
    x[0] = [p.side, 0.0, g.cell_side_big]
    x[1] = [p.side, p.side, p.side]
    x[250] = [g.cell_side_sml * 3, 0.0, p.side]
    e, o = pair(p, x)
    e.cell_locate()
    o.cell_locate()
    for sp in (0, 1):
        assert all(np.array_equal(a, b) for a, b in zip(e.cells(sp), o.cells(sp)))
    e.close()


# ------------------------------------------------------------------ forces
# sigma_pol = 4: cutoffs (1.89, 1.56, 1.25) keep the stencils within two
# cells (the slab walk); sigma_pol = 1: cutoffs (3.36, 2.64, 2.0) need the
# general stencil path, and its 4^3 big cells hold about 16 particles each.
AO_FAST = dict(sigma_pol=4.0, phi_pol=0.3)
AO_WIDE = dict(sigma_pol=1.0, phi_pol=0.4, ntags_big=32, ntags_sml=16)


@pytest.mark.parametrize("pot,kw", [(BD_POT_WCA, {}), (BD_POT_WCA_MODIFIED, {}),
                                    (BD_POT_HARMONIC, {}), (BD_POT_WCA49_50, {}),
                                    (BD_POT_WCA49_50_AO, AO_FAST),
                                    (BD_POT_WCA49_50_AO, AO_WIDE)])
def test_forces_and_neighbours(config1_positions, pot, kw):
    side, x = config1_positions
    p = make_params(1000, 1000, side=side, potential=pot, kappa=30.0, **kw)
    xx = x.copy()
    if pot in (BD_POT_HARMONIC, BD_POT_WCA49_50, BD_POT_WCA49_50_AO):
        # move some particles into contact so the short-range forces act
        # (the relaxed configuration keeps pairs beyond 2^(1/6) sigma)
        xx[:50] += 0.05 if pot == BD_POT_HARMONIC else 0.15
        xx %= side
    e, o = pair(p, xx)
    e.cell_locate()
    o.cell_locate()
    e.compute_forces()
    o.compute_forces()
    fe, ve, ue = e.forces()
    fo, vo, uo = o.forces()
    assert np.abs(fo).max() > 0
    assert_forces_close(fe, fo)
    assert_forces_close(ve, vo)
    assert_forces_close(ue, uo)
    exact = np.array_equal(fe, fo) and np.array_equal(ve, vo) and np.array_equal(ue, uo)
    print(f"forces bit-exact: {exact}")
    ce, le = e.neighbours(256)
    co, lo = o.neighbours(256)
    assert np.array_equal(ce, co)
    for m in range(len(x)):
        assert np.array_equal(le[m, :ce[m]], lo[m, :co[m]])
    e.close()


@pytest.mark.parametrize("pot", [BD_POT_WCA, BD_POT_WCA_MODIFIED])
def test_every_brick_shape_bit_identical(config1_positions, pot):
    side, x = config1_positions
    p = make_params(1000, 1000, side=side, potential=pot, seed=4)
    o = orc.System(p)
    o.set_positions(x)
    o.cell_locate()
    o.compute_forces()
    fo, vo, uo = o.forces()
    e = Engine(p)
    e.set_positions(x)
    e.cell_locate()
    tried = []
    for b in (0, 1, 2, 3, 4):
        try:
            e.set_brick(b)
        except abi.BDError as err:  # shape whose staged box exceeds the LDS table
            assert err.code == abi.BD_E_ARG
            continue
        tried.append(b)
        assert e.brick() == b
        walks = (abi.BD_WALK_ROWS, abi.BD_WALK_SLABS, abi.BD_WALK_CHUNK2, abi.BD_WALK_CHUNK4)
        for walk in (walks if b == 0 else (abi.BD_WALK_ROWS,)):
            e.set_walk(walk)
            assert e.walk() == walk
            for depth in ((0, 16, 32) if b == 0 else (16,)):
                e.set_defer_depth(depth)
                assert e.defer_depth() == depth
                for split in ((0, 1, 2) if b == 0 and walk != abi.BD_WALK_ROWS and depth
                              else (0,)):
                    e.set_pass_split(split)
                    assert e.pass_split() == split
                    e.compute_forces()
                    fe, ve, ue = e.forces()
                    assert np.array_equal(fe, fo) and np.array_equal(ve, vo) and \
                        np.array_equal(ue, uo), (b, walk, depth, split)
                e.set_pass_split(0)
        e.set_walk(abi.BD_WALK_ROWS)
    assert {0, 1, 2, 3} <= set(tried)
    for bad in (-1, 8, 64):
        with pytest.raises(BDError):
            e.set_defer_depth(bad)
    for bad in (-1, 6):
        with pytest.raises(BDError):
            e.set_walk(bad)
    for bad in (-1, 3):
        with pytest.raises(BDError):
            e.set_pass_split(bad)
    e.close()


def test_pairs_at_the_cutoff():
    """Pairs within 1e-9 of the cutoff on both sides, directly and through
    the periodic faces, for all three species pairs: every walk keeps
    exactly the oracle's partners (a pair term just beyond the cutoff is
    tiny but not zero, so one pair too many changes the bits)."""
    side = 13.0
    rng = np.random.default_rng(17)
    p = make_params(12, 12, side=side, seed=5)
    e = Engine(p)
    g = e.geometry()
    # grids of at least 2 nrange + 3 cells: the slab walks run (not the generic one)
    assert g.ncell_big >= 2 * max(g.nrange_bb, g.nrange_sb) + 3
    assert g.ncell_sml >= 2 * max(g.nrange_ss, g.nrange_bs) + 3
    rc = {"bb": g.rc_bb, "ss": g.rc_ss, "bs": g.rc_bs}
    x = np.empty((24, 3))
    x[:12] = rng.uniform(0, side, (12, 3))
    x[12:] = rng.uniform(0, side, (12, 3))
    # (i, j, cutoff, signed offset of r^2 from rc^2, axis, across a face)
    plan = [(0, 1, "bb", -5e-10, 0, False), (2, 3, "bb", 5e-10, 1, True),
            (12, 13, "ss", 4e-10, 2, True), (14, 15, "ss", -4e-10, 0, True),
            (4, 16, "bs", 3e-10, 1, False), (5, 17, "bs", -3e-10, 2, True),
            (6, 18, "bs", 0.0, 0, True), (7, 8, "bb", 0.0, 2, False)]
    for i, j, kind, off, axis, face in plan:
        r = np.sqrt(rc[kind] ** 2 + off)
        xi = rng.uniform(3.0, side - 3.0, 3)
        if face:
            xi[axis] = 0.3  # the partner sits across the x = 0 (or y, z) face
        xj = xi.copy()
        xj[axis] = (xi[axis] - r) % side
        x[i], x[j] = xi, xj
    o = orc.System(p)
    o.set_positions(x)
    o.cell_locate()
    o.compute_forces()
    ref = o.forces()
    counts, _ = o.neighbours(64)
    assert counts.sum() > 0
    e.set_positions(x)
    e.cell_locate()
    for walk, depth, split in ((abi.BD_WALK_SLABS, 16, 0), (abi.BD_WALK_ROWS, 0, 0),
                               (abi.BD_WALK_CHUNK4, 16, 2), (abi.BD_WALK_CHUNK2, 32, 1)):
        e.set_walk(walk)
        e.set_defer_depth(depth)
        e.set_pass_split(split)
        e.compute_forces()
        assert all(np.array_equal(a, b) for a, b in zip(e.forces(), ref)), (walk, depth, split)
    e.close()
    o.close()


def test_pass_split_column_overflow(config1_positions):
    """A big particle with more cross-species partners than the deferral
    column holds (40 small particles packed around it): the split kernel
    walks that pass again on the first wavefront, bit-identical."""
    side, x = config1_positions
    x = x.copy()
    rng = np.random.default_rng(5)
    k = 40
    d = rng.normal(size=(k, 3))
    d /= np.linalg.norm(d, axis=1)[:, None]
    x[1000:1000 + k] = (x[0] + d * rng.uniform(0.7, 1.3, k)[:, None]) % side
    p = make_params(1000, 1000, side=side, seed=3, ntags_big=24, ntags_sml=24)
    o = orc.System(p)
    o.set_positions(x)
    assert o.cell_locate() == 0
    o.compute_forces()
    fo, vo, uo = o.forces()
    counts, _ = o.neighbours(128)
    assert counts[0] > 32
    e = Engine(p)
    e.set_positions(x)
    e.cell_locate()
    e.set_walk(abi.BD_WALK_LIST)  # 40 + listed candidates: the particle walks its cells
    e.compute_forces()
    assert all(np.array_equal(a, b) for a, b in zip(e.forces(), (fo, vo, uo)))
    with pytest.raises(BDError):  # 24 slots per cell: beyond the flat walk's 6-bit range lengths
        e.set_walk(abi.BD_WALK_FLAT)
    assert e.walk() == abi.BD_WALK_LIST
    e.autotune(2)  # passes over the flat shapes it cannot run
    assert e.walk() != abi.BD_WALK_FLAT
    e.compute_forces()
    assert all(np.array_equal(a, b) for a, b in zip(e.forces(), (fo, vo, uo)))
    for walk in (abi.BD_WALK_SLABS, abi.BD_WALK_CHUNK2, abi.BD_WALK_CHUNK4):
        for depth in (16, 32):
            for split in (0, 1, 2):
                e.set_walk(walk)
                e.set_defer_depth(depth)
                e.set_pass_split(split)
                e.compute_forces()
                fe, ve, ue = e.forces()
                assert np.array_equal(fe, fo) and np.array_equal(ve, vo) and \
                    np.array_equal(ue, uo), (walk, depth, split)
    e.close()
    o.close()


@pytest.mark.parametrize("walk,depth,split", [(abi.BD_WALK_CHUNK4, 16, 0),
                                              (abi.BD_WALK_CHUNK4, 16, 1),
                                              (abi.BD_WALK_CHUNK4, 16, 2),
                                              (abi.BD_WALK_FLAT, 0, 0),
                                              (abi.BD_WALK_FLAT, 0, 2)])
def test_fused_update_bit_identical(config1_positions, walk, depth, split):
    """The update run inside the force kernel (default) against separate
    update launches: positions, raw positions, forces, rannew64 lanes and cell
    lists all equal after graph replays and eager steps, through a quench."""
    side, x = config1_positions
    p = make_params(1000, 1000, side=side, T0=1.0, Tf=0.3, t0=0.0, tf=0.004, seed=11)
    runs = []
    for fused in (0, 1):
        e = Engine(p)
        e.set_walk(walk)
        e.set_defer_depth(depth)
        e.set_pass_split(split)
        e.set_fused_update(fused)
        assert e.fused_update() == fused
        e.set_positions(x)
        e.step(37)  # unfused: two 16-step graphs, 5 eager steps; fused: the first update,
        #             two graphs of 16 fused pairs, 4 eager pairs, the last forces
        runs.append([e.positions(), e.raw_positions(), *e.forces(), *e.rng_lane(5),
                     *e.rng_lane(1999), *e.cells(0), *e.cells(1)])
        e.close()
    for a, b in zip(*runs):
        assert np.array_equal(np.asarray(a), np.asarray(b))
    e = Engine(p)
    e.set_walk(walk)
    e.set_defer_depth(depth)
    e.set_pass_split(split)
    # default mode 2: the flat walk, or a pass split
    assert e.fused_update() == (1 if split or walk == abi.BD_WALK_FLAT else 0)
    e.close()
    e = Engine(make_params(1000, 1000, side=side, seed=11, rng_lanes=512))
    e.set_fused_update(1)
    assert e.fused_update() == 0  # lanes serve two particles each: separate update
    e.close()


def test_list_walk_bit_identical(config1_positions):
    """Neighbour lists against the slab walk: the first forces equal the
    oracle's, and 400 steps at T=1 (lists rebuilt many times; a 0.05 skin
    rebuilds almost every step) leave positions, raw positions, forces and the
    rannew64 lanes identical.  The other walks' test covers the sort: a
    partner sorted into the wrong place changes the sums' last bits."""
    side, x = config1_positions
    p = make_params(1000, 1000, side=side, T0=1.0, Tf=1.0, seed=21)
    o = orc.System(p)
    o.set_positions(x)
    o.cell_locate()
    o.compute_forces()
    fo = o.forces()
    o.close()
    runs = []
    for walk, skin in ((abi.BD_WALK_SLABS, None), (abi.BD_WALK_LIST, None),
                       (abi.BD_WALK_LIST, 0.05)):
        e = Engine(p)
        e.set_walk(walk)
        assert e.walk() == walk
        if skin is not None:
            e.set_list_skin(skin)
            assert e.list_skin() == skin
        e.set_positions(x)
        e.cell_locate()
        e.compute_forces()
        assert all(np.array_equal(a, b) for a, b in zip(e.forces(), fo)), (walk, skin)
        e.step(400)
        runs.append([e.positions(), e.raw_positions(), *e.forces(), *e.rng_lane(7)])
        e.close()
    for other in runs[1:]:
        for a, b in zip(runs[0], other):
            assert np.array_equal(np.asarray(a), np.asarray(b))
    e = Engine(p)
    for bad in (0.0, -0.1, 2.5):
        with pytest.raises(BDError):
            e.set_list_skin(bad)
    e.close()


def test_list_walk_batch_bit_identical():
    """The list walk on an uneven batch (replicas rebuild at their own steps:
    T from 0.05 to 1.5), switched on after the cells are located and back and
    forth during the run, against the default slab walk: every replica's
    state equal after 300 steps of graph replays and eager steps."""
    specs = [(300, 340, 0.52, 1.5, 1), (700, 650, 0.55, 0.2, 2), (1000, 1000, 0.60, 0.05, 3),
             (520, 480, 0.58, 0.7, 4)]
    params, starts = [], []
    for nb, ns, phi, T, seed in specs:
        side = box_side(nb, ns, 1.4, 1.0, phi)
        starts.append(relax_overlaps(nb, ns, side, seed=seed))
        params.append(make_params(nb, ns, side=side, T0=1.0, Tf=T, t0=0.0, tf=0.0,
                                  first_step=-5, seed=3000 + seed))
    runs = []
    for mode in ("slabs", "list"):
        e = Engine(params)
        for r, x in enumerate(starts):
            e.set_positions(x, r)
        e.cell_locate()
        if mode == "list":
            e.set_walk(abi.BD_WALK_LIST)  # after the cells exist
        e.step(100)
        if mode == "list":
            e.set_walk(abi.BD_WALK_SLABS)
            e.step(37)
            e.set_walk(abi.BD_WALK_LIST)
            e.step(163)
        else:
            e.step(200)
        runs.append([[e.positions(r), e.raw_positions(r), *e.forces(r)] for r in range(len(specs))])
        e.close()
    for r in range(len(specs)):
        for a, b in zip(runs[0][r], runs[1][r]):
            assert np.array_equal(a, b), r


def test_flat_walk_bit_identical(config1_positions):
    """The flat walk (a wavefront's candidates tested together, partners
    recorded per particle) against the oracle: forces at every record depth
    (depth 1 and 4 send most particles through the direct re-walk) and at
    64, 32 and 16 particles per wavefront, then a 100-step trajectory (16
    per wavefront) identical to the oracle's, with the update fused into the
    walk (graphs of fused pairs and eager steps)."""
    side, x = config1_positions
    p = make_params(1000, 1000, side=side, T0=1.0, Tf=1.0, seed=17)
    e, o = pair(p, x)
    assert o.cell_locate() == 0
    o.compute_forces()
    fo = o.forces()
    e.set_walk(abi.BD_WALK_FLAT)
    assert e.walk() == abi.BD_WALK_FLAT
    e.cell_locate()
    for depth, split in ((32, 0), (1, 0), (4, 1), (12, 2), (32, 1), (32, 2)):
        e.set_flat_depth(depth)
        e.set_pass_split(split)  # 64, 32 or 16 particles per wavefront
        e.compute_forces()
        assert all(np.array_equal(a, b) for a, b in zip(e.forces(), fo)), (depth, split)
    with pytest.raises(BDError):
        e.set_flat_depth(0)
    with pytest.raises(BDError):
        e.set_flat_depth(33)
    assert e.fused_update()  # the flat walk moves the particles of the next step itself
    e.step(100)
    assert o.step(100) == 0
    assert np.array_equal(e.positions(), o.positions())
    assert np.array_equal(e.raw_positions(), o.raw_positions())
    assert all(np.array_equal(a, b) for a, b in zip(e.forces(), o.forces()))
    e.close()
    o.close()


def test_flat_walk_batch_bit_identical():
    """The flat walk on an uneven batch whose 64-particle blocks straddle the
    species boundary (big counts not multiples of 64), whose replicas need
    one or two stencil rings, and one of which is too small for the fast
    walks, switched on and off during the run, against the slab walk: every
    replica's state equal after 300 steps.  The same for the four-wavefront
    split of the chunked walk."""
    specs = [(300, 340, 0.52, 1.5, 1), (701, 650, 0.55, 0.2, 2), (1000, 1000, 0.60, 0.05, 3),
             (523, 480, 0.58, 0.7, 4), (100, 100, 0.50, 0.4, 5)]
    params, starts = [], []
    for nb, ns, phi, T, seed in specs:
        side = box_side(nb, ns, 1.4, 1.0, phi)
        starts.append(relax_overlaps(nb, ns, side, seed=seed))
        params.append(make_params(nb, ns, side=side, T0=1.0, Tf=T, t0=0.0, tf=0.0,
                                  first_step=-5, seed=3000 + seed))
    # the last replica's grids (4^3 big, 6^3 small cells) are too narrow for the
    # fast walks: its workgroups take the per-lane walk inside the flat kernel
    assert geometry(params[-1]).ncell_big < 5
    runs = []
    for mode in ("slabs", "flat", "split4"):
        e = Engine(params)
        for r, x in enumerate(starts):
            e.set_positions(x, r)
        if mode == "flat":
            e.set_walk(abi.BD_WALK_FLAT)
        elif mode == "split4":  # four wavefronts per block, the slab shares of walk_quad_part
            e.set_walk(abi.BD_WALK_CHUNK4)
            e.set_defer_depth(16)
            e.set_pass_split(2)
        e.step(100)
        if mode == "flat":
            e.set_walk(abi.BD_WALK_SLABS)
            e.step(37)
            e.set_walk(abi.BD_WALK_FLAT)
            e.step(163)
        elif mode == "split4":
            e.set_pass_split(0)
            e.step(37)
            e.set_pass_split(2)
            e.step(163)
        else:
            e.step(200)
        runs.append([[e.positions(r), e.raw_positions(r), *e.forces(r)] for r in range(len(specs))])
        e.close()
    for other in (1, 2):
        for r in range(len(specs)):
            for a, b in zip(runs[0][r], runs[other][r]):
                assert np.array_equal(a, b), (other, r)


@pytest.fixture(scope="module")
def plain_positions():
    """N=32768 (16384 + 16384) at phi=0.58, no overlaps: the bench workload."""
    side = box_side(16384, 16384, 1.4, 1.0, 0.58)
    return side, relax_overlaps(16384, 16384, side, seed=1)


def test_forces_bit_exact_at_bench_size(plain_positions):
    """Every force-kernel shape on the 23^3 + 32^3-cell grids, against the oracle."""
    side, x = plain_positions
    p = make_params(16384, 16384, side=side, seed=6)
    o = orc.System(p)
    o.set_positions(x)
    o.cell_locate()
    o.compute_forces()
    fo, vo, uo = o.forces()
    e = Engine(p)
    e.set_positions(x)
    e.cell_locate()
    shapes = [(0, 0, 0, 0), (0, 16, 0, 0), (0, 32, 0, 0), (0, 0, 1, 0), (0, 16, 1, 0),
              (0, 32, 1, 0), (0, 0, 2, 0), (0, 16, 2, 0), (0, 32, 2, 0), (0, 0, 3, 0),
              (0, 16, 3, 0), (0, 32, 3, 0), (0, 16, 1, 1), (0, 32, 1, 1), (0, 16, 2, 1),
              (0, 32, 2, 1), (0, 16, 3, 1), (0, 32, 3, 1), (0, 16, 1, 2), (0, 32, 1, 2),
              (0, 16, 2, 2), (0, 32, 2, 2), (0, 16, 3, 2), (0, 32, 3, 2), (0, 0, 4, 0),
              (0, 0, 5, 0), (0, 0, 5, 1), (0, 0, 5, 2), (1, 16, 0, 0), (2, 16, 0, 0)]
    for b, depth, walk, split in shapes:
        e.set_brick(b)
        e.set_defer_depth(depth)
        e.set_walk(walk)
        e.set_pass_split(split)
        e.compute_forces()
        fe, ve, ue = e.forces()
        assert np.array_equal(fe, fo) and np.array_equal(ve, vo) and np.array_equal(ue, uo), \
            (b, depth, walk, split)
    e.set_pass_split(0)
    for sp in (0, 1):
        ne, ce = e.cells(sp)
        no, co = o.cells(sp)
        assert np.array_equal(ne, no) and np.array_equal(ce, co)
    e.close()
    o.close()


# ------------------------------------------------------------------ update
def test_one_update_bit_exact(config1_positions):
    side, x = config1_positions
    p = make_params(1000, 1000, side=side, T0=1.0, Tf=1.0, seed=99)
    e, o = pair(p, x)
    e.cell_locate()
    o.cell_locate()
    e.compute_forces()
    o.compute_forces()
    e.update()
    assert o.update() == 0
    assert np.array_equal(e.positions(), o.positions())
    assert np.array_equal(e.raw_positions(), o.raw_positions())
    for lane in (0, 1, 777, 1999):
        ring_e, lcg_e, draws = e.rng_lane(lane)
        ring_o, lcg_o, npp, nq = o.rng_lane(lane)
        assert draws == 4
        assert np.array_equal(ring_e, ring_o) and lcg_e == lcg_o
    e.close()


@pytest.mark.parametrize("lanes", [0, 600])
def test_trajectory_100_steps(config1_positions, lanes):
    side, x = config1_positions
    p = make_params(1000, 1000, side=side, T0=1.0, Tf=1.0, seed=7, rng_lanes=lanes)
    e, o = pair(p, x)
    e.step(100)
    assert o.cell_locate() == 0
    o.compute_forces()
    assert o.step(100) == 0
    pe, po = e.positions(), o.positions()
    d = np.abs(pe - po)
    d = np.minimum(d, side - d)
    assert d.max() <= TRAJ_TOL
    assert np.abs(e.raw_positions() - o.raw_positions()).max() <= TRAJ_TOL
    print(f"trajectory bit-exact: {np.array_equal(pe, po)}")
    fe, _, _ = e.forces()
    fo, _, _ = o.forces()
    assert np.allclose(fe, fo, rtol=1e-9, atol=1e-9 * np.abs(fo).max())
    assert e.step_index() == o.step_index() == 100
    e.close()


@pytest.mark.parametrize("pot,kw", [(BD_POT_WCA49_50, {}), (BD_POT_WCA49_50_AO, AO_FAST),
                                    (BD_POT_WCA49_50_AO, AO_WIDE)])
def test_49_50_trajectory_100_steps(config1_positions, pot, kw):
    """potentials.cu:9-52 in the step: 100 steps against the oracle, whatever
    the shape settings (they select the same kernels for these potentials).
    dt = 1e-5: the s^50 wall is too stiff for explicit steps of 1e-4 (a 5 %
    overlap then moves a particle by 0.2 sigma in one step)."""
    side, x = config1_positions
    p = make_params(1000, 1000, side=side, T0=1.0, Tf=1.0, seed=17, potential=pot, dt=1e-5, **kw)
    e, o = pair(p, x)
    e.set_walk(abi.BD_WALK_CHUNK4)
    e.set_pass_split(2)
    e.set_fused_update(1)
    e.step(100)
    assert o.cell_locate() == 0
    o.compute_forces()
    assert o.step(100) == 0
    pe, po = e.positions(), o.positions()
    assert np.array_equal(pe, po)
    assert np.array_equal(e.raw_positions(), o.raw_positions())
    fe, ve, ue = e.forces()
    fo, vo, uo = o.forces()
    assert np.abs(fo).max() > 0
    assert np.array_equal(fe, fo) and np.array_equal(ve, vo) and np.array_equal(ue, uo)
    with pytest.raises(BDError):
        e.set_force_order(abi.BD_ORDER_SPLIT4)
    e.close()


def test_49_50_replicas_and_persistent_steps(config1_positions):
    """Two replicas with their own depletion constants in one engine, stepped by
    the persistent kernel and then the graph, each equal to its own oracle run."""
    side, x = config1_positions
    params = [make_params(1000, 1000, side=side, T0=1.0, Tf=0.5, first_step=-5, seed=41,
                          dt=1e-5, potential=BD_POT_WCA49_50_AO, **AO_FAST),
              make_params(1000, 1000, side=side, T0=0.5, Tf=0.5, first_step=-5, seed=42,
                          dt=1e-5, potential=BD_POT_WCA49_50_AO, sigma_pol=3.0, phi_pol=0.2)]
    e = Engine(params)
    for r in range(2):
        e.set_positions(x, r)
    e.set_persistent(16)
    e.step(37)
    e.set_persistent(0)
    e.step(5)
    for r in range(2):
        o = orc.System(params[r])
        o.set_positions(x)
        assert o.cell_locate() == 0
        o.compute_forces()
        assert o.step(42) == 0
        assert np.array_equal(e.positions(r), o.positions())
        fe, ve, ue = e.forces(r)
        fo, vo, uo = o.forces()
        assert np.array_equal(fe, fo) and np.array_equal(ve, vo) and np.array_equal(ue, uo)
    e.close()


def test_quench_and_sine_protocols(config1_positions):
    side, x = config1_positions
    # instant quench at t0 = tf = 0 after a 20-step transient (first_step = -20)
    p = make_params(1000, 1000, side=side, T0=1.0, Tf=0.1, t0=0.0, tf=0.0, first_step=-20,
                    seed=5)
    e, o = pair(p, x)
    e.step(40)
    o.cell_locate()
    o.compute_forces()
    o.step(40)
    assert np.array_equal(e.positions(), o.positions())
    ob_e, ob_o = e.observables(7.0), o.observables(7.0)
    assert ob_e["temperature"] == ob_o["temperature"] == 0.1
    e.close()
    p = make_params(1000, 1000, side=side, T0=0.2, Tf=1.0, t0=0.0, tf=1.0, period=0.003,
                    temp_protocol=BD_TEMP_SINE, seed=6)
    e, o = pair(p, x)
    e.step(30)
    o.cell_locate()
    o.compute_forces()
    o.step(30)
    assert np.array_equal(e.positions(), o.positions())
    e.close()


def test_graph_replay_equals_eager(config1_positions):
    side, x = config1_positions
    p = make_params(1000, 1000, side=side, seed=11)
    a = Engine(p)
    a.set_positions(x)
    a.step(48)  # three replays of the 16-step graph
    b = Engine(p)
    b.set_positions(x)
    for _ in range(48):
        b.step(1)
    assert np.array_equal(a.positions(), b.positions())
    a.close()
    b.close()


def test_graph_modes_bit_identical(config1_positions):
    """Eager launches, 7-step and 16-step graphs give the same trajectory."""
    side, x = config1_positions
    p = make_params(1000, 1000, side=side, seed=12)
    out = []
    for g in (0, 7, 16):
        e = Engine(p)
        assert e.graph_steps() == 16
        e.set_graph_steps(g)
        assert e.graph_steps() == g
        e.set_positions(x)
        e.step(40)
        out.append((e.positions(), e.forces()[0], e.step_index()))
        e.close()
    for pos, f, ni in out[1:]:
        assert np.array_equal(pos, out[0][0])
        assert np.array_equal(f, out[0][1])
        assert ni == out[0][2]
    e = Engine(p)
    with pytest.raises(BDError):
        e.set_graph_steps(-1)
    with pytest.raises(BDError):
        e.set_graph_steps(257)
    e.close()


# ------------------------------------------------------------------ batch
def test_batch_replicas_equal_single_runs():
    specs = [(300, 340, 0.52, 0.5, 1), (700, 650, 0.55, 0.2, 2), (1000, 1000, 0.60, 0.05, 3)]
    params, starts = [], []
    for nb, ns, phi, T, seed in specs:
        side = box_side(nb, ns, 1.4, 1.0, phi)
        starts.append(relax_overlaps(nb, ns, side, seed=seed))
        params.append(make_params(nb, ns, side=side, T0=1.0, Tf=T, t0=0.0, tf=0.0,
                                  first_step=-5, seed=1000 + seed))
    batch = Engine(params)
    for r, x in enumerate(starts):
        batch.set_positions(x, r)
    batch.step(37)
    for r, (p, x) in enumerate(zip(params, starts)):
        single = Engine(p)
        single.set_positions(x)
        single.step(37)
        assert np.array_equal(batch.positions(r), single.positions())
        assert np.array_equal(batch.raw_positions(r), single.raw_positions())
        single.close()
    # and replica 1 against the oracle
    o = orc.System(params[1])
    o.set_positions(starts[1])
    o.cell_locate()
    o.compute_forces()
    o.step(37)
    d = np.abs(batch.positions(1) - o.positions())
    assert np.minimum(d, params[1].side - d).max() <= TRAJ_TOL
    batch.close()


# ------------------------------------------------------------------ measurements
def test_gder_counts_bit_exact(config1_positions):
    side, x = config1_positions
    p = make_params(1000, 1000, side=side)
    e, o = pair(p, x)
    for nbins, rng in [(200, side / 2), (128, 5.0), (400, 15.0)]:
        ce = e.gder_counts(nbins, rng)
        co = o.gder_counts(nbins, rng)
        assert np.array_equal(ce, co), (nbins, rng)
    e.close()


def morton_order(x, side, bits=6):
    """Indices sorting positions along a Z-order curve of 2^bits cells per axis."""
    c = np.minimum((x / side * (1 << bits)).astype(np.int64), (1 << bits) - 1)
    key = np.zeros(len(x), dtype=np.int64)
    for b in range(bits):
        for a in range(3):
            key |= ((c[:, a] >> b) & 1) << (3 * b + a)
    return np.argsort(key, kind="stable")


def test_gder_uniform_image_tiles_bit_exact():
    """N=16384 in Z order (a wavefront's 64 particles and a tile's 256 are
    compact): most wavefront/tile blocks take the per-axis uniform image
    shortcut of k_gder, and the counts still equal the oracle's."""
    nb = ns = 8192
    p = make_params(nb, ns, phi=0.5)
    side = p.side
    rng = np.random.default_rng(77)
    x = rng.random((nb + ns, 3)) * side
    x[:nb] = x[:nb][morton_order(x[:nb], side)]
    x[nb:] = x[nb:][morton_order(x[nb:], side)]
    # how many (wavefront, tile) blocks have one image index per axis, by the
    # kernel's rule (64 consecutive i, 256 consecutive j)
    inv = 1.0 / side
    w = x.reshape(-1, 64, 3)
    t = x.reshape(-1, 256, 3)
    wlo, whi, tlo, thi = w.min(1), w.max(1), t.min(1), t.max(1)
    kmin = np.floor(inv * (wlo[:, None, :] - thi[None, :, :]) + 0.5)
    kmax = np.floor(inv * (whi[:, None, :] - tlo[None, :, :]) + 0.5)
    uniform = kmin == kmax
    assert uniform.mean() > 0.3 and uniform.any(axis=2).mean() > 0.6
    e, o = pair(p, x)
    for nbins, rg in [(128, 5.0), (200, side / 2)]:
        ce = e.gder_counts(nbins, rg)
        co = o.gder_counts(nbins, rg)
        assert np.array_equal(ce, co), (nbins, rg)
    e.close()


def test_observables_close(config1_positions):
    side, x = config1_positions
    p = make_params(1000, 1000, side=side, T0=0.8, Tf=0.8, seed=3)
    e, o = pair(p, x)
    e.step(25)
    o.cell_locate()
    o.compute_forces()
    o.step(25)
    for q in (7.0, 2 * math.pi):
        a, b = e.observables(q), o.observables(q)
        for k in a:
            assert math.isclose(a[k], b[k], rel_tol=1e-12, abs_tol=1e-14), (k, a[k], b[k])
    e.close()


# ------------------------------------------------------------------ failures
def test_cell_overflow_is_reported():
    p = make_params(100, 100, phi=0.1, ntags_big=2)
    rng = np.random.default_rng(2)
    x = rng.random((200, 3)) * p.side
    x[1:4] = x[0] + 0.01  # four big particles in one cell (capacity 2)
    e = Engine(p)
    e.set_positions(x)
    with pytest.raises(abi.BDError) as err:
        e.cell_locate()
    assert err.value.code == abi.BD_E_CELL_OVERFLOW
    e.close()


def test_error_names_first_failing_step_and_clears():
    """A cell overflow in the middle of a graph-replayed run is reported with
    the step index of the update that caused it (the oracle, stepped one
    step at a time, fails at the same step), and the record clears once the
    replica's positions are replaced."""
    p = make_params(200, 200, phi=0.05, ntags_big=3, ntags_sml=3, T0=300.0, Tf=300.0, seed=52)
    x = np.random.default_rng(12).random((400, 3)) * p.side
    o = orc.System(p)
    o.set_positions(x)
    assert o.cell_locate() == 0
    o.compute_forces()
    first = None
    for _ in range(60):
        ni = o.step_index()
        rc = o.step(1)
        if rc:
            first = (ni, rc)
            break
    o.close()
    assert first is not None and first[0] > 16, first  # inside the second graph replay
    e = Engine(p)
    e.set_positions(x)
    with pytest.raises(abi.BDError) as err:
        e.step(60)
    assert err.value.code == first[1]
    assert f"at step {first[0]}:" in str(err.value), str(err.value)
    with pytest.raises(abi.BDError):  # the record stays until the state is replaced
        e.cell_locate()
    e.set_positions(x)
    e.cell_locate()
    e.close()


def test_alternating_call_lengths_keep_both_graphs(config1_positions):
    """Calls of alternating lengths (fused pairs in graphs, whole-step graphs,
    eager remainders) against one eager run: the same trajectory."""
    side, x = config1_positions
    p = make_params(1000, 1000, side=side, T0=1.0, Tf=0.2, t0=0.0, tf=0.0, seed=23)
    runs = []
    for calls in ((17, 16, 40, 3, 16, 33), (125,)):
        e = Engine(p)
        e.set_walk(abi.BD_WALK_CHUNK4)
        e.set_pass_split(2)
        assert e.fused_update() == 1
        e.set_graph_steps(16 if len(calls) > 1 else 0)
        e.set_positions(x)
        for c in calls:
            e.step(c)
        runs.append([e.positions(), e.raw_positions(), *e.forces(), e.step_index()])
        e.close()
    for a, b in zip(*runs):
        assert np.array_equal(np.asarray(a), np.asarray(b))


def test_relaxed_configuration_has_no_overlaps(config1_positions):
    side, x = config1_positions
    d = x[:, None, :] - x[None, :, :]
    d -= side * np.floor(d / side + 0.5)
    r = np.sqrt((d ** 2).sum(-1))
    np.fill_diagonal(r, np.inf)
    big = np.arange(2000) < 1000
    sig = np.where(big[:, None] & big[None, :], 1.4, np.where(~big[:, None] & ~big[None, :], 1.0, 1.2))
    assert (r / sig).min() >= 1.0 - 1.5e-3


# ------------------------------------------------------------------ split order
@pytest.mark.parametrize("order", [abi.BD_ORDER_SPLIT4, abi.BD_ORDER_SPLIT8])
def test_split_order_forces_and_trajectory(config1_positions, order):
    side, x = config1_positions
    p = make_params(1000, 1000, side=side, T0=1.0, Tf=0.3, t0=0.0, tf=0.0, seed=21)
    o = orc.System(p)
    o.set_positions(x)
    o.cell_locate()
    o.compute_forces()
    fo, vo, uo = o.forces()
    e = Engine(p)
    e.set_positions(x)
    e.set_force_order(order)
    assert e.force_order() == order
    e.cell_locate()
    e.compute_forces()
    fe, ve, ue = e.forces()
    assert_forces_close(fe, fo)
    assert_forces_close(ve, vo)
    assert_forces_close(ue, uo)
    e.step(100)
    assert o.step(100) == 0
    d = np.abs(e.positions() - o.positions())
    assert np.minimum(d, side - d).max() <= TRAJ_TOL
    # deterministic: a second engine reproduces the run bit for bit
    e2 = Engine(p)
    e2.set_positions(x)
    e2.set_force_order(order)
    e2.step(100)
    assert np.array_equal(e.positions(), e2.positions())
    e.close()
    e2.close()


def test_autotune_keeps_results_and_caches(config1_positions, tmp_path):
    side, x = config1_positions
    p = make_params(1000, 1000, side=side, seed=2)
    e = Engine(p)
    e.set_positions(x)
    e.cell_locate()
    e.compute_forces()
    f0 = e.forces()
    cache = str(tmp_path / "tune.txt")
    b = e.autotune(5, cache)
    assert e.brick() == b
    e.compute_forces()
    f1 = e.forces()
    assert all(np.array_equal(a, c) for a, c in zip(f0, f1))
    last = open(cache).read().split()
    assert last[-4:] == [str(b), str(e.defer_depth()), str(e.walk()), str(e.pass_split())]
    e2 = Engine(p)
    e2.set_positions(x)
    assert e2.autotune(5, cache) == b  # served from the cache
    e.close()
    e2.close()


def test_autotune_under_49_50_times_its_one_kernel(config1_positions, tmp_path):
    """The 49-50 potentials run one force kernel whatever the shape: the
    autotuner times that shape only (slab walk, depth 16) and caches it."""
    side, x = config1_positions
    p = make_params(1000, 1000, side=side, seed=3, potential=BD_POT_WCA49_50_AO, **AO_FAST)
    e = Engine(p)
    e.set_positions(x)
    e.cell_locate()
    e.compute_forces()
    f0 = e.forces()
    cache = str(tmp_path / "tune.txt")
    assert e.autotune(3, cache) == 0
    assert (e.walk(), e.defer_depth(), e.pass_split()) == (abi.BD_WALK_SLABS, 16, 0)
    e.compute_forces()
    assert all(np.array_equal(a, c) for a, c in zip(f0, e.forces()))
    assert open(cache).read().split()[-4:] == ["0", "16", str(abi.BD_WALK_SLABS), "0"]
    e.close()


# ------------------------------------------------------------------ checkpoint
def test_checkpoint_resume_bit_exact(config1_positions, tmp_path):
    side, x = config1_positions
    specs = [(1000, 1000, side, 0.3, 21), (1000, 1000, side, 0.1, 22)]
    params = [make_params(nb, ns, side=sd, T0=1.0, Tf=T, first_step=-10, seed=sd_)
              for nb, ns, sd, T, sd_ in specs]
    a = Engine(params)
    for r in range(2):
        a.set_positions(x, r)
    a.step(30)
    a.set_origin(0)
    a.step(20)
    ck = str(tmp_path / "state.bdck")
    a.checkpoint_save(ck)
    a.step(30)
    b = Engine(params)
    b.checkpoint_load(ck)
    assert b.step_index() == a.step_index() - 30
    b.step(30)
    for r in range(2):
        assert np.array_equal(a.positions(r), b.positions(r))
        assert np.array_equal(a.raw_positions(r), b.raw_positions(r))
        assert all(np.array_equal(u, v) for u, v in zip(a.forces(r), b.forces(r)))
    assert np.array_equal(a.fself_origins(7.0, 1), b.fself_origins(7.0, 1))
    assert a.step_index() == b.step_index()
    # another engine's parameters, and a truncated file, are refused
    c = Engine([make_params(1000, 1000, side=side, seed=5)] * 2)
    with pytest.raises(BDError):
        c.checkpoint_load(ck)
    raw = open(ck, "rb").read()
    bad = str(tmp_path / "short.bdck")
    open(bad, "wb").write(raw[: len(raw) // 2])
    with pytest.raises(BDError):
        b.checkpoint_load(bad)
    for e in (a, b, c):
        e.close()


def test_persistent_steps_bit_identical(config1_positions):
    """One cooperative launch per 16 steps (grid barriers between the phases)
    reproduces the graph-replayed trajectory, also for a batch of replicas."""
    side, x = config1_positions
    params = [make_params(1000, 1000, side=side, T0=1.0, Tf=0.2, first_step=-5, seed=31),
              make_params(1000, 1000, side=side, T0=0.5, Tf=0.5, first_step=-5, seed=32)]
    out = []
    for persist in (0, 16):
        e = Engine(params)
        e.set_persistent(persist)
        assert e.persistent() == persist
        for r in range(2):
            e.set_positions(x, r)
        e.step(37)
        e.set_persistent(0)
        e.step(5)
        out.append(([e.positions(r) for r in range(2)], [e.forces(r)[0] for r in range(2)],
                    e.step_index()))
        e.close()
    for r in range(2):
        assert np.array_equal(out[0][0][r], out[1][0][r])
        assert np.array_equal(out[0][1][r], out[1][1][r])
    assert out[0][2] == out[1][2]
    e = Engine(params)
    with pytest.raises(BDError):
        e.set_persistent(-1)
    e.close()


def test_split_order_statistics_within_two_sigma(config1_positions):
    """Past the first steps the split order (not bit-identical) must agree with
    the reference order statistically: MSD and F_s(k=7) of a 16-replica
    ensemble after 1500 steps within 2 sigma of the replica spread."""
    side, x = config1_positions
    params = [make_params(1000, 1000, side=side, T0=1.0, Tf=1.0, seed=700 + r)
              for r in range(16)]
    stats = {}
    for order in (abi.BD_ORDER_REFERENCE, abi.BD_ORDER_SPLIT4):
        e = Engine(params)
        e.set_force_order(order)
        for r in range(16):
            e.set_positions(x, r)
        e.mark_origin()
        e.step(1500)
        ob = [e.observables(7.0, r) for r in range(16)]
        stats[order] = np.array([[o["msd_big"], o["msd_sml"], o["fself_big"]] for o in ob])
        e.close()
    a, b = stats[abi.BD_ORDER_REFERENCE], stats[abi.BD_ORDER_SPLIT4]
    err = np.sqrt(a.var(0, ddof=1) / 16 + b.var(0, ddof=1) / 16)
    assert np.all(np.abs(a.mean(0) - b.mean(0)) <= 2.0 * err + 1e-12), (a.mean(0), b.mean(0))


def test_stream_bandwidth_probe():
    """The HBM probe returns a plausible read + write rate and checks its arguments."""
    from bd_aging_amd.engine import stream_bandwidth
    gbs = stream_bandwidth(1 << 28, 3)
    assert 100.0 < gbs < 20000.0
    with pytest.raises(BDError):
        stream_bandwidth(1024, 3)


def test_grid_barrier_probe():
    """The barrier probe times co-resident grids, costs more with more
    workgroups, and refuses a grid that could not be co-resident."""
    from bd_aging_amd.engine import grid_barrier_cost
    us64, most = grid_barrier_cost(64, 200)
    assert most >= 256 and 0.0 < us64 < 1000.0
    us_most, _ = grid_barrier_cost(most, 200)
    assert us_most > us64
    with pytest.raises(BDError) as e:
        grid_barrier_cost(most + 1, 10)
    assert e.value.code == abi.BD_E_ARG


def test_launch_floor_probe():
    """The launch-floor probe times an almost empty kernel: a few microseconds,
    well below the force kernels it is compared with, and checks its arguments."""
    from bd_aging_amd.engine import launch_floor
    one = launch_floor(1, 64, 50)
    full = launch_floor(512, 256, 50)
    assert 0.0 < one < 100.0 and 0.0 < full < 100.0
    for bad in ((0, 64), (1, 32), (1, 100), (1, 2048)):
        with pytest.raises(BDError):
            launch_floor(*bad, 10)

"""GPU parity at the benchmark configurations, against the FP64 oracle.

* The N=262144 throughput-stress system (phi=0.58, T=0.5, dt=5e-5, 46^3 big
  and 64^3 small cells): cell lists and forces of the relaxed state bit-exact
  under the default and the batch-tuned force shapes, then 10 steps.
* A dilute N=120000 system whose small grid has 78 cells per row: the
  binning's row-chunked path (rows longer than one wavefront).
* The 256-replica uneven batch of the benchmark (N log-uniform in [1000,
  16000], phi in [0.52, 0.60], quench T in [0.05, 0.5]): the whole batch
  stepped in one engine (the slab walk, then the flat walk), eight replicas
  (smallest and largest N, extreme phi and T) compared with the oracle bit
  for bit.

Reference: cell_locate.cu:155-209, get_forces.cu:297-558, update_ermak.cu:91-160,
main.cu:213 (one seed per configuration) and 494-586 (the step loop).
"""
import numpy as np
import pytest

import orc
from bd_aging_amd import abi
from bd_aging_amd.engine import Engine, make_params, box_side
from bd_aging_amd.init import relax_overlaps, relax_batch

pytestmark = pytest.mark.gpu


def assert_cells_equal(e, o, r=0):
    for sp in (0, 1):
        ne, ce = e.cells(sp, r)
        no, co = o.cells(sp)
        assert np.array_equal(ne, no), ("occupancy", sp)
        assert np.array_equal(ce, co), ("cell lists", sp)


def assert_state_equal(e, o, r=0):
    assert np.array_equal(e.positions(r), o.positions())
    assert np.array_equal(e.raw_positions(r), o.raw_positions())
    fe, ve, ue = e.forces(r)
    fo, vo, uo = o.forces()
    assert np.array_equal(fe, fo) and np.array_equal(ve, vo) and np.array_equal(ue, uo)


@pytest.fixture(scope="module")
def stress_positions():
    side = box_side(131072, 131072, 1.4, 1.0, 0.58)
    return side, relax_overlaps(131072, 131072, side, seed=3)


def test_stress_config_cells_forces_and_steps(stress_positions):
    side, x = stress_positions
    p = make_params(131072, 131072, side=side, dt=5e-5, T0=0.5, Tf=0.5, seed=4242)
    e = Engine(p)
    g = e.geometry()
    assert (g.ncell_big, g.ncell_sml) == (46, 64)
    e.set_positions(x)
    e.cell_locate()
    o = orc.System(p)
    o.set_positions(x)
    assert o.cell_locate() == 0
    o.compute_forces()
    assert_cells_equal(e, o)
    fo, vo, uo = o.forces()
    # default slab walk, the chunked four-wavefront split, the row walk
    for walk, depth, split in ((abi.BD_WALK_SLABS, 16, 0), (abi.BD_WALK_CHUNK4, 16, 2),
                               (abi.BD_WALK_ROWS, 0, 0), (abi.BD_WALK_FLAT, 0, 0)):
        e.set_walk(walk)
        e.set_defer_depth(depth)
        e.set_pass_split(split)
        e.compute_forces()
        fe, ve, ue = e.forces()
        assert np.array_equal(fe, fo) and np.array_equal(ve, vo) and np.array_equal(ue, uo), \
            (walk, depth, split)
    e.set_walk(abi.BD_WALK_SLABS)
    e.set_defer_depth(16)
    e.set_pass_split(0)
    e.step(10)  # eager (shorter than a graph)
    assert o.step(10) == 0
    assert_state_equal(e, o)
    assert_cells_equal(e, o)
    e.close()
    o.close()


def test_long_rows_chunked_binning():
    """78 small cells per row: k_cells packs each row in chunks of 64 cells."""
    nb = ns = 60000
    side = box_side(nb, ns, 1.4, 1.0, 0.15)
    x = relax_overlaps(nb, ns, side, seed=8)
    p = make_params(nb, ns, side=side, T0=1.0, Tf=1.0, seed=808)
    e = Engine(p)
    assert e.geometry().ncell_sml > 64
    e.set_positions(x)
    e.cell_locate()
    o = orc.System(p)
    o.set_positions(x)
    assert o.cell_locate() == 0
    o.compute_forces()
    assert_cells_equal(e, o)
    e.compute_forces()
    assert_state_equal(e, o)
    e.step(20)  # one 16-step graph and four eager steps
    assert o.step(20) == 0
    assert_state_equal(e, o)
    assert_cells_equal(e, o)
    e.close()
    o.close()


def test_256_replica_batch_against_oracle():
    from bench import batch_specs
    n, phi, T = batch_specs()
    specs, params = [], []
    for r in range(len(n)):
        nb = int(n[r]) // 2
        ns = int(n[r]) - nb
        side = box_side(nb, ns, 1.4, 1.0, float(phi[r]))
        specs.append((nb, ns, side, 1.4, 1.0, 5000 + r))
        params.append(make_params(nb, ns, side=side, dt=1e-4, T0=1.0, Tf=float(T[r]), t0=0.0,
                                  tf=0.0, first_step=-5, seed=90000 + r))
    xs = relax_batch(specs)
    e = Engine(params)
    for r, x in enumerate(xs):
        e.set_positions(x, r)
    e.step(12)  # 5 steps at T = 1, the quench, 7 steps at T_r (eager)
    e.set_walk(abi.BD_WALK_FLAT)
    e.step(13)  # 13 more under the flat walk (eager)
    picks = []
    for r in [int(np.argmin(n)), int(np.argmax(n)), int(np.argmin(phi)), int(np.argmax(phi)),
              int(np.argmin(T)), int(np.argmax(T)), 0, len(n) - 1, *range(1, 9)]:
        if r not in picks and len(picks) < 8:
            picks.append(r)
    for r in picks:
        o = orc.System(params[r])
        o.set_positions(xs[r])
        assert o.cell_locate() == 0
        o.compute_forces()
        assert o.step(25) == 0
        assert_state_equal(e, o, r)
        assert_cells_equal(e, o, r)
        o.close()
    e.close()

"""Pins of the FP64 CPU oracle (no GPU).

Each test checks an oracle function against something independent of it:
golden vectors of the compiled reference, libm, brute force on small inputs,
closed forms and invariants.
"""
import math
import os
from fractions import Fraction

import numpy as np
import pytest

import orc
from bd_aging_amd.engine import make_params, box_side
from bd_aging_amd.abi import (BD_POT_WCA, BD_POT_WCA_MODIFIED, BD_POT_HARMONIC, BD_POT_WCA49_50,
                              BD_POT_WCA49_50_AO, BD_TEMP_SINE)

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


# ---------------------------------------------------------------- rannew64
def test_rannew64_matches_reference_golden_vectors():
    d = np.load(os.path.join(GOLDEN, "rannew64_ref.npz"))
    for seed, bits in zip(d["seeds"], d["bits"]):
        got = orc.Stream(int(seed)).doubles(bits.shape[0])
        assert np.array_equal(got.view(np.uint64), bits), f"seed {seed}"


def test_rannew64_double_is_u64_times_2m64():
    a = orc.Stream(99).u64(600)
    b = orc.Stream(99).doubles(600)
    assert np.array_equal(a.astype(np.float64) * 2.0 ** -64, b)


def test_rannew64_recurrence_independent_restatement():
    # x(n) = x(n-250) - x(n-103), y(n) = a y(n-1) + c, z = x - y  (mod 2^64),
    # restated in Python integers from the published recurrence.
    m = (1 << 64) - 1
    u = 4242
    for _ in range(1000):
        u = (u * 6364136223846793005 + 1442695040888963407) & m
    x = []
    for _ in range(250):
        u = (u * 6364136223846793005 + 1442695040888963407) & m
        x.append(u)
    y = u
    out = []
    # the ring is consumed from the top: x(n-250) is ring[np], x(n-103) is ring[nq]
    hist = list(reversed(x))  # hist[-250] = ring[249], the first x(n-250)
    for _ in range(700):
        xn = (hist[-250] - hist[-103]) & m
        hist.append(xn)
        y = (2862933555777941757 * y + 3037000493) & m
        out.append((xn - y) & m)
    got = orc.Stream(4242).u64(700)
    assert [int(v) for v in got] == out


def test_lane_seed_splitmix_known_answer():
    # SplitMix64 with state 0: first output 0xe220a8397b1dcdaf (Steele et al.)
    assert orc.lib().orc_lane_seed(0, 0) == 0xe220a8397b1dcdaf >> 1


# ---------------------------------------------------------------- Gaussian
def test_log_against_libm():
    rng = np.random.default_rng(1)
    xs = np.concatenate([rng.random(20000), 2.0 ** -rng.uniform(0, 53, 20000),
                         [1.0, 0.5, 2.0 ** -53, 0.7071067811865476, 0.7071067811865475]])
    xs = xs[xs > 0]
    f = orc.lib().orc_log
    for x in xs:
        ref = math.log(x)
        got = f(float(x))
        assert abs(got - ref) <= 4 * math.ulp(ref) + 1e-300, (x, got, ref)


def test_sincos_turn_against_exact():
    # exact reference: mpmath at 40 digits for the same double u
    import mpmath
    mpmath.mp.dps = 40
    rng = np.random.default_rng(2)
    us = np.concatenate([rng.random(4000), [0.0, 0.125, 0.25, 0.5, 0.75, 1.0 - 2 ** -53]])
    for u in us:
        s, c = orc.sincos_turn(u)
        a = 2 * mpmath.pi * mpmath.mpf(float(u))
        assert abs(s - float(mpmath.sin(a))) <= 2.5e-16
        assert abs(c - float(mpmath.cos(a))) <= 2.5e-16


def test_gauss3_distribution():
    from scipy import stats
    st = orc.Stream(2024)
    w = st.u64(4 * 30000).reshape(-1, 4)
    z = np.array([orc.gauss3(r) for r in w])
    for k in range(3):
        assert abs(z[:, k].mean()) < 0.03
        assert abs(z[:, k].var() - 1.0) < 0.03
        assert stats.kstest(z[:, k], "norm").pvalue > 1e-3
    # the three components are uncorrelated
    c = np.corrcoef(z.T)
    assert np.all(np.abs(c - np.eye(3)) < 0.03)


# ---------------------------------------------------------------- geometry
@pytest.mark.parametrize("pot,n,phi,exp", [
    (BD_POT_WCA_MODIFIED, 32768, 0.58, (25, 36, 2, 1)),   # reference-literal cells
    (BD_POT_WCA, 32768, 0.58, (23, 32, 2, 1)),
    (BD_POT_WCA, 2000, 0.50, (9, 13, 2, 1)),
    (BD_POT_WCA, 262144, 0.58, (46, 64, 2, 1)),
])
def test_geometry_reference_values(pot, n, phi, exp):
    p = make_params(n // 2, n // 2, phi=phi, potential=pot)
    rc, g = orc.geometry(p)
    assert rc == 0
    assert (g.ncell_big, g.ncell_sml, g.nrange_bs, g.nrange_sb) == exp
    assert g.nrange_bb == 1 and g.nrange_ss == 1
    # cells are at least the cutoff (main.cu:261-271 with 1.05 margin)
    assert g.cell_side_big >= 1.05 * g.rc_bb and g.cell_side_sml >= 1.05 * g.rc_ss


def test_box_side_config_values():
    assert abs(box_side(16384, 16384, 1.4, 1.0, 0.58) - 38.1161) < 1e-3
    assert abs(box_side(1000, 1000, 1.4, 1.0, 0.50) - 15.7684) < 1e-3
    assert abs(box_side(131072, 131072, 1.4, 1.0, 0.58) - 76.2323) < 1e-3


def test_geometry_rejects_tiny_box():
    p = make_params(4, 4, side=3.0)
    rc, _ = orc.geometry(p)
    assert rc != 0


# ---------------------------------------------------------------- helpers
def random_dilute(nb, ns, phi, seed, pot=BD_POT_WCA, **kw):
    p = make_params(nb, ns, phi=phi, potential=pot, seed=seed, **kw)
    rng = np.random.default_rng(seed)
    x = rng.random((nb + ns, 3)) * p.side
    return p, x


def np_cells(p, g, x, species):
    """numpy restatement of cell_locate_hst (cell_locate.cu:46-104)."""
    nb = p.n_big
    sl = slice(0, nb) if species == 0 else slice(nb, None)
    n = g.ncell_big if species == 0 else g.ncell_sml
    inv = 1.0 / (g.cell_side_big if species == 0 else g.cell_side_sml)
    ijk = (inv * x[sl]).astype(np.int64)
    ijk[ijk == n] = n - 1
    ci = ijk[:, 0] + n * (ijk[:, 1] + n * ijk[:, 2])
    return ci, n


# ---------------------------------------------------------------- cells
def test_cell_locate_ascending_and_complete():
    p, x = random_dilute(300, 300, 0.2, 5)
    # put some particles on cell faces and on the box edge (x == L)
    rc, g = orc.geometry(p)
    x[0] = [p.side, 0.0, g.cell_side_big]
    x[1] = [2 * g.cell_side_big, p.side, 0.5]
    x[400] = [g.cell_side_sml, p.side, p.side]
    s = orc.System(p)
    s.set_positions(x)
    assert s.cell_locate() == 0
    for sp, ntags in ((0, p.ntags_big), (1, p.ntags_sml)):
        nocup, cv = s.cells(sp)
        ci, n = np_cells(p, g, x, sp)
        cv = cv.reshape(-1, ntags)
        assert np.array_equal(nocup, np.bincount(ci, minlength=n ** 3))
        for c in np.nonzero(nocup)[0]:
            members = np.nonzero(ci == c)[0]
            assert np.array_equal(cv[c, :nocup[c]], members)  # ascending order
            assert np.all(cv[c, nocup[c]:] == -1)


def test_cell_overflow_reported():
    p, x = random_dilute(50, 50, 0.1, 3, ntags_big=2)
    x[:5] = x[0]  # five big particles in one cell
    s = orc.System(p)
    s.set_positions(x)
    assert s.cell_locate() == 3  # BD_E_CELL_OVERFLOW


# ---------------------------------------------------------------- forces
AO_KW = dict(sigma_pol=4.0, phi_pol=0.3)


@pytest.mark.parametrize("pot", [BD_POT_WCA, BD_POT_WCA_MODIFIED, BD_POT_HARMONIC,
                                 BD_POT_WCA49_50, BD_POT_WCA49_50_AO])
def test_forces_match_brute_force(pot):
    kw = AO_KW if pot == BD_POT_WCA49_50_AO else {}
    p, x = random_dilute(90, 110, 0.25, 11, pot=pot, kappa=50.0, **kw)
    s = orc.System(p)
    s.set_positions(x)
    s.cell_locate()
    s.compute_forces()
    f, v, u = s.forces()
    fb, vb, ub = orc.forces_brute(p, x)
    scale = np.abs(fb).max()
    assert scale > 0
    assert np.allclose(f, fb, rtol=1e-12, atol=1e-12 * scale)
    assert np.allclose(v, vb, rtol=1e-12, atol=1e-12 * np.abs(vb).max())
    assert np.allclose(u, ub, rtol=1e-12, atol=1e-12 * np.abs(ub).max())
    # Newton 3
    assert np.abs(f.sum(axis=0)).max() < 1e-9 * scale
    # neighbour sets equal the all-pairs sets
    counts, lst = s.neighbours(64)
    g = s.geometry()
    d = x[:, None, :] - x[None, :, :]
    d -= p.side * np.floor(d / p.side + 0.5)
    r2 = (d ** 2).sum(-1)
    big = np.arange(len(x)) < p.n_big
    rc = np.where(big[:, None] & big[None, :], g.rc_bb,
                  np.where(~big[:, None] & ~big[None, :], g.rc_ss, g.rc_bs))
    np.fill_diagonal(r2, np.inf)
    for m in range(len(x)):
        exp = set(np.nonzero(r2[m] < rc[m] ** 2)[0].tolist())
        assert set(lst[m, :counts[m]].tolist()) == exp


def pot_eval(kind, sigma, r, kappa=0.0, **kw):
    import ctypes as C
    p = make_params(4, 4, side=10.0, potential=kind, kappa=kappa, **kw)
    u = C.c_double(0.0)
    nf = C.c_double(0.0)
    orc.lib().orc_potential(C.byref(p), sigma, 1.0 / r, C.byref(u), C.byref(nf))
    return u.value, nf.value


def test_potential_closed_forms():
    c = 2.0 ** (1.0 / 6.0)
    for sigma in (1.0, 1.2, 1.4):
        for r in np.linspace(0.8 * sigma, c * sigma, 7):
            s = sigma / r
            u, nf = pot_eval(BD_POT_WCA, sigma, r)
            assert math.isclose(u, 4 * (s ** 12 - s ** 6) + 1, rel_tol=1e-13, abs_tol=1e-13)
            assert math.isclose(nf, -4 * (12 * s ** 12 - 6 * s ** 6) / r, rel_tol=1e-12,
                                abs_tol=1e-12)
            # wca_modified(sigma 2^(1/6)) == wca(sigma)
            um, nfm = pot_eval(BD_POT_WCA_MODIFIED, sigma * c, r)
            assert math.isclose(um, u, rel_tol=1e-12, abs_tol=1e-12)
            assert math.isclose(nfm, nf, rel_tol=1e-12, abs_tol=1e-12)
        u, nf = pot_eval(BD_POT_WCA, sigma, sigma)
        assert math.isclose(u, 1.0) and math.isclose(nf, -24.0 / sigma)
        u, nf = pot_eval(BD_POT_HARMONIC, sigma, 0.9 * sigma, kappa=10.0)
        assert math.isclose(u, 5.0 * (0.1 * sigma) ** 2) and math.isclose(nf, -10.0 * 0.1 * sigma)


def fd_slope(kind, sigma, r, h, **kw):
    """Central difference of the energy: dU/dr to O(h^2)."""
    return (pot_eval(kind, sigma, r + h, **kw)[0] - pot_eval(kind, sigma, r - h, **kw)[0]) / (2 * h)


@pytest.mark.parametrize("kind", [BD_POT_WCA49_50, BD_POT_WCA49_50_AO])
def test_49_50_repulsion_pinned(kind):
    """potentials.cu:9-23 / :27-41: U = invT (A (s^50 - s^49) + 1), s = sigma/r."""
    from bd_aging_amd.engine import PHS_AMP, PHS_INV_T
    kw = AO_KW if kind == BD_POT_WCA49_50_AO else {}
    for sigma in (1.0, 1.2, 1.4):
        # contact: s = 1, the bracket vanishes, dU/dr = -A invT / sigma
        u, nf = pot_eval(kind, sigma, sigma, **kw)
        assert u == PHS_INV_T
        assert math.isclose(nf, -PHS_AMP * PHS_INV_T / sigma, rel_tol=1e-14)
        # the force is minus the slope of the energy (exponents, signs)
        top = sigma * (50.0 / 49.0 if kind == BD_POT_WCA49_50 else 1.0)
        for r in np.linspace(0.97 * sigma, top - 1e-3 * sigma, 6):
            h = 1e-6 * sigma
            assert math.isclose(pot_eval(kind, sigma, r, **kw)[1], fd_slope(kind, sigma, r, h, **kw),
                                rel_tol=1e-6, abs_tol=1e-6)
        # closed form at a few s, exact rationals
        for r in (0.98 * sigma, 0.995 * sigma):
            sr = Fraction(sigma) / Fraction(r)
            exp_u = PHS_INV_T * (PHS_AMP * float(sr ** 50 - sr ** 49) + 1.0)
            assert math.isclose(pot_eval(kind, sigma, r, **kw)[0], exp_u, rel_tol=1e-12)
    if kind == BD_POT_WCA49_50:
        # the force vanishes at the cutoff (50/49) sigma: 50 s^50 - 49 s^49 = 0
        u, nf = pot_eval(kind, 1.0, 50.0 / 49.0)
        assert abs(nf) < 1e-12 * PHS_AMP
        # with the reference's A = 50 (50/49)^50 the energy there is -invT/49;
        # with A = 50 (50/49)^49 it is shifted to zero
        assert math.isclose(u, -PHS_INV_T / 49.0, rel_tol=1e-12)
        u0, _ = pot_eval(kind, 1.0, 50.0 / 49.0, pot_amp=50.0 * (50.0 / 49.0) ** 49)
        assert abs(u0) < 1e-12


def test_ao_depletion_branch_pinned():
    """potentials.cu:43-50: beyond contact U = -phi_pol (e^3 - 1.5 e^2 x + 0.5 x^3),
    e = 1 + sigma/sigma_pol, x = r/sigma."""
    for sigma, spol, phi in ((1.0, 1.0, 0.4), (1.4, 4.0, 0.3), (1.2, 2.5, 0.1)):
        kw = dict(sigma_pol=spol, phi_pol=phi)
        eta = 1.0 + sigma / spol
        # energy and force reach zero at the cutoff e sigma
        u, nf = pot_eval(BD_POT_WCA49_50_AO, sigma, eta * sigma, **kw)
        assert abs(u) < 1e-12 and abs(nf) < 1e-12
        for r in np.linspace(1.001 * sigma, 0.999 * eta * sigma, 7):
            u, nf = pot_eval(BD_POT_WCA49_50_AO, sigma, r, **kw)
            assert u < 0  # attraction
            h = 1e-6 * sigma
            assert math.isclose(nf, fd_slope(BD_POT_WCA49_50_AO, sigma, r, h, **kw),
                                rel_tol=1e-6, abs_tol=1e-9)
            if spol == sigma:
                # equal sizes: the textbook Asakura-Oosawa overlap volume,
                # U = -phi (1+q)^3/q^3 [1 - 3x/(2(1+q)) + x^3/(2(1+q)^3)], q = 1
                x = r / sigma
                exp = -phi * 8.0 * (1.0 - 3.0 * x / 4.0 + x ** 3 / 16.0)
                assert math.isclose(u, exp, rel_tol=1e-12, abs_tol=1e-14)
        # inside contact the 49-50 repulsion applies (same value as kind 3)
        u, nf = pot_eval(BD_POT_WCA49_50_AO, sigma, 0.99 * sigma, **kw)
        u3, nf3 = pot_eval(BD_POT_WCA49_50, sigma, 0.99 * sigma)
        assert math.isclose(u, u3, rel_tol=1e-14) and math.isclose(nf, nf3, rel_tol=1e-14)


def test_49_50_geometry():
    p = make_params(1000, 1000, phi=0.5, potential=BD_POT_WCA49_50)
    rc, g = orc.geometry(p)
    assert rc == 0
    assert (g.rc_bb, g.rc_ss, g.rc_bs) == tuple((50.0 / 49.0) * s for s in (1.4, 1.0, 1.2))
    p = make_params(1000, 1000, phi=0.5, potential=BD_POT_WCA49_50_AO, sigma_pol=4.0, phi_pol=0.3)
    rc, g = orc.geometry(p)
    assert rc == 0
    assert (g.rc_bb, g.rc_ss, g.rc_bs) == tuple(s * (1.0 + s / 4.0) for s in (1.4, 1.0, 1.2))
    for bad in (dict(sigma_pol=0.0, phi_pol=0.3), dict(sigma_pol=4.0, phi_pol=-1.0)):
        p = make_params(1000, 1000, phi=0.5, potential=BD_POT_WCA49_50_AO, **bad)
        assert orc.geometry(p)[0] != 0
    p = make_params(1000, 1000, phi=0.5, potential=BD_POT_WCA49_50, pot_inv_t=-1.0)
    assert orc.geometry(p)[0] != 0


# ---------------------------------------------------------------- update
def test_update_zero_temperature_closed_form():
    p, x = random_dilute(40, 60, 0.2, 21, T0=0.0, Tf=0.0)
    s = orc.System(p)
    s.set_positions(x)
    rng = np.random.default_rng(3)
    f = rng.normal(size=x.shape) * 50.0
    s.set_forces(f)
    assert s.update() == 0
    sig = np.where(np.arange(len(x)) < p.n_big, p.sigma_big, p.sigma_sml)[:, None]
    d = f * (1.0 / sig) * p.dt
    exp = x + d
    exp -= p.side * np.floor((1.0 / p.side) * exp)
    assert np.array_equal(s.positions(), exp)
    assert np.array_equal(s.raw_positions(), x + d)


def jittered_lattice(nb, ns, phi, seed, jitter=0.08, **kw):
    """nb + ns particles (big first) on random sites of a cubic lattice filling
    the box at packing fraction phi, each moved by up to `jitter` lattice
    spacings: pairs inside the cutoffs, none deeply overlapping."""
    p = make_params(nb, ns, phi=phi, seed=seed, **kw)
    n = nb + ns
    m = int(math.ceil(n ** (1.0 / 3.0)))
    a = p.side / m
    rng = np.random.default_rng(seed)
    sites = rng.permutation(m ** 3)[:n]
    ijk = np.stack([sites % m, (sites // m) % m, sites // (m * m)], axis=1)
    x = (ijk + 0.5 + rng.uniform(-jitter, jitter, (n, 3))) * a
    return p, np.mod(x, p.side)


def np_wca_pairs(p, x):
    """All ordered pairs i != j inside the cutoff with numpy: separation
    r_ij = r_i - r_j (minimum image), distance, U and dU/dr from the textbook
    WCA forms (not from the oracle's potential routine):
      WCA:          U = 4 (s^12 - s^6) + 1,   r < 2^(1/6) sigma_ij
      WCA modified: U = s^12 - 2 s^6 + 1,     r < sigma_ij
    with s = sigma_ij / r and sigma_ij = (sigma_i + sigma_j) / 2."""
    n = p.n_big + p.n_sml
    sig = np.where(np.arange(n) < p.n_big, p.sigma_big, p.sigma_sml)
    sij = 0.5 * (sig[:, None] + sig[None, :])
    d = x[:, None, :] - x[None, :, :]
    d -= p.side * np.round(d / p.side)
    r = np.sqrt((d ** 2).sum(-1))
    np.fill_diagonal(r, np.inf)
    if p.potential == BD_POT_WCA:
        rc = 2.0 ** (1.0 / 6.0) * sij
    else:
        rc = sij
    inside = r < rc
    s = np.where(inside, sij / np.where(inside, r, 1.0), 0.0)
    if p.potential == BD_POT_WCA:
        U = np.where(inside, 4.0 * (s ** 12 - s ** 6) + 1.0, 0.0)
        dU = np.where(inside, -(48.0 * s ** 12 - 24.0 * s ** 6) / np.where(inside, r, 1.0), 0.0)
    else:
        U = np.where(inside, s ** 12 - 2.0 * s ** 6 + 1.0, 0.0)
        dU = np.where(inside, -(12.0 * s ** 12 - 12.0 * s ** 6) / np.where(inside, r, 1.0), 0.0)
    return r, U, dU, inside


@pytest.mark.parametrize("pot", [BD_POT_WCA, BD_POT_WCA_MODIFIED])
@pytest.mark.parametrize("phi", [0.30, 0.55])
def test_energy_virial_z_pinned_independently(pot, phi):
    """main.cu:627-638 (e_s = sum over species s of pot / 2 / N_s, pot the sum of
    U over a particle's partners) and main.cu:801-806 with get_forces.cu:134,265
    (Z = 1 + sum_i vir_i / (3 N_tot T), vir_i = -1/2 sum_j (r_j - r_i).F_ij,
    i.e. sum_i vir_i = sum over unordered pairs of r_ij.F_ij = -r dU/dr) from
    an all-pairs numpy evaluation, on 80 big + 120 small particles (N_big,
    N_sml and N_tot all differ) at T = 0.7."""
    nb, ns, T = 80, 120, 0.7
    p, x = jittered_lattice(nb, ns, phi, 31, potential=pot, T0=T, Tf=T)
    s = orc.System(p)
    s.set_positions(x)
    assert s.cell_locate() == 0
    s.compute_forces()
    o = s.observables(7.0)
    s.close()
    r, U, dU, inside = np_wca_pairs(p, x)
    assert inside.sum() >= 50  # the check has pairs to see
    e_i = U.sum(axis=1)  # pot of particle i: sum over its partners
    e_b = 0.5 * e_i[:nb].sum() / nb
    e_s = 0.5 * e_i[nb:].sum() / ns
    w = -np.where(inside, r, 0.0) * dU  # r_ij . F_ij per ordered pair
    vir = 0.5 * w.sum()  # each unordered pair once
    z = 1.0 + vir / (3.0 * (nb + ns) * T)
    assert vir > 0  # repulsive pairs push outwards
    rel = lambda a, b: abs(a - b) / abs(b)  # noqa: E731
    assert rel(o["epot_big"], e_b) <= 1e-12
    assert rel(o["epot_sml"], e_s) <= 1e-12
    assert rel(o["virial"], vir) <= 1e-12
    assert rel(o["z"] - 1.0, z - 1.0) <= 1e-12
    assert o["temperature"] == T
    # each plausible slip moves a value far outside the tolerance: a
    # missing 1/2, a flipped virial, N_big for N_tot, the species swapped
    for wrong_e in (2 * e_b, e_i[:nb].sum() * 0.5 / ns):
        assert rel(o["epot_big"], wrong_e) > 1e-3
    for wrong_z in (1.0 - vir / (3.0 * (nb + ns) * T), 1.0 + vir / (3.0 * nb * T),
                    1.0 + 2 * vir / (3.0 * (nb + ns) * T), 1.0 + vir / (3.0 * (nb + ns))):
        assert rel(o["z"] - 1.0, wrong_z - 1.0) > 1e-3


def test_z_is_one_without_interactions():
    # no pair within any cutoff: every virial is zero and Z = 1 exactly
    p, x = jittered_lattice(40, 60, 0.01, 5, jitter=0.0, T0=0.4, Tf=0.4)
    r, _, _, inside = np_wca_pairs(p, x)
    assert not inside.any()
    s = orc.System(p)
    s.set_positions(x)
    assert s.cell_locate() == 0
    s.compute_forces()
    o = s.observables(7.0)
    s.close()
    assert o["virial"] == 0.0 and o["z"] == 1.0
    assert o["epot_big"] == 0.0 and o["epot_sml"] == 0.0


def test_free_diffusion_msd_and_fself():
    # zero forces: MSD = 6 D t and F_s(k, t) = exp(-k^2 D t), D = T / sigma
    p, x = random_dilute(1500, 1500, 0.05, 8, T0=0.7, Tf=0.7, dt=1e-3)
    s = orc.System(p)
    s.set_positions(x)
    nsteps = 200
    for _ in range(nsteps):
        assert s.update() == 0
    t = nsteps * p.dt
    o = s.observables(2.0)
    D_b, D_s = 0.7 / p.sigma_big, 0.7 / p.sigma_sml
    assert abs(o["msd_big"] / (6 * D_b * t) - 1) < 0.06
    assert abs(o["msd_sml"] / (6 * D_s * t) - 1) < 0.06
    assert abs(o["fself_big"] - math.exp(-4.0 * D_b * t)) < 0.02
    assert o["temperature"] == 0.7


def test_rng_lanes_lockstep_and_discards():
    # 7 particles on 3 lanes: every lane takes 4*ceil(7/3) = 12 draws per step
    p = make_params(3, 4, side=20.0, rng_lanes=3, seed=77)
    s = orc.System(p)
    s.set_positions(np.random.default_rng(0).random((7, 3)) * 20.0)
    s.update()
    s.update()
    for lane in range(3):
        ref = orc.Stream(orc.lib().orc_lane_seed(77, lane))
        ref.u64(24)
        ring, lcg, npp, nq = s.rng_lane(lane)
        assert np.array_equal(ring, np.frombuffer(bytes(ref.s.ring), dtype=np.uint64))
        assert lcg == ref.s.lcg and npp == ref.s.np and nq == ref.s.nq


# ---------------------------------------------------------------- temperature
def test_temperature_protocols():
    L = orc.lib()
    assert L.orc_temp_linear(1.0, 0.1, 2.0, 4.0, 1.0) == 1.0
    assert L.orc_temp_linear(1.0, 0.1, 2.0, 4.0, 5.0) == 0.1
    assert math.isclose(L.orc_temp_linear(1.0, 0.1, 2.0, 4.0, 3.0), 0.55)
    # instant quench: the reference returns NaN at t == t0 == tf
    assert L.orc_temp_linear(1.0, 0.1, 0.0, 0.0, 0.0) == 0.1
    assert L.orc_temp_linear(1.0, 0.1, 0.0, 0.0, -1e-4) == 1.0
    # sine a + b sin(2 pi (t/P - 1/4)): minimum T0 at t = 0, maximum Tf at t = P/2
    assert math.isclose(L.orc_temp_sine(0.2, 1.0, 0.0, 10.0, 4.0, 0.0), 0.2, abs_tol=1e-15)
    assert math.isclose(L.orc_temp_sine(0.2, 1.0, 0.0, 10.0, 4.0, 2.0), 1.0, abs_tol=1e-15)
    assert math.isclose(L.orc_temp_sine(0.2, 1.0, 0.0, 10.0, 4.0, 1.0), 0.6, abs_tol=1e-15)
    # update at first_step uses T0; later updates use T(dt*ni)
    p = make_params(2, 2, side=10.0, T0=1.0, Tf=0.1, t0=0.0, tf=0.0, first_step=-3, dt=0.5)
    assert L.orc_temperature(p, -3) == 1.0
    assert L.orc_temperature(p, -2) == 1.0   # t = -1.0 < t0
    assert L.orc_temperature(p, 0) == 0.1    # t = 0 == t0 == tf
    p.temp_protocol = BD_TEMP_SINE
    p.period = 4.0
    p.tf = 10.0
    assert math.isclose(L.orc_temperature(p, 5), 0.55 - 0.45 * math.sin(2 * math.pi * (2.5 / 4 - 0.25)),
                        abs_tol=1e-15)


# ---------------------------------------------------------------- sampling
def f32_round_half_away(x):
    q = Fraction(float(x))
    s = 1 if q >= 0 else -1
    q = abs(q)
    fl = q.numerator // q.denominator
    return s * (fl + 1 if q - fl >= Fraction(1, 2) else fl)


def numpy_sampling_indexes(minimal_decade, run_time, dt, n):
    f = np.float32
    last = f32_round_half_away(np.log10(f(run_time)))
    first = f32_round_half_away(np.log10(f(minimal_decade)))
    # powf(10, 1/n) correctly rounded (numpy's float32 power is not)
    import mpmath
    mpmath.mp.dps = 50
    r = f(float(mpmath.power(10, mpmath.mpf(float(f(1.0) / f(n))))))
    vals = []
    x = f(1.0)
    for _ in range(n + 1):
        vals.append(x)
        x = f(x * r)
    out = [0]
    for ip in range(first, last):
        power = f(10.0 ** ip)
        for j in range(n):
            st = f(vals[j] * power)
            out.append(f32_round_half_away(f(st / f(dt))))
    return out


def test_sampling_indexes():
    got = orc.sampling_indexes(1e-3, 10.0, 1e-4).tolist()
    assert got[:13] == [0, 10, 13, 16, 20, 25, 32, 40, 50, 63, 79, 100, 126]
    assert len(got) == 41 and got[-1] == 79433
    assert got == numpy_sampling_indexes(1e-3, 10.0, 1e-4, 10)
    for args in [(1e-2, 1e3, 1e-4), (1e-3, 1.0, 5e-5), (0.1, 100.0, 1e-3)]:
        assert orc.sampling_indexes(*args).tolist() == numpy_sampling_indexes(*args, 10)


# ---------------------------------------------------------------- g(r)
def np_gder(p, x, nbins, rng):
    nb = p.n_big
    out = np.zeros((4, nbins), dtype=np.uint64)
    bs = rng / nbins
    for m in range(len(x)):
        d = x[m] - x
        d -= p.side * np.floor((1.0 / p.side) * d + 0.5)
        dist = np.sqrt(d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1] + d[:, 2] * d[:, 2])
        idx = np.arange(len(x))
        ok = (dist < rng) & (idx != m)
        b = (dist[ok] / bs).astype(np.int64)
        b[b == nbins] = nbins - 1
        pr = 2 * (m >= nb) + (idx[ok] >= nb)
        np.add.at(out, (pr, b), 1)
    return out


def test_gder_brute_force_and_total():
    p, x = random_dilute(120, 150, 0.2, 13)
    s = orc.System(p)
    s.set_positions(x)
    nbins, rng = 50, p.side / 2
    c = s.gder_counts(nbins, rng)
    assert np.array_equal(c, np_gder(p, x, nbins, rng))
    # bs and sb count the same unordered pairs
    assert c[1].sum() == c[2].sum()


def test_gder_ideal_gas_normalisation():
    p, x = random_dilute(1000, 1000, 0.05, 17)
    s = orc.System(p)
    s.set_positions(x)
    nbins, rng = 20, p.side / 2
    c = s.gder_counts(nbins, rng)
    g = orc.gder_normalize(p, nbins, rng, c)
    vol = p.side ** 3
    pi = float(np.float32(3.14159266))
    for k, sig in enumerate([p.sigma_big, 1.2, 1.2, p.sigma_sml]):
        vfree = vol - 4.0 / 3.0 * pi * sig ** 3
        # uniform positions: g = V_free / V (the reference's normalisation)
        assert abs(g[k, 5:].mean() / (vfree / vol) - 1) < 0.02


# ---------------------------------------------------------------- contraction
def test_jacobi_reconstructs():
    rng = np.random.default_rng(4)
    a = rng.normal(size=(12, 12))
    a = a + a.T
    import ctypes as C
    w = np.zeros(12)
    v = np.zeros((12, 12))
    orc.lib().orc_jacobi_eig(12, orc.dp(np.ascontiguousarray(a)), orc.dp(w), orc.dp(v))
    assert np.allclose(v @ np.diag(w) @ v.T, a, atol=1e-12)
    assert np.allclose(np.sort(w), np.linalg.eigvalsh(a), atol=1e-12)


def test_cv3d_solve_matches_numpy_pinv():
    rng = np.random.default_rng(6)
    bins = 10
    ata = np.zeros((3, bins, bins))
    atb = rng.normal(size=(3, bins))
    for a in range(3):
        h = rng.normal(size=(30, bins))
        h[:, -2:] = 0.0  # rank deficient: two empty bins
        ata[a] = h.T @ h
    got = orc.cv3d_solve(ata, atb)
    exp = sum(np.linalg.pinv(ata[a], rcond=1e-6) @ atb[a] for a in range(3)) / 3.0
    assert np.allclose(got, exp, rtol=1e-9, atol=1e-10)


def test_cv3d_recovers_binwise_radial_force():
    # forces built as F_m = sum_b f_b * hit_b(m) are recovered exactly
    rng = np.random.default_rng(9)
    n, side, bins, rr = 300, 9.0, 8, 3.0
    x = rng.random((n, 3)) * side
    f_true = np.linspace(2.0, -1.0, bins)
    d = x[:, None, :] - x[None, :, :]
    d -= side * np.floor(d / side + 0.5)
    dist = np.sqrt((d ** 2).sum(-1))
    np.fill_diagonal(dist, np.inf)
    F = np.zeros((n, 3))
    for m in range(n):
        ok = dist[m] < rr
        b = np.minimum((dist[m, ok] * (bins / rr)).astype(int), bins - 1)
        F[m] = (f_true[b][:, None] * d[m, ok] / dist[m, ok][:, None]).sum(0)
    ata, atb = orc.cv3d_accumulate(bins, rr, x, F, side)
    assert np.allclose(ata, np.transpose(ata, (0, 2, 1)))
    got = orc.cv3d_solve(ata, atb)
    assert np.allclose(got, f_true, rtol=1e-8, atol=1e-8)

"""CPU checks of the two exact rewrites the GPU force walk relies on, against
the oracle (no GPU):

* packed rows: the stencil cells k_del = klo..khi of a row (x, y) are one or
  two contiguous slot ranges of a row packed in (z, ascending id) order, and
  walking them (wrapped low cells first, wrapped high cells last) meets the
  interacting partners in the reference's order (get_forces.cu:360-416);
* image shift: for grids of at least 2*nrange+3 cells,
  dx - L*floor(dx/L + 0.5) == dx + K*L with K the stencil cell's image index.

The walk below restates the kernel's index arithmetic in Python; the oracle's
neighbour lists (reference traversal) are the expected output.
"""
import numpy as np
import pytest

import orc
from bd_aging_amd.engine import make_params, box_side


def _cell(inv, n, x):
    ijk = [int(inv * c) for c in x]
    return [v - 1 if v == n else v for v in ijk]


def _slab_gap(d, cs, u):
    return d * cs - u if d > 0 else (u + (-d - 1) * cs if d < 0 else 0.0)


def _shift(hd, n, L):
    return -L if hd < 0 else (L if hd >= n else 0.0)


def _pack(o, sp, n, ntags, sp_off, xyz):
    """Row-packed slots from the oracle's reference-layout cell lists."""
    nocup, cv = o.cells(sp)
    cv = cv.reshape(n ** 3, ntags)
    rs = np.zeros(n * n * (n + 1), dtype=np.int64)
    cp = [None] * (n ** 3 * ntags)
    for row in range(n * n):
        base, carry = row * n * ntags, 0
        for z in range(n):
            c = row + n * n * z
            m = min(int(nocup[c]), ntags)
            rs[row * (n + 1) + z] = base + carry
            for k in range(m):
                idk = int(cv[c, k])
                cp[base + carry + k] = (*xyz[sp_off + idk], sp_off + idk)
            carry += m
        rs[row * (n + 1) + n] = base + carry
    return rs, cp


def _walk(NR, n, cs, rc2, self_, x, hw, rs, cp, L, out):
    lim = rc2 + 1e-9
    u = [x[a] - hw[a] * cs for a in range(3)]
    g2lo = [_slab_gap(-(k + 1), cs, u[2]) ** 2 for k in range(NR)]
    g2hi = [_slab_gap(k + 1, cs, u[2]) ** 2 for k in range(NR)]
    for di in range(-NR, NR + 1):
        gx2 = _slab_gap(di, cs, u[0]) ** 2
        if gx2 > lim:
            continue
        shx, a = _shift(hw[0] + di, n, L), (hw[0] + di) % n
        for dj in range(-NR, NR + 1):
            gxy2 = gx2 + _slab_gap(dj, cs, u[1]) ** 2
            if gxy2 > lim:
                continue
            shy, b = _shift(hw[1] + dj, n, L), (hw[1] + dj) % n
            klo = khi = 0
            for k in range(NR):
                if gxy2 + g2lo[k] <= lim:
                    klo = -(k + 1)
                if gxy2 + g2hi[k] <= lim:
                    khi = k + 1
            ro = (a + n * b) * (n + 1)
            z0, z1 = hw[2] + klo, hw[2] + khi
            if z0 < 0:
                segs = [(rs[ro + z0 + n], rs[ro + n], -L), (rs[ro], rs[ro + z1 + 1], 0.0)]
            elif z1 >= n:
                segs = [(rs[ro + z0], rs[ro + n], 0.0), (rs[ro], rs[ro + z1 - n + 1], L)]
            else:
                segs = [(rs[ro + z0], rs[ro + z1 + 1], 0.0)]
            center = self_ >= 0 and di == 0 and dj == 0
            for s0, s1, shz in segs:
                for s in range(s0, s1):
                    q = cp[s]
                    if center and q[3] == self_:
                        continue
                    dx, dy, dz = (q[0] - x[0]) + shx, (q[1] - x[1]) + shy, (q[2] - x[2]) + shz
                    if dx * dx + dy * dy + dz * dz < rc2:
                        out.append(q[3])


@pytest.mark.parametrize("nb,ns,phi,seed", [(1000, 1000, 0.50, 1), (300, 340, 0.52, 2),
                                            (500, 500, 0.60, 3)])
def test_row_walk_meets_partners_in_reference_order(nb, ns, phi, seed):
    side = box_side(nb, ns, 1.4, 1.0, phi)
    p = make_params(nb, ns, side=side)
    o = orc.System(p)
    rng = np.random.default_rng(seed)
    xyz = rng.uniform(0.0, side, (nb + ns, 3))
    xyz[:4] = [[0.0, 0.0, 0.0], [side * (1 - 1e-16), 1.0, 2.0],
               [1e-300, side / 2, side - 1e-12], [side / 2, 0.0, side * 0.999999]]
    o.set_positions(xyz)
    assert o.cell_locate() == 0
    xyz = o.positions()
    g = o.geometry()
    nr = {(0, 0): g.nrange_bb, (1, 1): g.nrange_ss, (0, 1): g.nrange_bs, (1, 0): g.nrange_sb}
    rc = {(0, 0): g.rc_bb, (1, 1): g.rc_ss, (0, 1): g.rc_bs, (1, 0): g.rc_bs}
    assert g.ncell_big >= 2 * max(g.nrange_bb, g.nrange_sb) + 3
    assert g.ncell_sml >= 2 * max(g.nrange_ss, g.nrange_bs) + 3
    grids = {}
    for sp in (0, 1):
        n = g.ncell_big if sp == 0 else g.ncell_sml
        cs = g.cell_side_big if sp == 0 else g.cell_side_sml
        grids[sp] = (n, cs, *_pack(o, sp, n, 8, 0 if sp == 0 else nb, xyz))
    counts, lst = o.neighbours(160)
    for i in range(nb + ns):
        gi = 0 if i < nb else 1
        out = []
        for gg in (gi, 1 - gi):
            n, cs, rs, cp = grids[gg]
            _walk(nr[(gi, gg)], n, cs, rc[(gi, gg)] ** 2, i if gg == gi else -1, xyz[i],
                  _cell(1.0 / cs, n, xyz[i]), rs, cp, side, out)
        assert out == list(lst[i, :counts[i]]), i
    o.close()


def test_image_shift_equals_floor_form():
    rng = np.random.default_rng(7)
    for n, nr in ((7, 2), (9, 1), (13, 2), (32, 2)):
        L = 17.3 + n
        cs = L / n
        x = rng.uniform(0.0, L, 200000)
        h = np.minimum((x / cs).astype(int), n - 1)
        for d in range(-nr, nr + 1):
            hd = h + d
            c = hd % n
            q = (c + rng.uniform(0.0, 1.0, x.size)) * cs  # a point of the stencil cell
            q = np.minimum(q, np.nextafter(L, 0.0))
            dx = q - x
            floor_form = dx - L * np.floor((1.0 / L) * dx + 0.5)  # get_forces.cu:389-391
            K = np.where(hd < 0, -1.0, np.where(hd >= n, 1.0, 0.0))
            shift_form = dx + K * L
            assert n >= 2 * nr + 3
            assert np.array_equal(floor_form, shift_form), (n, nr, d)

"""Initial configurations from a seed (no reference counterpart: the
reference's run-fself.sh:131-138 calls external set_init_config and
eliminate_overlap binaries that are not part of it).

Uniform random placement (one rannew64 stream, host) followed by overdamped
overlap removal on the GPU engine: BD steps at T = 0 with the harmonic
potential U = kappa/2 (sigma_ij - r)^2, r < sigma_ij, until the largest
overlap is below `tol * sigma_sml`.  The result is put in Z-order per
species (spatial_sort).
"""
import numpy as np

from .abi import BD_POT_HARMONIC
from .engine import Engine, make_params, random_positions


def _spread_bits(v):
    v = v.astype(np.uint64) & np.uint64(0x3ff)
    v = (v | (v << np.uint64(16))) & np.uint64(0x030000ff)
    v = (v | (v << np.uint64(8))) & np.uint64(0x0300f00f)
    v = (v | (v << np.uint64(4))) & np.uint64(0x030c30c3)
    v = (v | (v << np.uint64(2))) & np.uint64(0x09249249)
    return v


def morton_order(x, side):
    """Permutation putting positions in Z-order of a 1024^3 grid (stable)."""
    q = np.clip((x / side * 1024.0).astype(np.int64), 0, 1023)
    code = _spread_bits(q[:, 0]) | (_spread_bits(q[:, 1]) << np.uint64(1)) | \
        (_spread_bits(q[:, 2]) << np.uint64(2))
    return np.argsort(code, kind="stable")


def spatial_sort(x, n_big, side):
    """Reorder each species along the same Z-order curve, so that particles
    with neighbouring indices are neighbours in space (wavefront lanes then
    share cache lines in the cell gathers).  Species stay big first."""
    xb, xs = x[:n_big], x[n_big:]
    return np.concatenate([xb[morton_order(xb, side)], xs[morton_order(xs, side)]])


def relax_overlaps(n_big, n_sml, side, sigma_big=1.4, sigma_sml=1.0, seed=1, kappa=1.0,
                   dt=0.1, tol=1e-3, max_steps=200000, chunk=256, start=None, graph_steps=16):
    """Return an N x 3 array with no pair closer than sigma_ij (1 - tol).
    graph_steps: the engine's launch mode (steps per HIP graph, 0 = eager)."""
    p = make_params(n_big, n_sml, side=side, sigma_big=sigma_big, sigma_sml=sigma_sml,
                    potential=BD_POT_HARMONIC, kappa=kappa, dt=dt, T0=0.0, Tf=0.0,
                    ntags_big=24, ntags_sml=24, seed=seed)
    n = n_big + n_sml
    x = random_positions(n, side, seed) if start is None else np.asarray(start, float)
    # U per particle sums kappa/4 delta^2 over its overlaps, so max overlap
    # below tol*sigma_sml holds once every U < kappa/4 (tol*sigma_sml)^2.
    target = 0.25 * kappa * (tol * sigma_sml) ** 2
    with Engine(p) as eng:
        eng.set_graph_steps(graph_steps)
        eng.set_positions(x)
        done = 0
        while done < max_steps:
            eng.step(chunk)
            done += chunk
            _, _, u = eng.forces()
            if u.max() < target:
                break
        else:
            raise RuntimeError(f"overlaps remain after {max_steps} steps (max U {u.max():.3e})")
        return spatial_sort(eng.positions(), n_big, side)


def relax_batch(specs, kappa=1.0, dt=0.1, tol=1e-3, max_steps=200000, chunk=256,
                graph_steps=16):
    """Overlap removal for many systems in one batched engine.

    specs: list of (n_big, n_sml, side, sigma_big, sigma_sml, seed);
    graph_steps: the engine's launch mode (steps per HIP graph, 0 = eager).
    Returns a list of N_r x 3 arrays.
    """
    params = []
    for nb, ns, side, sb, ss, seed in specs:
        params.append(make_params(nb, ns, side=side, sigma_big=sb, sigma_sml=ss,
                                  potential=BD_POT_HARMONIC, kappa=kappa, dt=dt, T0=0.0,
                                  Tf=0.0, ntags_big=24, ntags_sml=24, seed=seed))
    sig_min = min(min(s[3], s[4]) for s in specs)
    target = 0.25 * kappa * (tol * sig_min) ** 2
    with Engine(params) as eng:
        eng.set_graph_steps(graph_steps)
        for r, (nb, ns, side, sb, ss, seed) in enumerate(specs):
            eng.set_positions(random_positions(nb + ns, side, seed), r)
        done = 0
        while done < max_steps:
            eng.step(chunk)
            done += chunk
            worst = max(eng.forces(r)[2].max() for r in range(len(specs)))
            if worst < target:
                break
        else:
            raise RuntimeError(f"overlaps remain after {max_steps} steps")
        return [spatial_sort(eng.positions(r), specs[r][0], specs[r][2])
                for r in range(len(specs))]

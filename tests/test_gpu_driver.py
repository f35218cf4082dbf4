"""The drop-in driver (bd_aging_self) against a replay of the main.cu flow on
the FP64 oracle: same init_config files, same stdin protocol, same samples."""
import os
import subprocess

import numpy as np
import pytest

import orc
from bd_aging_amd import abi
from bd_aging_amd.engine import make_params, lane_seed

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "bd_aging_amd", "bd_aging_self")

PROTOCOL = """linear
1.0  0.1  0.0  0.0  1
1e-4
linear
0.002  0.006  3
8 8
2
40 4.0
7.0
4321
"""


def read_config(path):
    lines = open(path).read().splitlines()
    nb, ns, nt = (int(v) for v in lines[2].split()[:3])
    side = float(lines[1].split()[0])
    sb, ss = (float(v) for v in lines[3].split()[:2])
    xyz = np.array([[float(v) for v in ln.split()[1:4]] for ln in lines[6:6 + nt]])
    return nb, ns, side, sb, ss, xyz


@pytest.mark.parametrize("pot_name,pot,dt", [("wca_modified", abi.BD_POT_WCA_MODIFIED, 1e-4),
                                             ("wca49_50", abi.BD_POT_WCA49_50, 1e-5)])
def test_driver_matches_oracle_replay(tmp_path, pot_name, pot, dt):
    """BD_POTENTIAL selects the pair potential of the configurations'
    equilibration and of the run.  The 49-50 potential takes dt = 1e-5, its
    wall being too stiff for explicit steps of 1e-4; its configurations are
    equilibrated with it (wca_modified leaves pairs at 0.9 sigma at T = 1,
    which the s^50 wall would throw apart)."""
    cfg = tmp_path / "configs"
    sim = tmp_path / "sim"
    (sim / "results").mkdir(parents=True)
    env = dict(os.environ, BD_POTENTIAL=pot_name)
    r = subprocess.run([EXE, "--make-configs", "250", "250", "0.45", "1.4", "1.0", "2", "77",
                        str(cfg), "1.0", "50", f"{dt:g}"], capture_output=True, text=True,
                       timeout=120, env=env)
    assert r.returncode == 0, r.stdout + r.stderr
    protocol = PROTOCOL
    if dt != 1e-4:  # the same 20 + 60 steps: (int)(time / dt) of the driver
        protocol = protocol.replace("\n1e-4\n", f"\n{dt:g}\n").replace(
            "0.002  0.006  3", f"{20.1 * dt:.6g}  {60.1 * dt:.6g}  3")
    r = subprocess.run([EXE], input=protocol, cwd=sim, capture_output=True, text=True,
                       timeout=120, env=env)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "TERMINADO" in r.stdout
    # replay on the oracle
    ntrans, niter, nsamples, nbins, rng, q = 20, 60, 3, 40, 4.0, 7.0
    samples = list(range(0, niter, niter // nsamples))
    acc = {k: np.zeros(len(samples)) for k in ("fs", "mb", "ms", "eb", "es", "z", "T")}
    gsum = np.zeros((len(samples), 4, nbins))
    for i in range(2):
        nb, ns, side, sb, ss, xyz = read_config(cfg / f"init_config_{i + 1}")
        p = make_params(nb, ns, side=side, sigma_big=sb, sigma_sml=ss, dt=dt,
                        potential=pot, T0=1.0, Tf=0.1, t0=0.0, tf=0.0,
                        seed=lane_seed(4321, i), first_step=-ntrans)
        o = orc.System(p)
        o.set_positions(xyz)
        assert o.cell_locate() == 0
        o.compute_forces()
        assert o.step(ntrans) == 0
        o.mark_origin()
        done = 0
        for k, target in enumerate(samples):
            assert o.step(target + 1 - done) == 0
            done = target + 1
            ob = o.observables(q)
            acc["fs"][k] += ob["fself_big"]
            acc["mb"][k] += ob["msd_big"]
            acc["ms"][k] += ob["msd_sml"]
            acc["eb"][k] += ob["epot_big"]
            acc["es"][k] += ob["epot_sml"]
            acc["z"][k] += ob["z"]
            acc["T"][k] = ob["temperature"]
            gsum[k] += orc.gder_normalize(p, nbins, rng, o.gder_counts(nbins, rng))
    res = sim / "results"
    t = dt * (1.0 + np.array(samples, float))
    fs = np.loadtxt(res / "fself.out")
    assert np.allclose(fs[:, 0], t, atol=2e-6)
    assert np.allclose(fs[:, 1], acc["fs"] / 2, atol=2e-6)
    msd = np.loadtxt(res / "msd_averaged.out")
    assert np.allclose(msd[:, 1:], np.stack([acc["mb"], acc["ms"]], 1) / 2, atol=2e-6)
    en = np.loadtxt(res / "energies_averaged.out")
    assert np.allclose(en[:, 1:3], np.stack([acc["eb"], acc["es"]], 1) / 2, atol=2e-6)
    assert np.allclose(en[:, 3], acc["T"], atol=2e-6)
    pr = np.loadtxt(res / "pressure.out")
    assert np.allclose(pr[:, 1], acc["z"] / 2, rtol=1e-6, atol=2e-6)
    for k in range(len(samples)):
        g = np.loadtxt(res / f"gder_{k}.out")
        assert g.shape == (nbins, 5)
        assert np.allclose(g[:, 1:], (gsum[k] / 2).T, atol=2e-6, rtol=1e-6)


AGING_PROTOCOL = """linear
1.0  0.1  0.0  0.0  1
1e-4
log
0.002  0.001 0.1
1  2  0
8 8
2
30 3.0
7.0
99
"""


def _sinc_stats(raw, org, nb, q):
    d = np.sqrt(((raw - org) ** 2).sum(1))
    qd = q * d[:nb]
    fs = np.where(qd > 0, np.sin(qd) / np.where(qd > 0, qd, 1.0), 1.0).mean()
    return fs, (d[:nb] ** 2).mean(), (d[nb:] ** 2).mean()


def test_aging_mode_one_run_origins_match_oracle(tmp_path):
    """--aging: F_s and the MSD from every dumped waiting time in one run."""
    cfg = tmp_path / "configs"
    sim = tmp_path / "sim"
    for k in range(3):
        (sim / "results" / f"decade{k}").mkdir(parents=True)
    r = subprocess.run([EXE, "--make-configs", "250", "250", "0.45", "1.4", "1.0", "2", "78",
                        str(cfg / "decade0"), "1.0", "50", "1e-4"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    for k in (1, 2):
        (cfg / f"decade{k}").mkdir()
    r = subprocess.run([EXE, "--aging"], input=AGING_PROTOCOL, cwd=sim, capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    dt, ntrans, q = 1e-4, 20, 7.0
    samples = orc.sampling_indexes(0.001, 0.1, dt, 10)
    origins = {}  # slot -> (t_w, [raw per config])
    got = {0: [], 1: []}
    systems = []
    for i in range(2):
        nb, ns, side, sb, ss, xyz = read_config(cfg / "decade0" / f"init_config_{i + 1}")
        p = make_params(nb, ns, side=side, sigma_big=sb, sigma_sml=ss, dt=dt,
                        potential=abi.BD_POT_WCA_MODIFIED, T0=1.0, Tf=0.1, t0=0.0, tf=0.0,
                        seed=lane_seed(99, i), first_step=-ntrans)
        o = orc.System(p)
        o.set_positions(xyz)
        assert o.cell_locate() == 0
        o.compute_forces()
        assert o.step(ntrans) == 0
        o.mark_origin()
        systems.append((o, nb))
    done, n_tw = 0, 1
    for k, target in enumerate(samples):
        if target + 1 < done:
            continue
        if target + 1 > done:
            for o, _ in systems:
                assert o.step(target + 1 - done) == 0
            done = target + 1
        time = dt * (1.0 + target)
        raws = [o.raw_positions() for o, _ in systems]
        for j, (tw, orgs) in origins.items():
            st = np.array([_sinc_stats(raws[i], orgs[i], systems[i][1], q) for i in range(2)])
            got[j].append((time - tw, *st.mean(0)))
        if k % 10 == 1 and n_tw <= 2:
            origins[n_tw - 1] = (time, raws)
            n_tw += 1
    for j in (0, 1):
        exp = np.array(got[j])
        fs = np.loadtxt(sim / "results" / f"decade{j + 1}" / "fself.out", ndmin=2)
        msd = np.loadtxt(sim / "results" / f"decade{j + 1}" / "msd_averaged.out", ndmin=2)
        assert fs.shape[0] == exp.shape[0] > 0
        assert np.allclose(fs[:, 0], exp[:, 0], atol=2e-6)
        assert np.allclose(fs[:, 1], exp[:, 1], atol=2e-6)
        assert np.allclose(msd[:, 1:], exp[:, 2:], atol=2e-6)
    # the decade dumps are still written (text-dump compatibility)
    assert (cfg / "decade1" / "init_config_1").exists()
    assert (cfg / "decade2" / "init_config_2").exists()


SYSTEM_PROTOCOL = """linear
1.0  1.0  0.0  0.0  1
1e-4
linear
0.002  0.004  2
8 8
../configs/
2
30 3.0
24 2.4
321
"""


def test_system_mode_force_contraction_matches_oracle(tmp_path):
    """--system (run-system.sh protocol): the contraction over every sample."""
    cfg = tmp_path / "configs"
    sim = tmp_path / "sim"
    (sim / "results").mkdir(parents=True)
    r = subprocess.run([EXE, "--make-configs", "200", "200", "0.42", "1.4", "1.0", "2", "79",
                        str(cfg), "1.0", "30", "1e-4"], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    r = subprocess.run([EXE, "--system"], input=SYSTEM_PROTOCOL, cwd=sim, capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    dt, ntrans, niter, nsamples, bins, rng = 1e-4, 20, 40, 2, 24, 2.4
    samples = list(range(0, niter, niter // nsamples))
    ata = np.zeros((3, bins, bins))
    atb = np.zeros((3, bins))
    systems = []
    for i in range(2):
        nb, ns, side, sb, ss, xyz = read_config(cfg / f"init_config_{i + 1}")
        p = make_params(nb, ns, side=side, sigma_big=sb, sigma_sml=ss, dt=dt,
                        potential=abi.BD_POT_WCA_MODIFIED, T0=1.0, Tf=1.0, t0=0.0, tf=0.0,
                        seed=lane_seed(321, i), first_step=-ntrans)
        o = orc.System(p)
        o.set_positions(xyz)
        assert o.cell_locate() == 0
        o.compute_forces()
        assert o.step(ntrans) == 0
        o.mark_origin()
        systems.append((o, side))
    done = 0
    for target in samples:
        for o, side in systems:
            assert o.step(target + 1 - done) == 0
        done = target + 1
        for o, side in systems:
            a, b = orc.cv3d_accumulate(bins, rng, o.positions(), o.forces()[0], side)
            ata += a
            atb += b
    exp = orc.cv3d_solve(ata, atb)
    got = np.loadtxt(sim / "results" / "force_contraction.out")
    assert got.shape == (bins, 2)
    assert np.allclose(got[:, 0], (np.arange(bins) + 0.5) * rng / bins, atol=2e-6)
    assert np.allclose(got[:, 1], exp, rtol=1e-6, atol=1e-8 * np.abs(exp).max())
    en = np.loadtxt(sim / "results" / "energies_averaged.out")
    assert en.shape == (nsamples, 4)
    assert (sim / "results" / "gder_1.out").exists()


def test_run_aging_script_end_to_end(tmp_path):
    """run_aging.sh (the run-fself.sh workflow, run-fself.sh:1-197) with a short
    parameter set: configurations made, one --aging run, and every decade's
    F_s and MSD written with the reference's layout.  The numbers themselves
    are checked against the oracle by the --aging test above; here F_s starts
    below 1 and stays in [-1, 1], the MSD grows from a positive first value."""
    script = os.path.join(ROOT, "bd_aging_amd", "run_aging.sh")
    params = tmp_path / "short.params"
    params.write_text("ngrain_big=200\nngrain_sml=200\nphi=0.45\nnumber_configs=2\n"
                      "eq_steps_configs=50\ntranstime_aging=0.002\nruntime_aging=0.05\n"
                      "initial_decade=1.0e-3\nnumber_initial_decades=2\ngr_nbins=50\n"
                      "gr_range=5.0\n")
    env = dict(os.environ, BD_AGING_PARAMS=str(params))
    r = subprocess.run(["bash", script, "t1"], cwd=tmp_path, capture_output=True, text=True,
                       timeout=300, env=env)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "Simulation finished successfully" in r.stdout
    run = [d for d in tmp_path.iterdir() if d.is_dir() and d.name.startswith("bigsml_200_200")]
    assert len(run) == 1
    res = run[0] / "simulation" / "results"
    for k in (1, 2):
        fs = np.loadtxt(res / f"decade{k}" / "fself.out", ndmin=2)
        msd = np.loadtxt(res / f"decade{k}" / "msd_averaged.out", ndmin=2)
        assert fs.shape[0] > 3 and msd.shape[0] == fs.shape[0]
        assert np.all(np.diff(fs[:, 0]) > 0)  # lags increase
        assert 0.0 < fs[0, 1] < 1.0 and np.all(np.abs(fs[:, 1]) <= 1.0)
        assert msd[0, 1] > 0.0 and msd[-1, 1] > msd[0, 1]
    assert (run[0] / "configs" / "decade0" / "init_config_2").exists()
    assert (run[0] / "configs" / "decade2" / "init_config_1").exists()

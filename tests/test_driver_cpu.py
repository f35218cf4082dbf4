"""Host-only behaviour of the bd_aging_self driver (no GPU call is reached)."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "bd_aging_amd", "bd_aging_self")

PROTOCOL = "linear\n1.0  0.1  0.0  0.0  1\n1e-4\nlinear\n0.002  0.006  3\n8 8\n2\n40 4.0\n7.0\n4321\n"


@pytest.fixture(scope="module")
def exe():
    if not os.path.exists(EXE):
        from bd_aging_amd import build
        build.build_driver(verbose=False)
    return EXE


def test_missing_init_config_exits_like_the_reference(exe, tmp_path):
    # main.cu:225-229: "Verify file path" and exit(-2)
    r = subprocess.run([exe], input=PROTOCOL, cwd=tmp_path, capture_output=True, text=True,
                       timeout=60)
    assert r.returncode == 254
    assert "Verify file path: ../configs/init_config_1" in r.stdout
    # the prompts of main.cu:93-148 are echoed in order
    prompts = ["temp_protocol ?", "temp_settings ?", "dt ?", "sample type (linear, log) ?",
               "ntags: big, sml ?", "n_configs ?", "nbins_gder, range_gder ?", "qmax ?",
               "idum ?"]
    pos = [r.stdout.index(p) for p in prompts]
    assert pos == sorted(pos)


def test_unknown_protocol_is_rejected(exe, tmp_path):
    r = subprocess.run([exe], input="cubic\n" + PROTOCOL.split("\n", 1)[1], cwd=tmp_path,
                       capture_output=True, text=True, timeout=60)
    assert r.returncode != 0


def test_unknown_potential_is_rejected_at_start(exe, tmp_path):
    env = dict(os.environ, BD_POTENTIAL="lennard_jones")
    r = subprocess.run([exe], input=PROTOCOL, cwd=tmp_path, capture_output=True, text=True,
                       timeout=60, env=env)
    assert r.returncode == 2
    assert "unknown BD_POTENTIAL 'lennard_jones'" in r.stderr
    assert "temp_protocol ?" not in r.stdout
    # the 49-50 names are accepted: the run goes on to the protocol prompts
    for name in ("wca", "wca_modified", "wca49_50", "wca49_50_ao"):
        env = dict(os.environ, BD_POTENTIAL=name)
        r = subprocess.run([exe], input=PROTOCOL, cwd=tmp_path, capture_output=True, text=True,
                           timeout=60, env=env)
        assert "temp_protocol ?" in r.stdout, name


def test_run_aging_script_parses():
    import subprocess
    script = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                          "bd_aging_amd", "run_aging.sh")
    r = subprocess.run(["bash", "-n", script], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr

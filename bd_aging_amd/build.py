"""In-tree build of libbd_hip.so (hipcc, gfx950) and of the bd_aging_self driver.

Flags: -ffp-contract=off / -fno-fast-math keep every a*b+c two roundings so
the kernels evaluate the same IEEE operations as the FP64 oracle;
-packed-fp32-ops is disabled for all device code.
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

COMMON = ["-O3", "-std=c++17", "-ffp-contract=off", "-fno-fast-math",
          f"--offload-arch={ARCH}",
          "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]

LIB_SOURCES = ["bd_engine.hip", "bd_cv3d.hip", "bd_mlp.hip"]


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, cwd=ROOT, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError(f"build failed: {' '.join(cmd)}")


def build_lib(verbose=True):
    """Each source compiled to an object concurrently, then one link."""
    out = os.path.join(HERE, "libbd_hip.so")
    objdir = os.path.join(HERE, "_build")
    os.makedirs(objdir, exist_ok=True)
    objs = [os.path.join(objdir, os.path.splitext(s)[0] + ".o") for s in LIB_SOURCES]
    cmds = [[HIPCC] + COMMON + ["-fPIC", "-c", "-o", o, os.path.join(CSRC, s)]
            for s, o in zip(LIB_SOURCES, objs)]
    with ThreadPoolExecutor(max_workers=len(cmds)) as ex:
        for f in [ex.submit(_run, c, verbose) for c in cmds]:
            f.result()
    _run([HIPCC, f"--offload-arch={ARCH}", "-fPIC", "-shared", "-o", out] + objs, verbose)
    return out


def build_driver(verbose=True):
    """bd_aging_self: host C++ on the C ABI (no device code)."""
    out = os.path.join(HERE, "bd_aging_self")
    _run(["g++", "-O2", "-std=c++17", "-Wall", "-o", out,
          os.path.join(CSRC, "bd_aging_self.cpp"), "-L" + HERE, "-lbd_hip",
          "-L/opt/rocm/lib", "-Wl,-rpath,$ORIGIN", "-Wl,-rpath-link,/opt/rocm/lib"], verbose)
    return out


def build_all(verbose=True):
    build_lib(verbose)
    build_driver(verbose)


if __name__ == "__main__":
    build_all()

"""Write tests/golden/rannew64_ref.npz from the reference rannew64.

Runs oracle/_ref/rannew64_ref (the reference rannew64.cu compiled by
`make -C oracle ref`) once per seed, in a fresh process each time because the
reference keeps its state in statics.  Stores the returned doubles as their
IEEE-754 bit patterns.

usage: python tests/golden/make_rannew64_vectors.py
"""
import os
import subprocess
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
REF = os.path.join(ROOT, "oracle", "_ref", "rannew64_ref")
COUNT = 1200  # more than four turns of the 250-word ring


def splitmix_lane_seed(seed, lane):
    m = (1 << 64) - 1
    z = (seed + (lane + 1) * 0x9e3779b97f4a7c15) & m
    z = ((z ^ (z >> 30)) * 0xbf58476d1ce4e5b9) & m
    z = ((z ^ (z >> 27)) * 0x94d049bb133111eb) & m
    z ^= z >> 31
    return z >> 1


SEEDS = [0, 1, 7, -7, 12345, 2147483647, -123456789,
         splitmix_lane_seed(12345, 0), splitmix_lane_seed(12345, 1),
         splitmix_lane_seed(2 ** 40, 3)]


def main():
    if not os.path.exists(REF):
        subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "ref"], check=True)
    rows = []
    for s in SEEDS:
        out = subprocess.run([REF, str(s), str(COUNT)], check=True, capture_output=True,
                             text=True).stdout.split()
        vals = np.array([float.fromhex(v) for v in out], dtype=np.float64)
        assert len(vals) == COUNT
        rows.append(vals.view(np.uint64))
    np.savez(os.path.join(HERE, "rannew64_ref.npz"),
             seeds=np.array(SEEDS, dtype=np.int64), bits=np.stack(rows))
    print("wrote", len(SEEDS), "seeds x", COUNT)


if __name__ == "__main__":
    sys.exit(main())

"""Writes the golden fixtures: seeded states of the drawn mechanisms with the FP64 oracle's outputs.

  python tests/golden/make_golden.py

The fixtures freeze the oracle (a regression pin: later changes to oracle/chem.c must reproduce them bit for
bit) and give the GPU tests known vectors that need no oracle build. Inputs are drawn with the same generator
as the benchmarks (combustion_amd.synth)."""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))

from combustion_amd import synth  # noqa: E402
from helpers import setup  # noqa: E402

CASES = {"h2o2-10": (64, 101), "gri-53": (32, 153), "large-160": (4, 1600)}


def main():
    for name, (n, seed) in CASES.items():
        M, mech, o = setup(name)
        T, P, Y = synth.config_states(name, n, seed)
        out, floor = o.batch(T, P, Y)
        np.savez_compressed(os.path.join(HERE, f"{name}.npz"), T=T, P=P, Y=Y, out=out, floor=floor, seed=seed)
        print(name, out.shape)


if __name__ == "__main__":
    main()

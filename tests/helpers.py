"""Shared test setup: mechanisms, oracle and error metric."""
import functools

import numpy as np

import oracle
from combustion_amd import Mechanism, synth
from oracle import mechanism as om


@functools.lru_cache(maxsize=None)
def setup(name, D=4):
    """(product Mechanism, oracle mechanism dict, Oracle with the product's fits) for a named configuration."""
    path = synth.mechanism_path(name)
    M = Mechanism(path)
    mech = om.load(path)
    fits = M.transport_fits(D, 1000.0)
    T0, _, _ = M.transport_scales()
    return M, mech, oracle.Oracle(mech, fits, T0)


def relative_error(x, ref, floor):
    """|x - r| / max(|r|, floor): relative error with a per-output floor (the term scale for ω)."""
    return np.abs(x - ref) / np.maximum(np.abs(ref), floor)

"""Shared test setup: mechanisms, oracle and error metric."""
import functools

import numpy as np

import oracle
from combustion_amd import Mechanism, synth
from oracle import mechanism as om
from combustion_amd.synth import models_variant, plog_variant, sri_variant  # noqa: F401 (re-exported for the tests)


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


def state_floor(scale, Ka, ratio=1e-30):
    """Mixed-form floor for ω: max(term scale s_k, ratio · max_j s_j) per state. The mixed form accumulates
    ν·q in FP32 under one exponent per state, so rates more than ~1e38 below the state's largest term cannot
    be represented; the stated floor puts them below 1e-30 of the state's largest term scale."""
    import numpy as np
    f = scale.copy()
    if Ka:
        f[:Ka] = np.maximum(scale[:Ka], ratio * scale[:Ka].max(axis=0, keepdims=True))
    return f


@functools.lru_cache(maxsize=None)
def setup_text(text, D=4):
    """setup() for a mechanism given as YAML text."""
    M = Mechanism(text=text)
    mech = om.load(text=text)
    fits = M.transport_fits(D, 1000.0)
    T0, _, _ = M.transport_scales()
    return M, mech, oracle.Oracle(mech, fits, T0)

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


def state_floor(scale, Ka, ratio=1e-30):
    """Mixed-form floor for ω: max(term scale s_k, ratio · max_j s_j) per state. The mixed form accumulates
    ν·q in FP32 under one exponent per state, so rates more than ~1e38 below the state's largest term cannot
    be represented; the stated floor puts them below 1e-30 of the state's largest term scale."""
    import numpy as np
    f = scale.copy()
    if Ka:
        f[:Ka] = np.maximum(scale[:Ka], ratio * scale[:Ka].max(axis=0, keepdims=True))
    return f


def sri_variant(name, seed=5, params=None):
    """YAML text of a named synthetic mechanism with every Troe block replaced by an SRI block: drawn
    (a ~ U[0.5, 2], b ~ U[100, 1000] K, c ~ U[500, 3000] K, d ~ U[0.8, 1.2], e ~ U[-0.5, 0.5]; the first one with
    c = 0), or the given (a, b, c, d, e) for all."""
    import re
    text = open(synth.mechanism_path(name)).read()
    rng = np.random.default_rng(seed)
    count = [0]

    def block(_):
        if params is not None:
            a, b, c, d, e = params
        else:
            a, b, c, d, e = rng.uniform(0.5, 2), rng.uniform(100, 1000), rng.uniform(500, 3000), rng.uniform(0.8, 1.2), rng.uniform(-0.5, 0.5)
            if count[0] == 0:
                c = 0.0
        count[0] += 1
        return f"SRI: {{A: {a!r}, B: {b!r}, C: {c!r}, D: {d!r}, E: {e!r}}}"
    return re.sub(r"Troe: \{[^}]*\}", block, text)


@functools.lru_cache(maxsize=None)
def setup_text(text, D=4):
    """setup() for a mechanism given as YAML text."""
    M = Mechanism(text=text)
    mech = om.load(text=text)
    fits = M.transport_fits(D, 1000.0)
    T0, _, _ = M.transport_scales()
    return M, mech, oracle.Oracle(mech, fits, T0)


def plog_variant(name, every=3, seed=6):
    """YAML text of a named synthetic mechanism with every `every`-th elementary reaction turned into a PLOG table:
    entries at 2, 8 and 30 atm (its rate constant scaled by drawn factors and exponents), plus a second entry at
    30 atm that the table sums."""
    import re
    text = open(synth.mechanism_path(name)).read()
    rng = np.random.default_rng(seed)
    count = [0]

    def table(m):
        count[0] += 1
        if count[0] % every:
            return m.group(0)
        A, b, Ea = float(m.group(2)), float(m.group(3)), float(m.group(4))
        rows = []
        for P in ("2.0 atm", "8.0 atm", "30.0 atm", "30.0 atm"):
            f, db = rng.uniform(0.2, 5.0), rng.uniform(-0.3, 0.3)
            rows.append(f"  - {{P: {P}, A: {A * f!r}, b: {b + db!r}, Ea: {Ea * rng.uniform(0.8, 1.2)!r}}}")
        return f"{m.group(1)}\n  type: pressure-dependent-Arrhenius\n  rate-constants:\n" + "\n".join(rows)
    return re.sub(r"(- equation: [^\n]*<=>[^\n(]*)\n  rate-constant: \{A: ([^,]*), b: ([^,]*), Ea: ([^}]*)\}", table, text)

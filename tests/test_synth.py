"""The seeded mechanism generator: element balance, active prefix, determinism, parameter ranges."""
import numpy as np
import pytest

from combustion_amd import synth
from oracle import mechanism as om


@pytest.mark.parametrize("name", ["h2o2-10", "gri-53", "large-160"])
def test_drawn_mechanism_properties(name):
    K, R, tb, troe, seed = synth.CONFIGS[name]
    m = om.load(synth.mechanism_path(name))
    assert (m["K"], m["R"], m["Ka"]) == (K, R, K - 1)
    assert (m["model"] == 2).sum() == tb and (m["model"] == 4).sum() == troe
    comp = m["composition"]
    nu = m["products"] - m["reactants"]
    np.testing.assert_array_equal(nu @ comp, 0)  # every reaction conserves every element
    assert np.all(np.abs(nu) <= 2) and set(np.unique(-nu.sum(axis=1))) <= {0, 1, -1, -2}
    assert m["names"][-1] == "N2" and not np.any(nu[:, -1])
    A_cgs = m["rate"][:, 0] / 1e-6 ** (m["reactants"].sum(axis=1) + (m["model"] == 2) - 1)
    assert np.all((np.log(A_cgs) > 20 - 1e-9) & (np.log(A_cgs) < 40 + 1e-9))
    assert np.all((m["rate"][:, 1] >= -1.5) & (m["rate"][:, 1] <= 2.5))
    assert np.all((m["rate"][:, 2] >= 0) & (m["rate"][:, 2] <= 25000))


def test_generation_is_deterministic():
    a = synth.config("h2o2-10").yaml()
    b = synth.config("h2o2-10").yaml()
    assert a == b
    with open(synth.mechanism_path("h2o2-10")) as f:
        assert f.read() == a


def test_nasa_pieces_continuous_at_split():
    m = om.load(synth.mechanism_path("gri-53"))
    T = 1000.0
    for lo, hi in m["nasa"]:
        h = lambda a: a[0] + a[1] * T / 2 + a[2] * T ** 2 / 3 + a[3] * T ** 3 / 4 + a[4] * T ** 4 / 5 + a[5] / T
        s = lambda a: a[0] * np.log(T) + a[1] * T + a[2] * T ** 2 / 2 + a[3] * T ** 3 / 3 + a[4] * T ** 4 / 4 + a[6]
        assert h(lo) == pytest.approx(h(hi), rel=1e-9) and s(lo) == pytest.approx(s(hi), rel=1e-9)

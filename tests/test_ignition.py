"""`.ignition` state files as cantera.rs:13-29 reads them: the first state's pressure, temperature and mole
proportions X (every species, non-negative, any normalisation, no volume) become the reference's State with V = 1 m³
and amounts n_k = (P/R) V / T · x_k / Σx in the mechanism's species order, which then feeds reaction::rates."""
import os

import numpy as np
import pytest

import oracle
from combustion_amd import Mechanism
from combustion_amd.native import Error
from oracle import mechanism as om

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")
R = 1.380649e-23 * 6.02214076e23
FILE_X = {"N2": 3.76, "O2": 1.0, "H2": 2.0, "OH": 0.0, "H2O": 0.05, "AR": 0.5}  # the first state of the fixture


def mechanism():
    return Mechanism(os.path.join(GOLDEN, "handwritten.yaml"))


def test_first_state_in_species_order():
    M = mechanism()
    T, P, n = M.ignition_state(os.path.join(GOLDEN, "handwritten.ignition"))
    assert (T, P) == (1150.0, 303975.0)
    x = np.array([FILE_X[name] for name in M.species_names])  # active species first (H2, O2, OH, H2O), then AR, N2
    assert M.species_names[-2:] == ["AR", "N2"]
    np.testing.assert_allclose(n, P / R / T * x / x.sum(), rtol=1e-15)
    assert n.sum() == pytest.approx(P / (R * T), rel=1e-15)  # the ideal-gas amount in 1 m³


@pytest.mark.parametrize("edit,what", [
    (lambda t: t.replace(", AR: 0.5}", "}", 1), "no entry for species AR"),
    (lambda t: t.replace("OH: 0.0", "OH: -0.1", 1), "non-negative"),
    (lambda t: t.replace("  temperature: 1150.0\n", "  temperature: 1150.0\n  volume: 2.0\n", 1), "no volume"),
    (lambda t: t.replace("- pressure: 303975.0\n", "- {}\n- pressure: 303975.0\n", 1), "pressure, temperature and X"),
    (lambda t: "pressure: 1.0\n", "sequence of states"),
])
def test_rejected_states_report_parse_errors(edit, what):
    text = edit(open(os.path.join(GOLDEN, "handwritten.ignition")).read())
    with pytest.raises(Error) as e:
        mechanism().ignition_state(text=text)
    assert e.value.kind == "parse" and what in str(e.value)


def test_ignition_state_feeds_the_rates_program():
    """cantera.rs:56-57 evaluates reaction::rates at the file's state: the product's program on the loaded amounts
    against the oracle at the same state (T, P, Y with Y_k = n_k W_k / Σ n_j W_j)."""
    M = mechanism()
    T, P, n = M.ignition_state(os.path.join(GOLDEN, "handwritten.ignition"))
    m = om.load(os.path.join(GOLDEN, "handwritten.yaml"))
    Y = n * m["W"] / np.dot(n, m["W"])
    omega, _ = oracle.Oracle(m).rhs(T, P, Y)
    _, scale = oracle.Oracle(m).batch(np.array([T]), np.array([P]), Y[:, None], transport=False)
    out = oracle.interpret(M.rates().linear(), np.concatenate([[P / R, R / P, T, 1.0], n[:-1]])[:, None])[:, 0]
    assert np.all(np.abs(out[2:] - omega) <= 1e-12 * np.maximum(np.abs(omega), scale[:, 0]))
    assert np.any(omega != 0)

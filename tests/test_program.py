"""The product's programs (reaction.rs, transport.rs builders restated in C++) evaluated by the FP64 interpreter
of their linear form (oracle/interpret.c) against the closed-form oracle."""
import numpy as np
import pytest

import oracle
from combustion_amd import synth
from helpers import relative_error, setup

R = 1.380649e-23 * 6.02214076e23


@pytest.mark.parametrize("name,n", [("h2o2-10", 4096), ("gri-53", 512)])
def test_chemistry_program_matches_oracle(name, n):
    M, mech, o = setup(name)
    T, P, Y = synth.config_states(name, n, 21)
    ref, floor = o.batch(T, P, Y, threads=4)
    got = oracle.interpret(M.chemistry().linear(), np.vstack([T, P, Y]))
    assert relative_error(got, ref, floor).max() <= 1e-12


def test_temperature_at_the_nasa_split():
    """T == 1000 K selects the low piece in the program (reaction.rs:106 <=) and in the oracle (lib.rs:15)."""
    M, mech, o = setup("h2o2-10")
    T, P, Y = synth.config_states("h2o2-10", 16, 5)
    T[:] = 1000.0
    ref, floor = o.batch(T, P, Y)
    got = oracle.interpret(M.chemistry().linear(), np.vstack([T, P, Y]))
    assert relative_error(got, ref, floor).max() <= 1e-12


def test_zero_and_negative_mass_fractions():
    """max(0, n_k) (reaction.rs:366) and the bulk closure (reaction.rs:372) on degenerate compositions."""
    M, mech, o = setup("h2o2-10")
    T, P, Y = synth.config_states("h2o2-10", 8, 6)
    Y[0, :] = 0.0
    Y[1, :4] = -1e-9
    ref, floor = o.batch(T, P, Y, transport=False)
    p = M.chemistry(1)
    got = oracle.interpret(p.linear(), np.vstack([T, P, Y]))
    assert relative_error(got, ref, floor).max() <= 1e-12


@pytest.mark.parametrize("name", ["h2o2-10", "gri-53"])
def test_reference_rates_function(name):
    """reaction.rs:356 rates with the reference's inputs [P/R, R/P, T, V, n_0..n_{K-2}] (V = 1 m³ so n = C):
    outputs [T dT/dt... as dT/dt, dV/dt, ω]."""
    M, mech, o = setup(name)
    T, P, Y = synth.config_states(name, 32, 31)
    prog = M.rates()
    lin = prog.linear()
    for s in range(32):
        n = Y[:, s] / mech["W"]
        C = P[s] / R / T[s] * n / n.sum()
        inp = np.concatenate([[P[s] / R, R / P[s], T[s], 1.0], C[:-1]])[:, None]
        out = oracle.interpret(lin, inp)[:, 0]
        omega, dTdt = o.rhs(T[s], P[s], Y[:, s])
        _, scale = o.batch(T[s:s + 1], P[s:s + 1], Y[:, s:s + 1], transport=False)
        assert np.all(np.abs(out[2:] - omega) <= 1e-12 * np.maximum(np.abs(omega), scale[:, 0]))
        assert out[0] == pytest.approx(dTdt, rel=1e-9, abs=1e-9 * np.abs(dTdt) + 1e-300)


def test_reactions_rates_function():
    """reaction.rs:329 reactions_rates: one rate of progress per reaction."""
    M, mech, o = setup("h2o2-10")
    T, P, Y = synth.config_states("h2o2-10", 8, 41)
    lin = M.reactions_rates().linear()
    for s in range(8):
        n = Y[:, s] / mech["W"]
        C = P[s] / R / T[s] * n / n.sum()
        out = oracle.interpret(lin, np.concatenate([[P[s] / R, T[s], 1.0], C[:-1]])[:, None])[:, 0]
        q = o.reaction_rates(T[s], P[s], Y[:, s])
        # floor: the forward and reverse magnitudes of each reaction
        assert np.all(np.abs(out - q) <= 1e-12 * np.maximum(np.abs(q), np.abs(out)) + 1e-300) or \
            np.max(np.abs(out - q) / np.maximum(np.abs(q), 1e-300)) < 1e-9


def test_transport_function_matches_scaled_oracle():
    """transport.rs:212 properties::<4> with inputs [Σx, T/T0, x_0..x_{K-2}] (x = concentrations)."""
    M, mech, o = setup("h2o2-10")
    T0, eta0, lam0 = M.transport_scales()
    prog = M.transport(4, T0, eta0, lam0)
    T, P, Y = synth.config_states("h2o2-10", 16, 51)
    ref, _ = o.batch(T, P, Y, rates=False)
    lin = prog.linear()
    for s in range(16):
        n = Y[:, s] / mech["W"]
        C = P[s] / R / T[s] * n / n.sum()
        out = oracle.interpret(lin, np.concatenate([[C.sum(), T[s] / T0], C[:-1]])[:, None])[:, 0]
        rho = np.dot(mech["W"], C)
        assert out[0] * lam0 == pytest.approx(ref[0, s], rel=1e-12)
        assert out[1] * eta0 == pytest.approx(ref[1, s], rel=1e-12)
        np.testing.assert_allclose(out[2:] * eta0 / rho, ref[2:, s], rtol=1e-12)


def test_interpreter_reports_nan_like_the_reference():
    """ast/interpret.rs:110,122 assert on NaN; the interpreter reports the state and instruction."""
    M, mech, o = setup("h2o2-10")
    T, P, Y = synth.config_states("h2o2-10", 4, 1)
    T[2] = np.nan
    with pytest.raises(FloatingPointError, match="state 2"):
        oracle.interpret(M.chemistry().linear(), np.vstack([T, P, Y]))


def test_program_round_trips_through_the_linear_form():
    """cb_program_from_linear(cb_program_linear(p)) emits the same straight-line HIP and interprets the same."""
    import oracle
    from combustion_amd.native import Program
    M, mech, o = setup("h2o2-10")
    p = M.chemistry()
    lin = p.linear()
    q = Program.from_linear(lin)
    assert (q.n_inputs, q.n_outputs, q.n_instructions) == (p.n_inputs, p.n_outputs, p.n_instructions)
    assert q.hip_source("straight") == p.hip_source("straight")
    T, P, Y = synth.config_states("h2o2-10", 3, 1)
    ins = np.vstack([T, P, Y])
    assert np.array_equal(oracle.interpret(q.linear(), ins), oracle.interpret(lin, ins))


def test_program_from_linear_rejects_bad_programs():
    import pytest
    from combustion_amd.native import Error, Program
    bad = {"op": [4], "dst": [2], "a": [0], "b": [5], "c": [0], "k": [0.0], "outputs": [2], "n_inputs": 2}
    with pytest.raises(Error, match="operand b"):
        Program.from_linear(bad)
    bad = {"op": [0], "dst": [3], "a": [0], "b": [0], "c": [0], "k": [1.0], "outputs": [2], "n_inputs": 2}
    with pytest.raises(Error, match="in order"):
        Program.from_linear(bad)

"""The CPU path (cb_cpu_assemble / cb_cpu_call): the program as C compiled by the host compiler and called once per
state, the counterpart of ir/lib.rs:225-263 and the loop of device.rs:28-48. FP64 with libm exp/log and no
contraction: within 1e-12 of the oracle like the FP64 GPU forms, the same bits on any number of threads."""
import numpy as np
import pytest

from combustion_amd import synth
from helpers import relative_error, setup

R = 1.380649e-23 * 6.02214076e23


@pytest.mark.parametrize("name,n", [("h2o2-10", 4096), ("gri-53", 1024)])
def test_host_program_matches_oracle(name, n):
    M, mech, o = setup(name)
    T, P, Y = synth.config_states(name, n, 91)
    ref, floor = o.batch(T, P, Y, threads=8)
    got = np.stack(M.chemistry().assemble_cpu()([T, P, *Y], threads=4))
    assert relative_error(got, ref, floor).max() <= 1e-12


def test_threads_give_the_same_bits():
    M, mech, o = setup("h2o2-10")
    T, P, Y = synth.config_states("h2o2-10", 1000, 92)
    mod = M.chemistry().assemble_cpu()
    one = np.stack(mod([T, P, *Y], threads=1))
    assert np.array_equal(one, np.stack(mod([T, P, *Y], threads=7)))
    assert np.array_equal(one[:, :1], np.stack(mod([T[:1], P[:1], *Y[:, :1]], threads=3)))


def test_reference_rates_function_on_the_host():
    """reaction.rs:356 rates with the reference's own inputs [P/R, R/P, T, V, n_0..n_{K-2}]: ω against the oracle."""
    M, mech, o = setup("h2o2-10")
    T, P, Y = synth.config_states("h2o2-10", 64, 93)
    n = Y / mech["W"][:, None]
    C = (P / R / T) * n / n.sum(axis=0)
    out = np.stack(M.rates().assemble_cpu()([P / R, R / P, T, np.ones_like(T), *C[:-1]]))
    for s in range(64):
        omega, dT, x = o.rhs_full(T[s], P[s], Y[:, s])
        assert np.all(np.abs(out[2:, s] - omega) <= 1e-12 * np.maximum(np.abs(omega), x["omega_scale"]))
        assert abs(out[0, s] - dT) <= 1e-12 * x["dTdt_scale"]


def test_bad_calls_return_status():
    from combustion_amd.native import Error
    M, mech, o = setup("h2o2-10")
    mod = M.chemistry().assemble_cpu()
    with pytest.raises(ValueError):
        mod([np.zeros(3)])
    T, P, Y = synth.config_states("h2o2-10", 4, 1)
    with pytest.raises(Error):
        mod([T, P, *Y], threads=0)

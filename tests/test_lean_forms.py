"""The program shapes that the one-lane-per-state FP64 kernels compile (CB_LEAN), run through the FP64 interpreter
on the CPU against the oracle: the restructured algebra (rate-of-progress products started from each reaction's
rate constant; Wilke's a_kj²(1 + ṽ_k/ṽ_j)² terms with ṽ_k = v_k W_k^(-1/4); the binary-diffusion fits normalized
by their constant term; four diffusion reciprocals from one reciprocal of their product) computes the reference's
quantities to the FP64 parity bar, 1e-12 (ω relative to the term scale, transport relative). The device-only
building blocks (polynomial exponential, one Newton step after v_rcp_f64) are checked on the GPU."""
import numpy as np
import oracle
import pytest

from combustion_amd import synth
from combustion_amd.native import LEAN, RATES, TRANSPORT
from helpers import relative_error, setup


@pytest.mark.parametrize("name", ["h2o2-10", "gri-53"])
def test_lean_program_matches_oracle(name):
    M, mech, o = setup(name)
    T, P, Y = synth.config_states(name, 48, 17)
    ref, floor = o.batch(T, P, Y)
    ins = np.vstack([T, P, Y])
    lean = M.chemistry(RATES | TRANSPORT | LEAN)
    plain = M.chemistry(RATES | TRANSPORT)
    got = oracle.interpret(lean.linear(), ins)
    base = oracle.interpret(plain.linear(), ins)
    Ka = M.active
    err = relative_error(got, ref, floor)
    print(name, "lean ω", err[:Ka].max(), "transport", err[Ka:].max(), "plain", relative_error(base, ref, floor).max(), flush=True)
    assert err.max() < 1e-12
    # a different program, not the plain one under another name
    assert lean.linear()["op"].size != plain.linear()["op"].size


def test_lean_flag_alone_is_rejected():
    from combustion_amd.native import Error
    M, mech, o = setup("h2o2-10")
    with pytest.raises(Error, match="flags must combine"):
        M.chemistry(LEAN)


@pytest.mark.parametrize("name", ["h2o2-10", "gri-53"])
def test_lean_program_at_extreme_states(name):
    """The states of test_gpu_extremes.py (200-6000 K, 1 Pa-100 MPa, zeros and 1e-300 traces) through the
    interpreter: the lean shapes keep the 1e-12 bar there too, above the subnormal floor that test states."""
    from test_gpu_extremes import extreme_states
    M, mech, o = setup(name)
    T, P, Y = extreme_states(M.K, 3)
    ref, floor = o.batch(T, P, Y)
    floor = np.abs(floor)  # a fit extrapolated past 3000 K can make the D floor 1/S_k negative
    floor[:M.active] = np.maximum(floor[:M.active], 1e-290)
    got = oracle.interpret(M.chemistry(RATES | TRANSPORT | LEAN).linear(), np.vstack([T, P, Y]))
    with np.errstate(invalid="ignore"):
        err = np.where(got == ref, 0.0, relative_error(got, ref, floor))
    assert err.max() < 1e-12

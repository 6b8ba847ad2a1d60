"""The FP64 and fast forms at the edges of the state space, against the oracle: temperatures from 200 to 6000 K
(where |G/RT| runs past the clamps of exp(±G/RT) and the exponential's reduced range is exercised at both ends),
pressures from 1 Pa to 100 MPa, and compositions with exact zeros and with one species and traces at 1e-300. The bar
is the same as on the configuration's own states: 1e-12 for FP64 (ω relative to the term scale, transport
relative), the stated error bound for the fast form."""
import numpy as np
import pytest

from helpers import relative_error, setup, state_floor
from test_gpu_parity import run


def extreme_states(K, seed):
    rng = np.random.default_rng(seed)
    T = np.array([200., 250., 300., 400., 1000., 2500., 4000., 5000., 6000.])
    P = np.array([1., 1e3, 101325., 1e6, 1e8])
    Ts, Ps = np.meshgrid(T, P)
    T, P = Ts.ravel(), Ps.ravel()
    n = T.size
    Y = rng.dirichlet(np.full(K, 0.3), n).T.copy()
    Y[0, 0::4] = 0.0                       # every fourth state: the first species absent
    Y[:, 1::5] = 1e-300                    # every fifth: traces everywhere but the bulk species (a state of the
    Y[-1, 1::5] = 1.0                      # bulk alone is undefined in the reference: its D is 0/0)
    Y[:, 2::7] = 1e-300                    # every seventh: traces everywhere but one reactive species
    Y[K // 2, 2::7] = 1.0
    Y /= Y.sum(axis=0, keepdims=True)
    return T, P, Y


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["h2o2-10", "gri-53"])
# The mixed (MFMA) form is not among them: its log-domain contraction carries log2 C in FP32, whose rounding at the
# 1e-300 traces here (|log2 C| ~ 1000, steps of 2^-14) is ~2e-5 per factor, above its 1e-5 bound; exact zeros it
# handles (test_gpu_parity.py::test_mixed_forms_on_states_with_absent_species).
# fp64+tiled+tiled-rates: the large-mechanism kernels (table-driven rates, MFMA transport in FP64) at these sizes
@pytest.mark.parametrize("form", ["fp64", "fast", "fp64+tiled+tiled-rates"])
def test_extreme_states_match_oracle(name, form, gpu):
    M, mech, o = setup(name)
    T, P, Y = extreme_states(M.K, 3)
    ref, floor = o.batch(T, P, Y)
    got = run(M.chemistry(3), form, T, P, Y, gpu)
    Ka = M.active
    exact = form.startswith("fp64")
    # |floor|: beyond the fits' 300-3000 K range a binary-diffusion fit can turn negative, and with it the oracle's
    # D floor 1/S_k
    f = np.abs(floor if exact else state_floor(floor, Ka))
    # Below 1e-290 mol/m³/s the terms of ω are subnormal numbers, which carry fewer than the 40 significant bits of a
    # 1e-12 comparison in any evaluation order (the oracle's own included): the trace states at 200 K reach 5e-315.
    f[:Ka] = np.maximum(f[:Ka], 1e-290)
    tol_w, tol_t = (1e-12, 1e-12) if exact else (1e-5, 1e-3)
    with np.errstate(invalid="ignore"):
        err = np.where(got == ref, 0.0, relative_error(got, ref, f))  # a rate whose terms all underflow is 0 in both
    print(name, form, "ω", err[:Ka].max(), "transport", err[Ka:].max(), flush=True)
    bad = np.argwhere(~np.isfinite(got) | (err > (np.where(np.arange(len(got))[:, None] < Ka, tol_w, tol_t))))
    for o_, s_ in bad[:12]:
        print("  output", o_, "state", s_, "T", T[s_], "P", P[s_], "got", got[o_, s_], "ref", ref[o_, s_], "floor", f[o_, s_],
              "min Y", Y[:, s_].min(), flush=True)
    assert np.isfinite(got).all()
    assert err[:Ka].max() < tol_w and err[Ka:].max() < tol_t

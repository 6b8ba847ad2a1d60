"""The batched Rosenbrock reactor integrator (combustion_amd/reactor.py) against scipy's BDF on the same
right-hand side (the oracle's reaction rates and energy balance, oracle/chem.c), on the CPU."""
import numpy as np
import pytest
import torch
from scipy.integrate import solve_ivp

from combustion_amd import reactor, synth
from helpers import setup


def oracle_rhs(o):
    def rhs(T, P, Y):
        Tn, Pn, Yn = T.numpy(), P.numpy(), Y.numpy()
        omega, dT = [], []
        for i in range(len(Tn)):
            w, d = o.rhs(Tn[i], Pn[i], np.ascontiguousarray(Yn[:, i]))
            omega.append(w)
            dT.append(d)
        return torch.from_numpy(np.array(omega).T.copy()), torch.from_numpy(np.array(dT))
    return rhs


def test_rosenbrock_matches_scipy_bdf():
    M, mech, o = setup("h2o2-10")
    reactors = reactor.Reactors(oracle_rhs(o), mech["W"], mech["Ka"], rtol=1e-7, atol_Y=1e-14)
    y0, P = reactor.initial_states(M, 3, seed=5, T_range=(1200.0, 1500.0))
    # the drawn kinetics are fast (k up to ~1e20 /s with A T^β at β = 2.5): the first 330 K of relaxation take ~1e-19 s
    t_end = 5e-20
    y, t, steps = reactors.integrate(torch.from_numpy(y0), torch.from_numpy(P), t_end, h0=1e-24, max_steps=3000)
    assert np.all(t.numpy() >= t_end * (1 - 1e-12)), steps
    f = lambda P: (lambda _, z: reactors.f(torch.from_numpy(z)[None], torch.tensor([P]))[0].numpy())
    for i in range(3):
        ref = solve_ivp(f(P[i]), (0, t_end), y0[i], method="BDF", rtol=1e-10, atol=1e-14)
        assert ref.success
        yT = ref.y[:, -1]
        assert abs(y[i, 0].item() - yT[0]) < 1e-3 * yT[0]  # temperature
        assert np.max(np.abs(y[i, 1:].numpy() - yT[1:])) < 1e-4  # mass fractions


def test_reactor_conserves_mass_and_enthalpy_direction():
    """Σ Y stays 1 (element-balanced reactions conserve mass: Σ W_k ω_k = 0) and the bulk species is inert."""
    M, mech, o = setup("h2o2-10")
    reactors = reactor.Reactors(oracle_rhs(o), mech["W"], mech["Ka"])
    y0, P = reactor.initial_states(M, 2, seed=9)
    y, t, steps = reactors.integrate(torch.from_numpy(y0), torch.from_numpy(P), 1e-18, h0=1e-24, max_steps=300)
    Y = y[:, 1:].numpy()
    assert np.allclose(Y.sum(axis=1), 1.0, atol=1e-9)
    assert np.allclose(Y[:, -1], y0[:, -1], rtol=0, atol=1e-15)


def test_integrator_options_are_checked_before_any_compilation():
    """cb_integrator_create rejects flags outside cb_integrator_flags and states beyond one wave (K+1 > 64) with its
    own errors, before it compiles or touches a device (the out == NULL call only fills the code-object cache)."""
    from combustion_amd import Mechanism, native
    M, mech, o = setup("h2o2-10")
    for flags in (0x3, 0x40, 0x100):
        with pytest.raises(native.Error) as e:
            native._check(native.lib().cb_integrator_create(M._h, 1e-6, 1e-6, 1e-12, flags, None))
        assert e.value.kind == "argument"
    assert native.Integrator.flags("differences", True) == 0x12
    assert native.Integrator.flags("analytic", method="ode23s") == 0x21
    assert native.Integrator.flags() == 0  # automatic Jacobian, packed waves, RODAS4
    big = Mechanism(synth.mechanism_path("large-160"))
    with pytest.raises(native.Error) as e:
        native.Integrator.compile(big)
    assert e.value.kind == "unsupported"

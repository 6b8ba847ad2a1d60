"""Writes reactor_handwritten.npz: constant-pressure reactor trajectories of the hand-written H2/O2 mechanism
(handwritten.yaml) through ignition, integrated by scipy's BDF (solve_ivp, rtol 1e-10, atol 1e-14 on mass
fractions and 1e-10 K) on the FP64 oracle's right-hand side (dT/dt of the energy balance, dY_k/dt = W_k ω_k / ρ):
the fixtures of the GPU integrator's test. Initial states: T0 ~ U[1300, 1600] K, P ~ U[1, 5] atm, H2/O2/N2/AR
mixtures (φ ~ U[0.5, 2], N2/O2 ~ U[1, 4], AR/O2 ~ U[0, 1]) with an OH seed of 1e-3 of the O2; output times 0.1, 0.3
and 1 of t_end = 1e-2 s. (The synthetic mechanisms' rate constants, up to 1e20 /s, make every scipy stiff solver stop
within a microsecond, so they serve the rates benchmarks, not this test.)
Run from the repository root: python tests/golden/make_reactor_golden.py"""
import os
import sys

import numpy as np
from scipy.integrate import solve_ivp

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))

from helpers import setup_text  # noqa: E402

R = 1.380649e-23 * 6.02214076e23
ATM = 101325.0
N_REACTORS = 12
T_END = 1e-2
FRACTIONS = (0.1, 0.3, 1.0)


def initial_states(mech, rng):
    names = mech["names"]
    K = mech["K"]
    W = np.asarray(mech["W"])
    T0 = rng.uniform(1300.0, 1600.0, N_REACTORS)
    P = rng.uniform(1.0, 5.0, N_REACTORS) * ATM
    Y = np.zeros((K, N_REACTORS))
    for s in range(N_REACTORS):
        phi, dil, ar = rng.uniform(0.5, 2.0), rng.uniform(1.0, 4.0), rng.uniform(0.0, 1.0)
        X = np.zeros(K)
        X[names.index("H2")] = 2.0 * phi
        X[names.index("O2")] = 1.0
        X[names.index("OH")] = 1e-3
        X[names.index("N2")] = dil
        X[names.index("AR")] = ar
        X /= X.sum()
        Y[:, s] = X * W / np.dot(X, W)
    return T0, P, Y


def main():
    M, mech, o = setup_text(open(os.path.join(ROOT, "tests", "golden", "handwritten.yaml")).read())
    W = np.asarray(mech["W"])
    Ka = mech["Ka"]
    rng = np.random.default_rng(2718)
    T0, P, Y0 = initial_states(mech, rng)
    times = np.array([f * T_END for f in FRACTIONS])
    out = np.zeros((len(times), 1 + mech["K"], N_REACTORS))
    for s in range(N_REACTORS):
        def f(t, y):
            T, Y = y[0], y[1:]
            omega, dT = o.rhs(T, P[s], Y)
            rho = P[s] / (R * T * np.sum(Y / W))
            dy = np.zeros_like(y)
            dy[0] = dT
            dy[1:1 + Ka] = W[:Ka] * omega / rho
            return dy
        atol = np.full(1 + mech["K"], 1e-14)
        atol[0] = 1e-10
        sol = solve_ivp(f, (0.0, T_END), np.concatenate([[T0[s]], Y0[:, s]]), method="BDF", t_eval=times, rtol=1e-10, atol=atol)
        assert sol.success, sol.message
        out[:, :, s] = sol.y.T
    path = os.path.join(ROOT, "tests", "golden", "reactor_handwritten.npz")
    np.savez(path, T0=T0, P=P, Y0=Y0, times=times, states=out)
    print(path, "final T", out[-1, 0])


if __name__ == "__main__":
    main()

"""Writes the fixtures of the GPU integrator's tests: constant-pressure reactor trajectories integrated by scipy's
BDF (solve_ivp, rtol 1e-10, atol 1e-14 on mass fractions and 1e-10 K) on the FP64 oracle's right-hand side (dT/dt
of the energy balance, dY_k/dt = W_k ω_k / ρ for the active species).

reactor_handwritten.npz: the hand-written H2/O2 mechanism (handwritten.yaml) through ignition. Initial states:
T0 ~ U[1300, 1600] K, P ~ U[1, 5] atm, H2/O2/N2/AR mixtures (φ ~ U[0.5, 2], N2/O2 ~ U[1, 4], AR/O2 ~ U[0, 1]) with an
OH seed of 1e-3 of the O2; output times 0.1, 0.3 and 1 of t_end = 1e-2 s.

reactor_gri53.npz: the 53-species synthetic mechanism, the first four of eight fuel/O2/N2 mixtures of
combustion_amd.reactor's initial_states (seed 11: T0 ~ U[900, 1600] K, P ~ U[1, 40] atm, φ ~ U[0.3, 3]) whose BDF
run reaches t_end (the others stop with a step size below the spacing of the time values); output times 0.1, 0.3 and
1 of t_end = 1e-3 s (a few minutes of BDF per reactor). Random mixtures of every species (synth.states) are out of
reach: their rates of 1e20 /s and more stop scipy's stiff solvers within a microsecond.
Run from the repository root: python tests/golden/make_reactor_golden.py [handwritten|gri-53]"""
import os
import sys
from multiprocessing import Pool

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


def trajectory(job):
    """One reactor from y0 at pressure P, sampled at times (scipy BDF on the oracle's right-hand side)."""
    text, y0, P, times = job
    M, mech, o = setup_text(text)
    W = np.asarray(mech["W"])
    Ka = mech["Ka"]

    def f(t, y):
        T, Y = y[0], y[1:]
        omega, dT = o.rhs(T, P, Y)
        rho = P / (R * T * np.sum(Y / W))
        dy = np.zeros_like(y)
        dy[0] = dT
        dy[1:1 + Ka] = W[:Ka] * omega / rho
        return dy
    atol = np.full(1 + mech["K"], 1e-14)
    atol[0] = 1e-10
    sol = solve_ivp(f, (0.0, times[-1]), y0, method="BDF", t_eval=times, rtol=1e-10, atol=atol)
    if not sol.success:
        print("dropped: T0", y0[0], "P", P, sol.message, flush=True)
        return None
    return sol.y.T


def write(name, text, T0, P, Y0, t_end, keep=None):
    """Integrates every initial state; keeps the first `keep` whose BDF run reached t_end (all when None)."""
    times = np.array([f * t_end for f in FRACTIONS])
    jobs = [(text, np.concatenate([[T0[s]], Y0[:, s]]), P[s], times) for s in range(len(T0))]
    with Pool(min(8, len(jobs))) as pool:
        done = pool.map(trajectory, jobs)
    ok = [s for s, r in enumerate(done) if r is not None][:keep]
    assert keep is None and len(ok) == len(jobs) or keep is not None and len(ok) == keep, "too few trajectories"
    T0, P, Y0 = T0[ok], P[ok], np.ascontiguousarray(Y0[:, ok])
    out = np.stack([done[s] for s in ok], axis=-1)  # [time][K+1][reactor]
    path = os.path.join(ROOT, "tests", "golden", f"reactor_{name}.npz")
    np.savez(path, T0=T0, P=P, Y0=Y0, times=times, states=out)
    print(path, "final T", out[-1, 0])


def main(which):
    if which == "handwritten":
        text = open(os.path.join(ROOT, "tests", "golden", "handwritten.yaml")).read()
        M, mech, o = setup_text(text)
        T0, P, Y0 = initial_states(mech, np.random.default_rng(2718))
        write("handwritten", text, T0, P, Y0, T_END)
    elif which == "gri-53":
        from combustion_amd import reactor, synth
        text = open(synth.mechanism_path("gri-53")).read()
        M, mech, o = setup_text(text)
        y0, P = reactor.initial_states(M, 8, 11)
        write("gri53", text, y0[:, 0].copy(), P, np.ascontiguousarray(y0[:, 1:].T), 1e-3, keep=4)
    else:
        raise SystemExit(f"unknown fixture {which}")


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "handwritten")

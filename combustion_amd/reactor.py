"""Batched constant-pressure reactor trajectories: the states the transport surrogate is trained on.

The reference integrates reactors with SUNDIALS CVODE (cvode/lib.rs:17-47) or explicit Euler (plot.rs:80-118);
neither is available here, so this module integrates many adiabatic constant-pressure reactors at once with a
linearly implicit Rosenbrock method on the GPU. State y = [T, Y_0..Y_{K-1}] at the trajectory's fixed pressure:

    dT/dt   = the chemistry program's energy balance (reaction.rs:375-386: -T Σ ω_k h_k/RT / Σ C_k cp_k/R)
    dY_k/dt = W_k ω_k / ρ (active species; the inert bulk species stays fixed), ρ = P / (R T Σ_j Y_j/W_j)

Method: the L-stable (2,3) Rosenbrock pair of Shampine & Reichelt, "The MATLAB ODE Suite" (SIAM J. Sci.
Comput. 18, 1997) — the method of ode23s — with d = 1/(2+√2), e32 = 6+√2:

    W = I - h d J,  k1 = W⁻¹ f(y),  F1 = f(y + h k1/2),  k2 = W⁻¹ (F1 - k1) + k1,  y1 = y + h k2,
    k3 = W⁻¹ (f(y1) - e32 (k2 - F1) - 2 (k1 - f(y))),  error = h (k1 - 2 k2 + k3) / 6,

with a finite-difference Jacobian J (one batched evaluation of K+2 perturbed states per trajectory) and a
per-trajectory step size. Right-hand sides are evaluated by the chemistry program's HIP kernel; LU
factorisations and solves of the small (K+1)×(K+1) systems use torch's batched linear algebra.
"""
from __future__ import annotations

import math

import numpy as np

R_GAS = 1.380649e-23 * 6.02214076e23
ATM = 101325.0
D_ROS = 1.0 / (2.0 + math.sqrt(2.0))
E32 = 6.0 + math.sqrt(2.0)

FUELS = ("H2", "CH4", "C2H6", "C3H8", "CH3OH", "C2H4", "CH2O", "C2H2")


class ChemistryRHS:
    """f(T, P, Y) -> (ω_active, dT/dt) on device tensors, by the chemistry program (rates | rhs) on the GPU."""

    def __init__(self, mechanism, form="fp64"):
        from . import native
        self.program = mechanism.chemistry(native.RATES | native.RHS)
        self.module = self.program.assemble(form=form)
        self.K = mechanism.K
        self.Ka = mechanism.active

    def __call__(self, T, P, Y):
        import torch
        n = T.shape[0]
        outs = [torch.empty(n, dtype=torch.float64, device=T.device) for _ in range(self.Ka + 1)]
        self.module([T, P, *Y], outs)
        return torch.stack(outs[:self.Ka]), outs[self.Ka]


class Reactors:
    """A batch of constant-pressure reactors integrated together (each with its own time and step size).

    rhs(T[n], P[n], Y[K][n]) -> (ω[Ka][n], dT/dt[n]) evaluates the chemistry on torch tensors; W are the molar
    masses [kg/mol]."""

    def __init__(self, rhs, W, Ka, rtol=1e-6, atol_T=1e-6, atol_Y=1e-12):
        import torch
        self.rhs = rhs
        self.W = torch.as_tensor(np.asarray(W, np.float64))
        self.K = len(W)
        self.Ka = Ka
        self.rtol, self.atol_T, self.atol_Y = rtol, atol_T, atol_Y

    def f(self, y, P):
        """dy/dt for y [n][K+1] at pressures P [n]."""
        import torch
        W = self.W.to(y.device)
        T = y[:, 0].contiguous()
        Y = y[:, 1:].t().contiguous()
        omega, dT = self.rhs(T, P.contiguous(), Y)
        rho = P / (R_GAS * T * (Y / W[:, None]).sum(0))
        out = torch.zeros_like(y)
        out[:, 0] = dT
        out[:, 1:1 + self.Ka] = (omega * W[:self.Ka, None] / rho).t()
        return out

    def jacobian(self, y, P):
        """Forward-difference Jacobian [n][m][m] (m = K+1) and f(y), from one batched evaluation."""
        import torch
        n, m = y.shape
        delta = 1.5e-8 * torch.maximum(y.abs(), torch.tensor([1.0] + [1e-6] * (m - 1), dtype=y.dtype, device=y.device))
        Yp = y[:, None, :].expand(n, m + 1, m).clone()
        idx = torch.arange(m, device=y.device)
        Yp[:, idx + 1, idx] += delta
        F = self.f(Yp.reshape(n * (m + 1), m), P[:, None].expand(n, m + 1).reshape(-1)).reshape(n, m + 1, m)
        f0 = F[:, 0]
        J = (F[:, 1:] - f0[:, None, :]).transpose(1, 2) / delta[:, None, :]
        return J, f0

    def step(self, y, P, h):
        """One Rosenbrock step of size h [n] for every trajectory: (y1, RMS error norm against the tolerances)."""
        import torch
        n, m = y.shape
        J, f0 = self.jacobian(y, P)
        Wm = torch.eye(m, dtype=y.dtype, device=y.device)[None] - (h * D_ROS)[:, None, None] * J
        LU, piv = torch.linalg.lu_factor(Wm)
        solve = lambda b: torch.linalg.lu_solve(LU, piv, b[:, :, None])[:, :, 0]
        k1 = solve(f0)
        F1 = self.f(y + 0.5 * h[:, None] * k1, P)
        k2 = solve(F1 - k1) + k1
        y1 = y + h[:, None] * k2
        F2 = self.f(y1, P)
        k3 = solve(F2 - E32 * (k2 - F1) - 2.0 * (k1 - f0))
        err = h[:, None] / 6.0 * (k1 - 2.0 * k2 + k3)
        atol = torch.full((m,), self.atol_Y, dtype=y.dtype, device=y.device)
        atol[0] = self.atol_T
        scale = atol + self.rtol * torch.maximum(y.abs(), y1.abs())
        norm = torch.sqrt(torch.mean((err / scale) ** 2, dim=1))
        return y1, norm

    def integrate(self, y0, P, t_end, h0=1e-10, max_steps=4000, record=None):
        """Integrates every trajectory to t_end (or max_steps); record(y, mask) sees each accepted step's states."""
        import torch
        y = y0.clone()
        n = y.shape[0]
        t = torch.zeros(n, dtype=y.dtype, device=y.device)
        h = torch.full((n,), h0, dtype=y.dtype, device=y.device)
        steps = 0
        while steps < max_steps:
            active = t < t_end * (1 - 1e-12)
            if not bool(active.any()):
                break
            h = torch.where(active, torch.minimum(h, t_end - t), h)
            y1, norm = self.step(y, P, h)
            ok = active & (norm <= 1.0) & torch.isfinite(norm)
            t = torch.where(ok, t + h, t)
            y = torch.where(ok[:, None], y1, y)
            factor = torch.where(torch.isfinite(norm), 0.8 * torch.clamp(norm, min=1e-12) ** (-1.0 / 3.0), torch.full_like(norm, 0.1))
            h = h * torch.clamp(factor, 0.2, 5.0)
            if record is not None and bool(ok.any()):
                record(y, ok)
            steps += 1
        return y, t, steps


def initial_states(mechanism, n, seed, T_range=(900.0, 1600.0), P_range_atm=(1.0, 40.0)):
    """Random fuel/O2/N2 mixtures: a fuel among the mechanism's named fuels, equivalence ratio ~ U[0.3, 3]
    (against the fuel's oxygen demand), N2/O2 molar ratio ~ U[1, 6]. Returns (y0 [n][K+1], P [n]) as numpy."""
    from .synth import parse_formula
    rng = np.random.Generator(np.random.PCG64(seed))
    names = mechanism.species_names
    K = len(names)
    W = mechanism.molar_mass
    fuels = [f for f in FUELS if f in names]
    o2, n2 = names.index("O2"), K - 1
    y0 = np.zeros((n, K + 1))
    P = np.zeros(n)
    for i in range(n):
        fuel = fuels[int(rng.integers(len(fuels)))]
        c = parse_formula(fuel)
        demand = c.get("C", 0) + c.get("H", 0) / 4.0 - c.get("O", 0) / 2.0  # mol O2 per mol fuel
        phi = rng.uniform(0.3, 3.0)
        X = np.zeros(K)
        X[names.index(fuel)] = phi / demand
        X[o2] = 1.0
        X[n2] = rng.uniform(1.0, 6.0)
        X /= X.sum()
        Y = X * W / np.dot(X, W)
        y0[i, 0] = rng.uniform(*T_range)
        y0[i, 1:] = Y
        P[i] = rng.uniform(*P_range_atm) * ATM
    return y0, P


def trajectories(mechanism, n_trajectories, n_states, seed, t_end=1e-2, batch=2048, device="cuda", max_steps=3000, keep=1.0,
                 log=None):
    """Samples n_states states (T [n], P [n], Y [K][n], numpy) from constant-pressure trajectories started at
    initial_states(seed): each accepted step's state is a candidate with probability keep (stored in FP32), and a
    seeded uniform draw picks n_states of the candidates."""
    import torch
    K, Ka = mechanism.K, mechanism.active
    rhs = ChemistryRHS(mechanism)
    reactors = Reactors(rhs, mechanism.molar_mass, Ka)
    y0, P0 = initial_states(mechanism, n_trajectories, seed)
    gen = torch.Generator(device=device)
    gen.manual_seed(seed)
    chunks = []
    for b in range(0, n_trajectories, batch):
        y = torch.from_numpy(y0[b:b + batch]).to(device)
        P = torch.from_numpy(P0[b:b + batch]).to(device)
        saved = []

        def record(y_now, ok, P=P, saved=saved):
            if keep < 1.0:
                ok = ok & (torch.rand(ok.shape, generator=gen, device=ok.device, dtype=torch.float64) < keep)
            saved.append(torch.cat([y_now[ok], P[ok, None]], dim=1).to(torch.float32).cpu())

        _, t, steps = reactors.integrate(y, P, t_end, record=record, max_steps=max_steps)
        states = torch.cat(saved).numpy()
        chunks.append(states)
        if log:
            log(f"trajectories {b}..{b + len(y)}: {steps} steps, {len(states)} states, "
                f"{float((t >= t_end * (1 - 1e-12)).float().mean()):.3f} reached t_end")
    allstates = np.concatenate(chunks).astype(np.float64)
    rng = np.random.Generator(np.random.PCG64(seed + 1))
    pick = rng.choice(len(allstates), size=n_states, replace=len(allstates) < n_states)
    s = allstates[pick]
    T, Y, P = s[:, 0].copy(), np.ascontiguousarray(s[:, 1:1 + K].T), s[:, 1 + K].copy()
    Y = np.clip(Y, 0.0, None)
    Y /= Y.sum(axis=0, keepdims=True)
    return T, P, Y

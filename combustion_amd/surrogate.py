"""Learned mixture-averaged transport: training data from reactor trajectories, training with the project's
MLP kernels (kernels/mlp.hip through cb_mlp_*), and the surrogate form (cb_hip_assemble_surrogate).

The network maps the state's features f = (T, ln P, X_0..X_{K-1}) (mole fractions from the mass fractions Y) to
r_o = ln(p_o / b_o) for the K+2 transport outputs p = [λ, μ, D_0..D_{K-1}] of the chemistry program
(transport.rs:141-177), relative to baselines b that cost O(K·M) per state instead of O(K²): the mixture rules
with the pair sums over j restricted to M major species S (the M largest mole fractions seen in the training
states) and the remaining mole fraction X_rest = 1 - Σ_S X_j given to the bulk species b, from the transport fits
in ℓ = ln(T/T0) (transport.rs:125-136; λ_k = c_k(ℓ)√T, η_k = v_k(ℓ)²√T, D_kj = d_kj(ℓ) T^1.5/P):
    λ_b = ½ (Σ X_k λ_k + 1/Σ X_k/λ_k)          the conductivity rule itself, which is O(K)
    μ_b = Σ_k X_k η_k / den_k,  den_k = Σ_S X_j φ_kj + X'_k φ_kb + [k ∉ S] X_k      (Wilke, transport.rs:148-158)
    D_b,k = (1 - Y_k) / (Σ_{S, j≠k} X_j / D_kj + X'_k / D_k,r(k))                    (transport.rs:162-177)
with X'_k = X_rest - [k ∉ S] X_k and r(k) = b, or for k = b the first other major species; so the network learns
what the minor species add. Features and targets are standardised with the training set's means and deviations;
the box [min, max] of the training features is the domain where the network is used — states outside it are
evaluated by the plain transport kernel.
"""
from __future__ import annotations

import ctypes as C
import math

import numpy as np

from . import native


class Trained:
    """A trained surrogate: network parameters (flat FP32, cb_mlp layout), normalisation and the major species of
    its baselines."""

    FIELDS = ("params", "f_mean", "f_rstd", "lo", "hi", "o_mean", "o_std")

    def __init__(self, width, depth, params, f_mean, f_rstd, lo, hi, o_mean, o_std, major):
        self.width, self.depth = int(width), int(depth)
        self.params, self.f_mean, self.f_rstd = (np.ascontiguousarray(x, np.float32) for x in (params, f_mean, f_rstd))
        self.lo, self.hi, self.o_mean, self.o_std = (np.ascontiguousarray(x, np.float32) for x in (lo, hi, o_mean, o_std))
        self.major = np.ascontiguousarray(major, np.int32)

    def save(self, path):
        np.savez(path, width=self.width, depth=self.depth, major=self.major, **{k: getattr(self, k) for k in self.FIELDS})

    @classmethod
    def load(cls, path):
        z = np.load(path, allow_pickle=False)
        return cls(int(z["width"]), int(z["depth"]), *(z[k] for k in cls.FIELDS), z["major"])


def mole_fractions(mechanism, Y):
    n = np.asarray(Y) / mechanism.molar_mass[:, None]
    return n / n.sum(axis=0, keepdims=True)


def features(mechanism, T, P, Y):
    """[n][K+2] float64: (T, ln P, X_0..X_{K-1})."""
    X = mole_fractions(mechanism, Y)
    return np.concatenate([np.asarray(T)[:, None], np.log(np.asarray(P))[:, None], X.T], axis=1)


def choose_major(mechanism, Y, m=16):
    """The m species with the largest mole fraction over the states Y [K][n], largest first (int32)."""
    return np.argsort(-mole_fractions(mechanism, Y).max(axis=1), kind="stable")[:m].astype(np.int32)


def rest_partner(K, major):
    """r(k): the bulk species K-1, or for the bulk itself the first other major species."""
    r = np.full(K, K - 1, np.int32)
    r[K - 1] = next(int(j) for j in major if j != K - 1)
    return r


def wilke_coefficients(W):
    """(a, b) [K][K] with Wilke's φ_kj = (a_kj + b_kj v_k/v_j)², v = √η (transport.rs:148-158)."""
    a = 1 / (8 * (1 + W[:, None] / W[None, :])) ** 0.25
    return a, a * (W[None, :] / W[:, None]) ** 0.25


def log_baseline(mechanism, T, P, Y, major, D=4, T0=1000.0):
    """ln b_o [n][K+2] (see the module docstring) for the major species `major`."""
    c, v, d = mechanism.transport_fits(D, T0)
    K = mechanism.K
    b = K - 1
    major = np.asarray(major, np.int64)
    in_S = np.zeros(K, bool)
    in_S[major] = True
    r = rest_partner(K, major)
    T = np.asarray(T, np.float64)
    Y = np.asarray(Y, np.float64)
    X = mole_fractions(mechanism, Y)                   # [K][n]
    l = np.log(T / T0)
    pw = np.stack([l ** i for i in range(D)], axis=1)  # [n][D]
    sqrtT = np.sqrt(T)
    lam = (pw @ c.T).T * sqrtT                          # [K][n]
    vk = (pw @ v.T).T                                   # [K][n]
    a, bb = wilke_coefficients(mechanism.molar_mass)
    X_rest = np.maximum(1 - X[major].sum(0), 0)
    X_other = np.maximum(X_rest[None, :] - np.where(in_S[:, None], 0, X), 0)  # X'_k
    den = np.zeros_like(X)
    S = np.zeros_like(X)
    for j in major:
        den += X[j] * (a[:, j, None] + bb[:, j, None] * vk / vk[j]) ** 2
        dkj = np.einsum("kd,nd->kn", d[:, j], pw)
        S += np.where(np.arange(K)[:, None] == j, 0, X[j] / np.where(np.arange(K)[:, None] == j, 1, dkj))
    den += X_other * (a[:, b, None] + bb[:, b, None] * vk / vk[b]) ** 2 + np.where(in_S[:, None], 0, X)
    S += X_other / np.einsum("kd,nd->kn", d[np.arange(K), r], pw)
    out = np.empty((len(T), K + 2))
    out[:, 0] = np.log(0.5 * ((X * lam).sum(0) + 1 / (X / lam).sum(0)))
    out[:, 1] = np.log(sqrtT * (X * vk ** 2 / den).sum(0))
    out[:, 2:] = (np.log(np.maximum(1 - Y, 1e-30)) - np.log(np.maximum(S, 1e-300)) + 1.5 * np.log(T) - np.log(np.asarray(P))).T
    return out


def exact_transport(mechanism, T, P, Y, device="cuda", chunk=1 << 20):
    """The plain FP64 transport outputs [K+2][n] (tiled FP64 form on the GPU)."""
    import torch
    prog = mechanism.chemistry(native.TRANSPORT)
    mod = prog.assemble(form="fp64")
    outs = []
    for i in range(0, len(T), chunk):
        ins = [torch.from_numpy(np.ascontiguousarray(x[i:i + chunk])).to(device) for x in (T, P)]
        ins += [torch.from_numpy(np.ascontiguousarray(y[i:i + chunk])).to(device) for y in Y]
        outs.append(torch.stack(mod(ins)).cpu().numpy())
    return np.concatenate(outs, axis=1)


def initial_params(net, rng):
    """LeCun-normal weights on the valid inputs, zero biases and zero padding."""
    params = np.zeros(net.n_params, np.float32)
    for l, (w, b, o, i) in enumerate(net.layers):
        fan_in = net.n_in if l == 0 else i
        W = rng.normal(0, 1 / math.sqrt(fan_in), (o, i)).astype(np.float32)
        if l == 0:
            W[:, net.n_in:] = 0
        if l == len(net.layers) - 1:
            W[net.n_out:] = 0
        params[w:w + o * i] = W.ravel()
    return params


def train(mechanism, T, P, Y, props, width=256, depth=3, steps=4000, batch=16384, lr=2e-3, seed=0, device="cuda", log=None,
          hard=0.0, refresh=1000, common_scale=False, n_major=20):
    """Fits the network to props [K+2][n] (the exact FP64 transport kernel's outputs) on the GPU with the project's kernels.

    hard > 0 mines hard examples: every `refresh` steps the network is evaluated on a random sample of the training
    states and a fraction `hard` of each later batch is drawn from the 5% of that sample with the largest error
    max_o |Δ ln p_o|. common_scale standardises every output by the largest deviation instead of its own, so that the
    loss weighs each output's error in ln p_o alike (the relative error that the bound is stated in). The baselines
    keep the pair terms of the n_major species with the largest mole fractions in the training states."""
    import torch
    K = mechanism.K
    NF = K + 2
    F = features(mechanism, T, P, Y)
    major = choose_major(mechanism, Y, n_major)
    R = np.log(np.asarray(props).T) - log_baseline(mechanism, T, P, Y, major)
    f_mean, f_std = F.mean(0), F.std(0)
    f_rstd = np.where(f_std > 0, 1 / np.maximum(f_std, 1e-300), 0.0)
    o_mean, o_std = R.mean(0), np.maximum(R.std(0), 1e-6)
    if common_scale:
        o_std = np.full_like(o_std, o_std.max())
    net = native.Network(NF, width, depth, NF)
    X = np.zeros((len(T), net.in_pad), np.float32)
    X[:, :NF] = (F - f_mean) * f_rstd
    Tg = np.zeros((len(T), net.out_pad), np.float32)
    Tg[:, :NF] = (R - o_mean) / o_std
    dev = torch.device(device)
    Xd, Td = torch.from_numpy(X).to(dev), torch.from_numpy(Tg).to(dev)
    rng = np.random.default_rng(seed)
    p = torch.from_numpy(initial_params(net, rng)).to(dev)
    p16 = p.to(torch.bfloat16)
    m1, m2, grad = torch.zeros_like(p), torch.zeros_like(p), torch.zeros_like(p)
    y = torch.empty((batch, net.out_pad), dtype=torch.float32, device=dev)
    ws = torch.empty(net.workspace_bytes(batch), dtype=torch.uint8, device=dev)
    gen = torch.Generator(device=dev)
    gen.manual_seed(seed)
    n = len(T)
    perm = torch.randperm(n, generator=gen, device=dev)
    at = 0
    history = []
    n_hard = int(round(hard * batch))
    pool = None
    scale = torch.from_numpy(o_std.astype(np.float32)).to(dev)
    for step in range(1, steps + 1):
        if n_hard and step % refresh == 0 and step < steps:
            pool = _hard_pool(net, p, p16, Xd, Td, scale, NF, batch, ws, y, gen)
        m = batch - n_hard if pool is not None else batch
        if at + m > n:
            perm = torch.randperm(n, generator=gen, device=dev)
            at = 0
        idx = perm[at:at + m]
        at += m
        if pool is not None:
            idx = torch.cat([idx, pool[torch.randint(len(pool), (n_hard,), generator=gen, device=dev)]])
        xb, tb = Xd[idx].contiguous(), Td[idx].contiguous()
        net.forward(p, p16, xb, y, ws)
        net.backward_mse(p16, y, tb, ws, grad)
        rate = lr * 0.5 * (1 + math.cos(math.pi * step / steps)) + 1e-5
        net.adam(p, p16, grad, m1, m2, rate, step)
        if step == 1 or step % 500 == 0 or step == steps:
            loss = float(((y[:, :NF] - tb[:, :NF]) ** 2).mean())
            history.append((step, loss))
            if log:
                log(f"step {step}: batch loss {loss:.3e}")
    trained = Trained(width, depth, p.cpu().numpy(), f_mean, f_rstd, F.min(0), F.max(0), o_mean, o_std, major)
    trained.history = history
    return trained


def _hard_pool(net, p, p16, Xd, Td, scale, NF, batch, ws, y, gen, sample=1 << 19, top=0.05):
    """Indices of the training states with the largest max_o |Δ ln p_o| in a random sample."""
    import torch
    n = Xd.shape[0]
    idx = torch.randperm(n, generator=gen, device=Xd.device)[:min(sample, n)]
    errs = []
    for i in range(0, len(idx), batch):
        j = idx[i:i + batch]
        xb = Xd[j].contiguous()
        if len(j) < batch:
            xb = torch.cat([xb, xb[:1].expand(batch - len(j), -1)])
        net.forward(p, p16, xb, y, ws)
        errs.append(((y[:len(j), :NF] - Td[j, :NF]).abs() * scale).amax(1))
    e = torch.cat(errs)
    k = max(1, int(top * len(e)))
    return idx[torch.topk(e, k).indices]


def assemble(program, trained, rates="fast"):
    """The surrogate form of a chemistry program (cb_hip_assemble_surrogate) as a native.Module; rates is "fast"
    (the FP32-core straight-line rates kernel) or "mixed" (the MFMA mixed-precision one)."""
    if rates not in ("fast", "mixed"):
        raise ValueError("rates is 'fast' or 'mixed'")
    K = program.mechanism.K
    arrays = [np.ascontiguousarray(getattr(trained, k), np.float32) for k in Trained.FIELDS]
    for a in arrays[1:]:
        if a.shape != (K + 2,):
            raise ValueError("normalisation arrays need K+2 entries")
    fp = C.POINTER(C.c_float)
    w = native.SurrogateWeights(trained.width, trained.depth, *(a.ctypes.data_as(fp) for a in arrays), len(trained.major),
                                trained.major.ctypes.data_as(C.POINTER(C.c_int32)), native.FORMS[rates])
    h = native._P()
    native._check(native.lib().cb_hip_assemble_surrogate(program._h, C.byref(w), C.byref(h)))
    mod = native.Module(h, program)
    mod._keep = arrays
    return mod


def fallback_count(module, stream=None):
    c = C.c_uint32()
    native._check(native.lib().cb_module_fallback_count(module._h, stream, C.byref(c)))
    return c.value


def main(argv=None):
    """Generates trajectory states, computes the exact FP64 transport on them, trains the network and saves it:
    python -m combustion_amd.surrogate --mechanism gri-53 --out surrogate_gri53.npz"""
    import argparse
    import time

    from . import reactor, synth
    ap = argparse.ArgumentParser()
    ap.add_argument("--mechanism", default="gri-53")
    ap.add_argument("--trajectories", type=int, default=8192)
    ap.add_argument("--states", type=int, default=1 << 22)
    ap.add_argument("--max-steps", type=int, default=1200)
    ap.add_argument("--t-end", type=float, default=1e-3)
    ap.add_argument("--steps", type=int, default=6000)
    ap.add_argument("--batch", type=int, default=16384)
    ap.add_argument("--width", type=int, default=256)
    ap.add_argument("--depth", type=int, default=3)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--hard", type=float, default=0.0, help="fraction of each batch drawn from mined hard states")
    ap.add_argument("--common-scale", action="store_true", help="one output scale for every output")
    ap.add_argument("--major", type=int, default=20, help="major species kept in the baselines' pair sums")
    ap.add_argument("--out", required=True)
    a = ap.parse_args(argv)
    log = lambda m: print(f"[{time.strftime('%H:%M:%S')}] {m}", flush=True)
    mech = native.Mechanism(synth.mechanism_path(a.mechanism))
    t = time.time()
    T, P, Y = reactor.trajectories(mech, a.trajectories, a.states, a.seed, t_end=a.t_end, max_steps=a.max_steps, log=log)
    log(f"{len(T)} training states from {a.trajectories} trajectories in {time.time() - t:.1f} s; T in [{T.min():.0f}, {T.max():.0f}]")
    t = time.time()
    props = exact_transport(mech, T, P, Y)
    log(f"exact transport in {time.time() - t:.1f} s")
    t = time.time()
    trained = train(mech, T, P, Y, props, width=a.width, depth=a.depth, steps=a.steps, batch=a.batch, lr=a.lr, seed=a.seed, log=log,
                    hard=a.hard, common_scale=a.common_scale, n_major=a.major)
    log(f"trained in {time.time() - t:.1f} s")
    trained.save(a.out)
    log(f"saved {a.out}")


if __name__ == "__main__":
    main()

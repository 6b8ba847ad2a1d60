"""The transport surrogate end to end on the GPU: trajectory states from the project's reactor integrator,
training with the project's kernels, the surrogate form against the FP64 oracle, and its fallback."""
import os

import numpy as np
import pytest

from combustion_amd import native, reactor, surrogate, synth
from helpers import relative_error, setup
from oracle import mlp as omlp

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
TRAINED = os.path.join(ROOT, "combustion_amd", "data", "surrogate_gri53.npz")
SURROGATE_TOL = 1e-3  # the stated bound: transport outputs against the FP64 oracle


def evaluate(mod, T, P, Y, gpu):
    import torch
    ins = [torch.from_numpy(np.ascontiguousarray(x)).to(gpu) for x in (T, P, *Y)]
    out = mod(ins)
    torch.cuda.synchronize()
    return np.stack([o.cpu().numpy() for o in out])


@pytest.fixture(scope="module")
def small(gpu):
    """A network trained briefly on 2^14 states of 64 trajectories (the states, their exact transport, the network)."""
    M, mech, o = setup("gri-53")
    T, P, Y = reactor.trajectories(M, 64, 1 << 14, seed=3, max_steps=200, batch=64)
    assert np.all(np.isfinite(T)) and T.min() > 200 and np.allclose(Y.sum(0), 1)
    props = surrogate.exact_transport(M, T, P, Y)
    trained = surrogate.train(M, T, P, Y, props, steps=300, batch=4096)
    return T, P, Y, trained


@pytest.mark.gpu
def test_trajectories_train_and_fallback(gpu, small):
    """The loss falls, and states outside the training box come back from the plain FP64 transport kernel (the
    FP64 forms' 1e-12 of the oracle)."""
    M, mech, o = setup("gri-53")
    T, P, Y, trained = small
    first, last = trained.history[0][1], trained.history[-1][1]
    print("loss", trained.history, flush=True)
    assert last < first
    mod = surrogate.assemble(M.chemistry(2), trained)
    Th, Ph, Yh = synth.config_states("gri-53", 4096, 21)  # far outside the trajectory box: all fall back
    got = evaluate(mod, Th, Ph, Yh, gpu)
    assert surrogate.fallback_count(mod) == 4096
    ref, floor = o.batch(Th, Ph, Yh, rates=False, threads=8)
    assert relative_error(got, ref, floor).max() <= 1e-12


@pytest.mark.gpu
def test_inference_kernel_matches_oracle_network(gpu, small):
    """kernels/surrogate.hip against the oracle's forward pass (oracle/mlp.py, the kernels' BF16 rounding points
    reproduced) on the numpy features and baselines, for the training states inside the box. The kernel computes
    features and baselines in FP32, so a feature or an activation can round to the neighbouring BF16 value: the
    bar is statistical (99th percentile 1e-4 in ln p, max 1e-2), and a wrong operand layout, row or bias shows as
    O(1)."""
    M, mech, o = setup("gri-53")
    T, P, Y, trained = small
    NF = M.K + 2
    mod = surrogate.assemble(M.chemistry(2), trained)
    n = len(T)
    got = evaluate(mod, T, P, Y, gpu)
    print("fallback", surrogate.fallback_count(mod), "of", n, flush=True)
    F = surrogate.features(M, T, P, Y)
    # The kernel tests its FP32 features against the FP32 box: compare only the states clear of the box's edges
    # by more than FP32 rounding (1e-5 of the bound; a bound of 0 is exact).
    lo, hi = trained.lo.astype(np.float64), trained.hi.astype(np.float64)
    inside = np.all((F >= lo + 1e-5 * np.abs(lo)) & (F <= hi - 1e-5 * np.abs(hi)), axis=1)
    assert inside.sum() > 0.9 * n
    net = native.Network(NF, trained.width, trained.depth, NF)
    x = np.zeros((n, net.in_pad))
    x[:, :NF] = (F - trained.f_mean) * trained.f_rstd
    y, _ = omlp.forward(trained.params, net.layers, x, rounding=True)
    expect = y[:, :NF] * trained.o_std + trained.o_mean + surrogate.log_baseline(M, T, P, Y, trained.major)
    d = np.abs(np.log(got.T[inside]) - expect[inside])
    print(f"|ln kernel - ln oracle|: 99th percentile {np.percentile(d, 99):.2e}, max {d.max():.2e}", flush=True)
    assert np.percentile(d, 99) < 1e-4 and d.max() < 1e-2


def held_out(gpu):
    """The committed network on states of 256 held-out trajectories (seed 99): errors against the FP64 oracle of
    the surrogate form and of its O(K) baselines alone."""
    if not os.path.exists(TRAINED):
        pytest.skip("no trained network committed")
    M, mech, o = setup("gri-53")
    trained = surrogate.Trained.load(TRAINED)
    T, P, Y = reactor.trajectories(M, 256, 1 << 15, seed=99, max_steps=600, batch=256)
    ref, floor = o.batch(T, P, Y, rates=False, threads=8)
    mod = surrogate.assemble(M.chemistry(2), trained)
    got = evaluate(mod, T, P, Y, gpu)
    err = relative_error(got, ref, floor)
    base = relative_error(np.exp(surrogate.log_baseline(M, T, P, Y, trained.major)).T, ref, floor)
    fb = surrogate.fallback_count(mod)
    print(f"held-out: max {err.max():.2e}, p99 {np.percentile(err, 99):.2e}, median {np.median(err):.2e}, "
          f"λ {err[0].max():.2e}, μ {err[1].max():.2e}, D {err[2:].max():.2e}, fallback {fb}/{len(T)}; "
          f"baselines alone: max {base.max():.2e}, p99 {np.percentile(base, 99):.2e}, median {np.median(base):.2e}",
          flush=True)
    return err, base


@pytest.mark.gpu
def test_trained_network_gain_on_held_out_trajectories(gpu):
    """What the network gains over the method without it (the O(K·M) baselines alone), on held-out trajectory
    states: at least 5x lower max and 2x lower 99th-percentile relative error. (The median rises: where the
    baselines are exact to 1e-11 the network's FP32/BF16 output adds ~1e-7.)"""
    err, base = held_out(gpu)
    assert err.max() * 5 <= base.max()
    assert np.percentile(err, 99) * 2 <= np.percentile(base, 99)


@pytest.mark.gpu
def test_trained_network_meets_the_stated_bound(gpu):
    """The stated bound: every held-out output within 1e-3 of the FP64 oracle (fallback states included)."""
    err, _ = held_out(gpu)
    assert err.max() <= SURROGATE_TOL


@pytest.mark.gpu
def test_surrogate_with_the_mfma_mixed_rates(gpu):
    """The mixed form's pairing: the network's transport beside the MFMA mixed-precision rates kernel (split-BF16
    contractions, FP32 exp/log, FP64 ΔG/RT and forward-minus-reverse), on held-out trajectory states: ω within the
    mixed form's 1e-5 of the state floor and transport within 1e-3 of the FP64 oracle; ω the same bits as the mixed
    form's own."""
    if not os.path.exists(TRAINED):
        pytest.skip("no trained network committed")
    from helpers import state_floor
    M, mech, o = setup("gri-53")
    trained = surrogate.Trained.load(TRAINED)
    T, P, Y = reactor.trajectories(M, 64, 1 << 13, seed=98, max_steps=600, batch=64)
    ref, floor = o.batch(T, P, Y, threads=8)
    prog = M.chemistry()
    got = evaluate(surrogate.assemble(prog, trained, rates="mixed"), T, P, Y, gpu)
    mixed = evaluate(prog.assemble(form="mixed"), T, P, Y, gpu)
    Ka = M.active
    assert np.array_equal(got[:Ka], mixed[:Ka])
    err = relative_error(got, ref, state_floor(floor, Ka))
    print(f"surrogate + mixed rates: ω {err[:Ka].max():.2e}, transport {err[Ka:].max():.2e}", flush=True)
    assert err[:Ka].max() <= 1e-5 and err[Ka:].max() <= SURROGATE_TOL


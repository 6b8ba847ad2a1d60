"""The transport surrogate end to end on the GPU: trajectory states from the project's reactor integrator,
training with the project's kernels, the surrogate form against the FP64 oracle, and its fallback."""
import os

import numpy as np
import pytest

from combustion_amd import reactor, surrogate, synth
from helpers import relative_error, setup

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
TRAINED = os.path.join(ROOT, "combustion_amd", "data", "surrogate_gri53.npz")
SURROGATE_TOL = 1e-3  # transport outputs against the FP64 oracle


def evaluate(mod, T, P, Y, gpu):
    import torch
    ins = [torch.from_numpy(np.ascontiguousarray(x)).to(gpu) for x in (T, P, *Y)]
    out = mod(ins)
    torch.cuda.synchronize()
    return np.stack([o.cpu().numpy() for o in out])


@pytest.mark.gpu
def test_trajectories_train_and_fallback(gpu):
    """A small network trained on states of 64 trajectories: the loss falls, and states outside the training box
    come back from the plain FP32 transport kernel (within its FP32 accuracy of the oracle)."""
    M, mech, o = setup("gri-53")
    T, P, Y = reactor.trajectories(M, 64, 1 << 14, seed=3, max_steps=200, batch=64)
    assert np.all(np.isfinite(T)) and T.min() > 200 and np.allclose(Y.sum(0), 1)
    props = surrogate.judge(M, T, P, Y)
    trained = surrogate.train(M, T, P, Y, props, steps=300, batch=4096)
    first, last = trained.history[0][1], trained.history[-1][1]
    print("loss", trained.history, flush=True)
    assert last < first
    prog = M.chemistry(2)
    mod = surrogate.assemble(prog, trained)
    Th, Ph, Yh = synth.config_states("gri-53", 4096, 21)  # far outside the trajectory box: all fall back
    got = evaluate(mod, Th, Ph, Yh, gpu)
    assert surrogate.fallback_count(mod) == 4096
    ref, floor = o.batch(Th, Ph, Yh, rates=False, threads=8)
    assert relative_error(got, ref, floor).max() < 1e-5


@pytest.mark.gpu
def test_trained_surrogate_meets_the_transport_bound(gpu):
    """The committed network (trained on 2^22 trajectory states by `python -m combustion_amd.surrogate`) on held-out
    trajectory states: every output within 1e-3 of the FP64 oracle, with the fallback states included."""
    if not os.path.exists(TRAINED):
        pytest.skip("no trained network committed")
    M, mech, o = setup("gri-53")
    trained = surrogate.Trained.load(TRAINED)
    T, P, Y = reactor.trajectories(M, 256, 1 << 15, seed=99, max_steps=600, batch=256)
    ref, floor = o.batch(T, P, Y, rates=False, threads=8)
    mod = surrogate.assemble(M.chemistry(2), trained)
    got = evaluate(mod, T, P, Y, gpu)
    err = relative_error(got, ref, floor)
    fb = surrogate.fallback_count(mod)
    print(f"held-out: max {err.max():.2e}, median {np.median(err):.2e}, λ {err[0].max():.2e}, μ {err[1].max():.2e}, "
          f"D {err[2:].max():.2e}, fallback {fb}/{len(T)}", flush=True)
    assert err.max() <= SURROGATE_TOL

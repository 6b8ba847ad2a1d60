"""The batched reactor integrator on the GPU (cb_integrator_*: Rosenbrock ode23s with the analytic Jacobian, one wave
per reactor for the LU and the solves), the counterpart of the reference's CVODE wrapper (cvode/lib.rs:17-47) and
reactor loop (plot.rs:80-118): twelve constant-pressure H2/O2 reactors of the hand-written mechanism integrated
through ignition, against the committed scipy BDF trajectories (tests/golden/make_reactor_golden.py, rtol 1e-10 on
the FP64 oracle's right-hand side) at 1, 3 and 10 ms. Bars: with the integrator's tolerances at rtol 1e-8, atol 1e-6 K
and 1e-12, temperatures within 2e-6 relative and mass fractions within 1e-6 absolute of the reference trajectory
(measured: 1.6e-7 and 8.5e-8 at 1 ms, mid-ignition for some reactors, where timing differences weigh most)."""
import os

import numpy as np
import pytest

from combustion_amd import Mechanism
from combustion_amd.native import Integrator

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")


@pytest.mark.gpu
def test_reactors_through_ignition_match_scipy_bdf(gpu):
    import torch
    d = np.load(os.path.join(GOLDEN, "reactor_handwritten.npz"))
    M = Mechanism(os.path.join(GOLDEN, "handwritten.yaml"))
    n = len(d["T0"])
    y = torch.from_numpy(np.vstack([d["T0"], d["Y0"]])).to(gpu).contiguous()
    P = torch.from_numpy(d["P"]).to(gpu)
    t = torch.zeros(n, dtype=torch.float64, device=gpu)
    h = torch.full((n,), 1e-9, dtype=torch.float64, device=gpu)
    status = torch.zeros(n, dtype=torch.int32, device=gpu)
    steps = torch.zeros(n, dtype=torch.int32, device=gpu)
    integ = Integrator(M, rtol=1e-8, atol_T=1e-6, atol_Y=1e-12)
    for i, t_out in enumerate(d["times"]):
        status.zero_()
        taken = integ.advance(y, P, t, h, status, steps, float(t_out), max_steps=200000)
        torch.cuda.synchronize()
        assert torch.all(status == 1), status.cpu().numpy()
        np.testing.assert_allclose(t.cpu().numpy(), t_out, rtol=1e-12)
        got = y.cpu().numpy()
        ref = d["states"][i]
        dT = np.abs(got[0] - ref[0]) / ref[0]
        dY = np.abs(got[1:] - ref[1:])
        print(f"t = {t_out:.1e} s: {taken} batch steps, max |ΔT|/T {dT.max():.2e}, max |ΔY| {dY.max():.2e}, "
              f"steps per reactor {steps.cpu().numpy().min()}..{steps.cpu().numpy().max()}", flush=True)
        assert dT.max() <= 2e-6 and dY.max() <= 1e-6
    assert np.all(d["states"][-1][0] > d["T0"] + 800)  # every reactor ignited


@pytest.mark.gpu
def test_integrator_limits_are_reported(gpu):
    import torch
    from combustion_amd import synth
    from combustion_amd.native import Error
    big = Mechanism(synth.mechanism_path("large-160"))
    with pytest.raises(Error) as e:
        Integrator(big)
    assert e.value.kind == "unsupported"
    M = Mechanism(os.path.join(GOLDEN, "handwritten.yaml"))
    integ = Integrator(M)
    x = torch.zeros((M.K + 1, 4), dtype=torch.float64, device=gpu)
    with pytest.raises(ValueError):
        integ.advance(x[:-1], x[0].clone(), x[0].clone(), x[0].clone(), torch.zeros(4, dtype=torch.int32, device=gpu),
                      torch.zeros(4, dtype=torch.int32, device=gpu), 1e-3)

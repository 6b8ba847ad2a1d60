"""The batched reactor integrator on the GPU (cb_integrator_*: Rosenbrock RODAS4 or ode23s, ⌊64/(K+1)⌋ reactors per
wave for the LU and the solves), the counterpart of the reference's CVODE wrapper (cvode/lib.rs:17-47) and reactor loop
(plot.rs:80-118): twelve constant-pressure H2/O2 reactors of the hand-written mechanism integrated through ignition,
against the committed scipy BDF trajectories (tests/golden/make_reactor_golden.py, rtol 1e-10 on the FP64 oracle's
right-hand side) at 1, 3 and 10 ms, with the analytic Jacobian and with difference quotients, several reactors per
wave and one. Bars: with the integrator's tolerances at rtol 1e-8, atol 1e-6 K and 1e-12, temperatures within 2e-6
relative and mass fractions within 1e-6 absolute of the reference trajectory for ode23s (measured: 2.9e-7 and 1.6e-7
at 1 ms, mid-ignition for some reactors, where timing differences weigh most), within 1e-8 for RODAS4 (measured:
3.6e-10 and 1.3e-9)."""
import os

import numpy as np
import pytest

from combustion_amd import Mechanism
from combustion_amd.native import Integrator

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")


def against_fixture(gpu, d, M, integ, dT_max, dY_max):
    """Advances the fixture's reactors to each of its output times, checking every reactor's state there."""
    import torch
    n = len(d["T0"])
    y = torch.from_numpy(np.vstack([d["T0"], d["Y0"]])).to(gpu).contiguous()
    P = torch.from_numpy(d["P"]).to(gpu)
    t = torch.zeros(n, dtype=torch.float64, device=gpu)
    h = torch.full((n,), 1e-9, dtype=torch.float64, device=gpu)
    status = torch.zeros(n, dtype=torch.int32, device=gpu)
    steps = torch.zeros(n, dtype=torch.int32, device=gpu)
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
        assert dT.max() <= dT_max and dY.max() <= dY_max


@pytest.mark.gpu
@pytest.mark.parametrize("jacobian,one_per_wave,method", [(j, one, method) for method in ("rodas4", "ode23s") for j, one in
                                                          (("automatic", False), ("differences", False), ("analytic", True),
                                                           ("differences", True))])
def test_reactors_through_ignition_match_scipy_bdf(gpu, jacobian, one_per_wave, method):
    d = np.load(os.path.join(GOLDEN, "reactor_handwritten.npz"))
    M = Mechanism(os.path.join(GOLDEN, "handwritten.yaml"))
    integ = Integrator(M, rtol=1e-8, atol_T=1e-6, atol_Y=1e-12, jacobian=jacobian, one_per_wave=one_per_wave, method=method)
    against_fixture(gpu, d, M, integ, *((2e-6, 1e-6) if method == "ode23s" else (1e-8, 1e-8)))
    assert np.all(d["states"][-1][0] > d["T0"] + 800)  # every reactor ignited


@pytest.mark.gpu
@pytest.mark.parametrize("method", ["ode23s", "rodas4"])
def test_large_mechanism_reactors_match_scipy_bdf(gpu, method):
    """The 53-species mechanism (state size 54: difference quotients, one reactor per wave) against the scipy BDF
    trajectories of tests/golden/reactor_gri53.npz (four fuel/O2/N2 mixtures to 1 ms), at ode23s's bars for both
    methods (measured at 0.1 ms: ΔT/T 4.9e-7 for ode23s, 4.8e-7 for RODAS4, ΔY 9.0e-8 for both: the fixture's own
    error on one reactor: a CPU run of RODAS4 at rtol 1e-11 differs from it by 4.4e-7 and 9.0e-8)."""
    from combustion_amd import synth
    d = np.load(os.path.join(GOLDEN, "reactor_gri53.npz"))
    M = Mechanism(synth.mechanism_path("gri-53"))
    integ = Integrator(M, rtol=1e-8, atol_T=1e-6, atol_Y=1e-12, method=method)
    against_fixture(gpu, d, M, integ, 2e-6, 1e-6)


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["handwritten", "h2o2-10"])
def test_jacobian_kernel_matches_the_interpreter(gpu, name):
    """The analytic Jacobian program of the integrator (cb_build_jacobian of cb_build_reactor's program over T and
    every Y_k) as the FP64 straight-line kernel, against the oracle's interpreter of the same program (whose
    derivatives tests/test_jacobian.py checks against closed forms and central differences): within 1e-10 of each
    row's largest entry (the kernel's exp and log are the device library's, the interpreter's the host's)."""
    import torch

    import oracle
    from combustion_amd import synth
    if name == "handwritten":
        d = np.load(os.path.join(GOLDEN, "reactor_handwritten.npz"))
        M = Mechanism(os.path.join(GOLDEN, "handwritten.yaml"))
        X = np.vstack([d["states"][:, 0].reshape(-1), np.repeat(d["P"][None], 3, 0).reshape(-1),
                       d["states"][:, 1:].transpose(1, 0, 2).reshape(M.K, -1)])
    else:
        M = Mechanism(synth.mechanism_path(name))
        T, P, Y = synth.config_states(name, 509, 21)
        X = np.vstack([T, P, Y])
    m = M.K + 1
    J = M.reactor().jacobian([0] + [2 + k for k in range(M.K)])
    ref = oracle.interpret(J.linear(), X)
    got = torch.stack(J.assemble(form="fp64")([torch.from_numpy(np.ascontiguousarray(r)).to(gpu) for r in X])).cpu().numpy()
    np.testing.assert_array_equal(np.isfinite(got), np.isfinite(ref))
    A, B = got[m:].reshape(m, m, -1), ref[m:].reshape(m, m, -1)
    scale = np.max(np.abs(B), axis=1, keepdims=True) + 1e-300
    err = np.abs(A - B) / scale
    f_err = np.abs(got[:m] - ref[:m]) / (np.max(np.abs(ref[:m]), axis=1, keepdims=True) + 1e-300)
    print(name, "Jacobian kernel against the interpreter: max error", err.max(), "right-hand side", f_err.max())
    assert err.max() <= 1e-10 and f_err.max() <= 1e-10


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


def integrate(gpu, M, y0, P, t_end, **flags):
    """Reactors from y0 [K+1][n] at pressures P [n] to t_end: (y, steps) on the host."""
    import torch
    n = y0.shape[1]
    y = torch.from_numpy(np.ascontiguousarray(y0)).to(gpu)
    t = torch.zeros(n, dtype=torch.float64, device=gpu)
    h = torch.full((n,), 1e-9, dtype=torch.float64, device=gpu)
    status = torch.zeros(n, dtype=torch.int32, device=gpu)
    steps = torch.zeros(n, dtype=torch.int32, device=gpu)
    taken = Integrator(M, rtol=1e-8, atol_T=1e-6, atol_Y=1e-12, **flags).advance(y, torch.from_numpy(np.ascontiguousarray(P)).to(gpu), t,
                                                                                 h, status, steps, t_end, max_steps=200000)
    torch.cuda.synchronize()
    assert torch.all(status == 1)
    assert taken >= int(steps.max().item())  # the batch's steps cover every reactor's, whichever stream it ran on
    return y.cpu().numpy(), steps.cpu().numpy()


@pytest.mark.gpu
@pytest.mark.parametrize("mechanism,jacobian,t_end,method", [("handwritten", "analytic", 1e-3, "ode23s"),
                                                             ("handwritten", "differences", 1e-3, "ode23s"),
                                                             ("gri-53", "differences", 1e-5, "ode23s"),
                                                             ("handwritten", "analytic", 1e-3, "rodas4"),
                                                             ("gri-53", "differences", 1e-5, "rodas4")])
def test_a_reactor_does_not_depend_on_its_batch(gpu, mechanism, jacobian, t_end, method):
    """Each reactor's steps are its own: integrated alone, or at another place in a batch of another size (another
    lane group of a packed wave, another block and XCD), a reactor ends in the same state to the bit after the same
    number of steps."""
    from combustion_amd import synth
    if mechanism == "handwritten":
        d = np.load(os.path.join(GOLDEN, "reactor_handwritten.npz"))
        M = Mechanism(os.path.join(GOLDEN, "handwritten.yaml"))
    else:
        d = np.load(os.path.join(GOLDEN, "reactor_gri53.npz"))
        M = Mechanism(synth.mechanism_path(mechanism))
    y0, P = np.vstack([d["T0"], d["Y0"]]), d["P"]
    y, steps = integrate(gpu, M, y0, P, t_end, jacobian=jacobian, method=method)
    one, one_steps = integrate(gpu, M, y0[:, 1:2], P[1:2], t_end, jacobian=jacobian, method=method)
    np.testing.assert_array_equal(one[:, 0], y[:, 1])
    assert one_steps[0] == steps[1]
    order = np.arange(len(P))[::-1]
    rev, rev_steps = integrate(gpu, M, np.tile(y0[:, order], 3), np.tile(P[order], 3), t_end, jacobian=jacobian, method=method)
    np.testing.assert_array_equal(rev[:, :len(P)], y[:, order])
    np.testing.assert_array_equal(rev_steps[:len(P)], steps[order])


@pytest.mark.gpu
def test_integrator_keeps_the_order_of_the_callers_stream(gpu):
    """advance() on a stream of the caller's (not the null stream): the states that the stream's earlier work (a few
    milliseconds of matrix products, then the copy of the initial states) leaves are the ones both halves of the batch
    start from, and the result equals the null-stream run to the bit."""
    import torch
    d = np.load(os.path.join(GOLDEN, "reactor_handwritten.npz"))
    M = Mechanism(os.path.join(GOLDEN, "handwritten.yaml"))
    y0, P0 = np.tile(np.vstack([d["T0"], d["Y0"]]), 3), np.tile(d["P"], 3)  # 36 reactors: parts of 24 and 12
    ref, ref_steps = integrate(gpu, M, y0, P0, 1e-3)
    n = len(P0)
    s = torch.cuda.Stream()
    src = torch.from_numpy(np.ascontiguousarray(y0)).to(gpu)
    P = torch.from_numpy(np.ascontiguousarray(P0)).to(gpu)
    a = torch.randn(4096, 4096, device=gpu)
    y = torch.zeros_like(src)
    t = torch.zeros(n, dtype=torch.float64, device=gpu)
    h = torch.full((n,), 1e-9, dtype=torch.float64, device=gpu)
    status = torch.zeros(n, dtype=torch.int32, device=gpu)
    steps = torch.zeros(n, dtype=torch.int32, device=gpu)
    integ = Integrator(M, rtol=1e-8, atol_T=1e-6, atol_Y=1e-12)
    s.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(s):
        for _ in range(8):
            a = a @ a / 64.0
        y.copy_(src)
        integ.advance(y, P, t, h, status, steps, 1e-3, max_steps=200000, stream=s.cuda_stream)
    torch.cuda.synchronize()
    assert torch.all(status == 1)
    np.testing.assert_array_equal(y.cpu().numpy(), ref)
    np.testing.assert_array_equal(steps.cpu().numpy(), ref_steps)


@pytest.mark.gpu
def test_mid_size_mechanism_packings_and_methods_agree(gpu):
    """The 10-species mechanism (state size 11: five reactors per wave; 16 copies per reactor for the difference
    quotients, four reactors to a block of copies), which no trajectory fixture covers: eight fuel/O2/N2 reactors
    (no radicals at the start) to 1 µs with RODAS4 on packed waves, with one reactor per wave, and with ode23s. The
    packings agree to the bit (the same arithmetic in either, the error norm summed in row order); at rtol 1e-8 the
    methods agree within 1e-6 (relative T, absolute Y; measured 3.4e-7 and 8.9e-8).
    (With the analytic Jacobian these reactors stall at steps below 1e-19 s: the mechanism's radical rates are so
    nonlinear at zero concentration that the exact Jacobian at the state misses the stiffness a trace of radicals
    brings, which the difference quotients' increments reach into; the automatic choice takes those above 8 rows.)"""
    from combustion_amd import reactor, synth
    M = Mechanism(synth.mechanism_path("h2o2-10"))
    y0, P = reactor.initial_states(M, 8, 5)
    y0 = np.ascontiguousarray(y0.T)
    runs = {}
    for name, flags in (("packed", {}), ("one per wave", {"one_per_wave": True}), ("ode23s", {"method": "ode23s"})):
        y, steps = integrate(gpu, M, y0, P, 1e-6, **flags)
        print(name, "steps per reactor", steps.min(), "..", steps.max(), "T", y[0].min(), "..", y[0].max(), flush=True)
        runs[name] = y
    np.testing.assert_array_equal(runs["packed"], runs["one per wave"])
    a, b = runs["packed"], runs["ode23s"]
    dT, dY = np.max(np.abs(a[0] - b[0]) / b[0]), np.max(np.abs(a[1:] - b[1:]))
    print("RODAS4 against ode23s: max |ΔT|/T", dT, "max |ΔY|", dY)
    assert dT <= 1e-6 and dY <= 1e-6

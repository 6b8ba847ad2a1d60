"""The reference's own program shapes and rate models through the HIP forms, against the oracle:

* the rate models the synthetic draws leave out (Lindemann falloff reaction.rs:153-158, Troe with a constant
  high-pressure limit :172-174, irreversible reactions :272) and the hand-written mechanism (single-range NASA-7,
  a polar species, inert argon) through every form;
* the reference-convention programs as device::assemble callers hand them over (perf.rs:14, cantera.rs:56-57,79):
  rates (reaction.rs:356-391, with T dT/dt and dV/dt), reactions_rates (reaction.rs:329-354) and
  properties::<4> (transport.rs:179-214), through the faithful straight-line form and the FP64 form;
* the transport section through the tiled MFMA kernel at K <= 64 (v_mfma_f64_16x16x4f64 in the FP64 form,
  v_mfma_f32_16x16x4f32 in the FP32-core form), which by default serves only larger mechanisms.

Bars: 1e-12 relative for the FP64 forms (ω and q floored by their term scales, dT/dt and dV/dt by theirs);
the mixed-precision forms ω within 1e-5 of the state floor and transport within 1e-3."""
import os

import numpy as np
import pytest

from combustion_amd import synth
from helpers import models_variant, relative_error, setup, setup_text, state_floor

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
R = 1.380649e-23 * 6.02214076e23
EXACT = ("straight", "fp64", "fp64+tiled")


def device(arrays, gpu):
    import torch
    return [torch.from_numpy(np.ascontiguousarray(x, np.float64)).to(gpu) for x in arrays]


def run(prog, form, arrays, gpu):
    import torch
    mod = prog.assemble(form=form)
    out = mod(device(arrays, gpu))
    torch.cuda.synchronize()
    return np.stack([o.cpu().numpy() for o in out])


def check(got, ref, floor, Ka, form, label):
    assert np.all(np.isfinite(got))
    if form in EXACT:
        err = relative_error(got, ref, floor)
        print(label, form, "max error", float(err.max()), flush=True)
        assert err.max() <= 1e-12
    else:
        err = relative_error(got, ref, state_floor(floor, Ka))
        print(label, form, "ω", float(err[:Ka].max()), "transport", float(err[Ka:].max()), flush=True)
        assert err[:Ka].max() <= 1e-5 and err[Ka:].max() <= 1e-3


@pytest.mark.gpu
@pytest.mark.parametrize("name,n", [("h2o2-10", 4096), ("gri-53", 4096)])
@pytest.mark.parametrize("form", ["straight", "fp64", "mixed", "fast"])
def test_reference_rate_models_through_every_form(name, n, form, gpu):
    M, mech, o = setup_text(models_variant(name))
    T, P, Y = synth.config_states(name, n, 61)
    ref, floor = o.batch(T, P, Y, threads=8)
    got = run(M.chemistry(), form, [T, P, *Y], gpu)
    check(got, ref, floor, M.active, form, f"models:{name}")


@pytest.mark.gpu
@pytest.mark.parametrize("form", ["straight", "fp64", "mixed", "fast"])
def test_handwritten_mechanism_through_every_form(form, gpu):
    M, mech, o = setup_text(open(os.path.join(ROOT, "tests", "golden", "handwritten.yaml")).read())
    T, P, Y = synth.states(M.K, 2048, 12, (300.0, 3000.0), (0.1, 50.0), 0.5)
    ref, floor = o.batch(T, P, Y, threads=8)
    got = run(M.chemistry(), form, [T, P, *Y], gpu)
    check(got, ref, floor, M.active, form, "handwritten")


def amounts(mech, T, P, Y, V):
    """The reference's inputs (amounts in a volume V, device.rs callers' convention) for states given as T, P, Y:
    n_k = C_k V with C_k = (P/RT) X_k."""
    n = Y / mech["W"][:, None]
    C = (P / R / T) * n / n.sum(axis=0)
    return C * V


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["h2o2-10", "gri-53"])
@pytest.mark.parametrize("form", ["straight", "fp64"])
def test_reference_rates_program(name, form, gpu):
    """reaction.rs:356-391: inputs [P/R, R/P, T, V, n_0..n_{K-2}], outputs [T dT/dt/T = dT/dt, dV/dt, ω]."""
    M, mech, o = setup(name)
    N = 2048
    T, P, Y = synth.config_states(name, N, 71)
    V = np.random.default_rng(71).uniform(0.25, 4.0, N)
    n = amounts(mech, T, P, Y, V)
    got = run(M.rates(), form, [P / R, R / P, T, V, *n[:-1]], gpu)
    ref, floor = np.zeros_like(got), np.zeros_like(got)
    for s in range(N):
        omega, dT, x = o.rhs_full(T[s], P[s], Y[:, s])
        ref[:, s] = np.concatenate([[dT, V[s] * x["dlnV"]], omega])
        floor[:, s] = np.concatenate([[x["dTdt_scale"], V[s] * x["dlnV_scale"]], x["omega_scale"]])
    err = relative_error(got, ref, floor)
    print(name, form, "max error dT/dt", float(err[0].max()), "dV/dt", float(err[1].max()), "ω", float(err[2:].max()), flush=True)
    assert err.max() <= 1e-12


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["h2o2-10", "gri-53"])
@pytest.mark.parametrize("form", ["straight", "fp64"])
def test_reference_reactions_rates_program(name, form, gpu):
    """reaction.rs:329-354: inputs [P/R, T, V, n_0..n_{K-2}], one rate of progress per reaction."""
    M, mech, o = setup(name)
    N = 1024
    T, P, Y = synth.config_states(name, N, 72)
    V = np.random.default_rng(72).uniform(0.25, 4.0, N)
    n = amounts(mech, T, P, Y, V)
    got = run(M.reactions_rates(), form, [P / R, T, V, *n[:-1]], gpu)
    ref, floor = np.zeros_like(got), np.zeros_like(got)
    for s in range(N):
        ref[:, s], floor[:, s] = o.reaction_rates(T[s], P[s], Y[:, s], scale=True)
    err = relative_error(got, ref, floor)
    print(name, form, "max error q", float(err.max()), flush=True)
    assert err.max() <= 1e-12


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["h2o2-10", "gri-53"])
@pytest.mark.parametrize("form", ["straight", "fp64"])
def test_reference_transport_program(name, form, gpu):
    """transport.rs:179-214 properties::<4>: inputs [Σx, T/T0, x_0..x_{K-2}] with x the concentrations,
    outputs [λ/λ0, μ/η0, ρD_k/η0]."""
    M, mech, o = setup(name)
    T0, eta0, lam0 = M.transport_scales()
    N = 4096
    T, P, Y = synth.config_states(name, N, 73)
    x = amounts(mech, T, P, Y, 1.0)
    got = run(M.transport(4, T0, eta0, lam0), form, [x.sum(axis=0), T / T0, *x[:-1]], gpu)
    ref, floor = o.batch(T, P, Y, rates=False, threads=8)
    rho = (mech["W"][:, None] * x).sum(axis=0)
    scaled = np.vstack([got[:1] * lam0, got[1:2] * eta0, got[2:] * eta0 / rho])
    err = relative_error(scaled, ref, floor)
    print(name, form, "max error λ", float(err[0].max()), "μ", float(err[1].max()), "D", float(err[2:].max()), flush=True)
    assert err.max() <= 1e-12


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["h2o2-10", "gri-53"])
@pytest.mark.parametrize("form", ["fp64+tiled", "fast+tiled"])
def test_tiled_transport_at_small_species_counts(name, form, gpu):
    """The MFMA transport kernel (Wilke sums as three K×K contractions, binary-diffusion polynomials as
    (16 pairs × 4)·(4 × 16 states) tiles) on mechanisms that the straight-line transport kernel serves by default:
    v_mfma_f64_16x16x4f64 to 1e-12 in the FP64 form, v_mfma_f32_16x16x4f32 within 1e-3 in the FP32-core form."""
    M, mech, o = setup(name)
    T, P, Y = synth.config_states(name, 4096, 74)
    ref, floor = o.batch(T, P, Y, threads=8)
    src = M.chemistry().hip_source(form)
    assert "transport_tiled" in src and ("#define F64 1" in src) == (form == "fp64+tiled")
    got = run(M.chemistry(), form, [T, P, *Y], gpu)
    check(got, ref, floor, M.active, form, name)


@pytest.mark.gpu
@pytest.mark.parametrize("spec", ["h2o2-10", "gri-53", "models:h2o2-10", "models:gri-53", "handwritten"])
def test_tiled_rates_at_small_species_counts(spec, gpu):
    """The table-driven FP64 rates kernel that serves mechanisms above 64 species (8 states × 8 reaction groups per
    wave, concentrations and clamped exp(±G/RT) in LDS, ω by LDS atomic adds) on mechanisms that the straight-line
    kernel serves by default, every rate model it takes among them (elementary, irreversible, three-body, Lindemann,
    Troe with and without a constant k∞): to 1e-12 of the oracle."""
    if spec == "handwritten":
        M, mech, o = setup_text(open(os.path.join(ROOT, "tests", "golden", "handwritten.yaml")).read())
        T, P, Y = synth.states(M.K, 2045, 13, (300.0, 3000.0), (0.1, 50.0), 0.5)
    else:
        variant, _, name = spec.rpartition(":")
        M, mech, o = setup_text(models_variant(name)) if variant else setup(name)
        T, P, Y = synth.config_states(name, 4093, 76)  # a ragged count: the last wave's states are partly absent
    ref, floor = o.batch(T, P, Y, threads=8)
    src = M.chemistry().hip_source("fp64+tiled-rates")
    assert "rates_tiled" in src and "rates_fp64" not in src
    got = run(M.chemistry(), "fp64+tiled-rates", [T, P, *Y], gpu)
    check(got, ref, floor, M.active, "fp64", spec)


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["h2o2-10", "gri-53"])
def test_exported_device_functions_on_the_gpu(name, gpu):
    """The nekRK export (cuda.rs:296-337: rates.c, conductivity.c, viscosity.c, diffusivity.c) compiled by hiprtc
    as it is written to disk (dfloat = double, __NEKRK_DEVICE__ a device function) and run on the GPU over SoA states,
    its arguments formed as cuda.rs:350-369 forms them: to 1e-12 of the oracle with fits of length 5."""
    M, mech, o = setup(name, D=5)
    T, P, Y = synth.config_states(name, 4096, 75)
    ref, floor = o.batch(T, P, Y, threads=8)
    got = run_module(M.export_module(5), [T, P, *Y], gpu)
    err = relative_error(got, ref, floor)
    Ka = M.active
    print(name, "export: max error ω", float(err[:Ka].max()), "transport", float(err[Ka:].max()), flush=True)
    assert err.max() <= 1e-12


def run_module(mod, arrays, gpu):
    import torch
    out = mod(device(arrays, gpu))
    torch.cuda.synchronize()
    return np.stack([o.cpu().numpy() for o in out])

"""The FP64 building blocks that the FP64 chemistry kernels inline instead of the device library's: the polynomial
exponential (Cody-Waite reduction, degree-10 polynomial, v_ldexp_f64) and the reciprocal by v_rcp_f64 and one Newton
step. Each runs as a one-instruction program in form 1 (the straight-line form that uses them) over dense and edge
inputs, against numpy's correctly rounded (to within an ulp) results: the bounds written here are what the
chemistry kernels' 1e-12 parity relies on, with three orders of magnitude to spare."""
import numpy as np
import pytest

from combustion_amd.native import Program

EXP, DIV = 10, 8  # ast::LOp


def run(program, inputs, gpu):
    import torch
    mod = program.assemble(form="fp64")
    out = mod([torch.from_numpy(np.ascontiguousarray(x)).to(gpu) for x in inputs])
    torch.cuda.synchronize()
    return out[0].cpu().numpy()


@pytest.mark.gpu
def test_polynomial_exponential(gpu):
    p = Program.from_linear({"op": [EXP], "dst": [1], "a": [0], "b": [0], "c": [0], "k": [0.0], "outputs": [1], "n_inputs": 1})
    assert "exp64(" in p.hip_source("fp64")
    rng = np.random.default_rng(5)
    edges = [-np.inf, -1e300, -1000., -746., -745.2, -745.1, -744.5, -720., -708.5, -708.39, -700., -30., -1., -1e-300, 0.,
             1e-300, 0.5, 1., 30., 700., 709.7, 709.78, 709.79, 710., 1000., 1e300, np.inf]
    x = np.concatenate([rng.uniform(-745., 709.7, 1 << 20), rng.uniform(-1., 1., 1 << 16), np.array(edges)])
    got = run(p, [x], gpu)
    with np.errstate(over="ignore"):
        ref = np.exp(x)
    normal = (ref >= np.finfo(np.float64).tiny) & np.isfinite(ref)
    rel = np.abs(got[normal] - ref[normal]) / ref[normal]
    print("exp64 max relative error (normal results)", rel.max(), flush=True)
    assert rel.max() < 6.7e-16  # three ulps: the polynomial's 4.3e-16 and the reduction's and Horner steps' rounding
    sub = ref < np.finfo(np.float64).tiny  # denormal and zero results: within two units of the smallest denormal
    assert np.all(np.abs(got[sub] - ref[sub]) <= 2 * 2.0 ** -1074)
    assert np.all(got[x >= 709.79] == np.inf) and np.all(got[x <= -746.] == 0.)
    assert got[np.flatnonzero(x == 0.)[0]] == 1.


@pytest.mark.gpu
def test_reciprocal_with_one_newton_step(gpu):
    p = Program.from_linear({"op": [DIV], "dst": [2], "a": [0], "b": [1], "c": [0], "k": [0.0], "outputs": [2], "n_inputs": 2})
    assert "rcp64(" in p.hip_source("fp64")
    rng = np.random.default_rng(6)
    n = 1 << 20
    a = rng.uniform(-1., 1., n) * 2.0 ** rng.integers(-300, 300, n)
    b = rng.uniform(0.5, 1., n) * 2.0 ** rng.integers(-300, 300, n) * rng.choice([-1., 1.], n)
    got = run(p, [a, b], gpu)
    ref = a / b
    rel = np.abs(got - ref) / np.abs(ref)
    print("a * rcp64(b) max relative error", rel.max(), flush=True)
    assert rel.max() < 5e-15

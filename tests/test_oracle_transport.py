"""Pins the transport restatements (oracle/fits.py, oracle/chem.c) and the product's species physics.

Known answers from the literature (N2 viscosity, conductivity, N2-O2 diffusion at 300 K) pin the
Chapman-Enskog formulas and collision-integral interpolation; two independent implementations (C++ in the
product, numpy in the oracle) must agree to rounding; the mixture rules reduce to closed forms for pure and
binary mixtures.
"""
import json
import math
import os
import re

import numpy as np
import pytest

import oracle
from combustion_amd import Mechanism, synth
from oracle import fits as ofits
from oracle import mechanism as om

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
# GRI-Mech 3.0 published NASA-7 coefficients and transport parameters (public data)
N2_NASA = ([3.298677, 0.0014082404, -3.963222e-06, 5.641515e-09, -2.444854e-12, -1020.8999, 3.950372],
           [2.92664, 0.0014879768, -5.68476e-07, 1.0097038e-10, -6.753351e-15, -922.7977, 5.980528])
O2_NASA = ([3.78245636, -0.00299673416, 9.84730201e-06, -9.68129509e-09, 3.24372837e-12, -1063.94356, 3.65767573],
           [3.28253784, 0.00148308754, -7.57966669e-07, 2.09470555e-10, -2.16717794e-14, -1088.45772, 5.45323129])
H2O_NASA = ([4.19864056, -0.0020364341, 6.52040211e-06, -5.48797062e-09, 1.77197817e-12, -30293.7267, -0.849032208],
            [3.03399249, 0.00217691804, -1.64072518e-07, -9.7041987e-11, 1.68200992e-14, -30004.2971, 4.9667701])


def real_yaml():
    d = lambda a: "[" + ", ".join(repr(x) for x in a) + "]"
    sp = lambda name, comp, nasa, tr: (f"- name: {name}\n  composition: {comp}\n  thermo:\n    model: NASA7\n"
                                       f"    temperature-ranges: [200.0, 1000.0, 3500.0]\n    data:\n    - {d(nasa[0])}\n"
                                       f"    - {d(nasa[1])}\n  transport:\n{tr}")
    return ("units: {length: cm, time: s, quantity: mol, activation-energy: cal/mol}\nspecies:\n" +
            sp("O2", "{O: 2}", O2_NASA, "    model: gas\n    geometry: linear\n    well-depth: 107.4\n    diameter: 3.458\n"
                                        "    polarizability: 1.6\n    rotational-relaxation: 3.8\n") +
            sp("H2O", "{H: 2, O: 1}", H2O_NASA, "    model: gas\n    geometry: nonlinear\n    well-depth: 572.4\n    diameter: 2.605\n"
                                                "    dipole: 1.844\n    rotational-relaxation: 4.0\n") +
            sp("H2O2", "{H: 2, O: 2}", H2O_NASA, "    model: gas\n    geometry: nonlinear\n    well-depth: 107.4\n    diameter: 3.458\n"
                                                 "    rotational-relaxation: 3.8\n") +
            sp("N2", "{N: 2}", N2_NASA, "    model: gas\n    geometry: linear\n    well-depth: 97.53\n    diameter: 3.621\n"
                                        "    polarizability: 1.76\n    rotational-relaxation: 4.0\n") +
            "reactions:\n- equation: 2 H2O + O2 <=> 2 H2O2\n  rate-constant: {A: 1.0e+10, b: 0.0, Ea: 1000.0}\n")


@pytest.fixture(scope="module")
def real():
    text = real_yaml()
    return Mechanism(text=text), om.load(text=text)


def test_known_viscosity_diffusion_conductivity(real):
    M, m = real
    i = {n: k for k, n in enumerate(M.species_names)}
    eta = M.transport_species("viscosity", i["N2"], i["N2"], 300.0)
    assert eta == pytest.approx(1.78e-5, rel=0.02)  # N2, 300 K: 17.8 µPa s
    PD = M.transport_species("diffusivity", i["N2"], i["O2"], 300.0)
    assert PD / 101325.0 == pytest.approx(2.0e-5, rel=0.06)  # N2-O2, 300 K, 1 atm: 0.20 cm²/s
    lam = M.transport_species("conductivity", i["N2"], i["N2"], 300.0)
    assert lam == pytest.approx(0.0259, rel=0.06)  # N2, 300 K: 25.9 mW/m/K


def test_collision_integrals_reproduce_table_nodes(real):
    """At a tabulated T* (no dipole), the quadratic interpolation returns the table value."""
    M, m = real
    i = {n: k for k, n in enumerate(M.species_names)}
    k = i["N2"]
    for T_star, row in [(2.0, 14), (3.0, 16), (10.0, 24)]:
        T = T_star * 97.53
        assert M.transport_species("omega22", k, k, T) == pytest.approx(ofits.OMEGA22[row][0], rel=1e-12)


def test_product_physics_matches_numpy_restatement(real):
    """C++ (product) and numpy (oracle) restatements of transport.rs:57-117, with a polar species (H2O)
    exercising the polar/non-polar correction and the δ* polynomials."""
    M, m = real
    ph = ofits.Physics(m)
    for T in (300.0, 1234.5, 2900.0):
        for a in range(M.K):
            assert M.transport_species("viscosity", a, a, T) == pytest.approx(ph.viscosity(a, T), rel=1e-12)
            assert M.transport_species("conductivity", a, a, T) == pytest.approx(ph.conductivity(a, T), rel=1e-12)
            for b in range(M.K):
                assert M.transport_species("diffusivity", a, b, T) == pytest.approx(ph.diffusivity(a, b, T), rel=1e-11)
                assert M.transport_species("omega11", a, b, T) == pytest.approx(ph.omega11(a, b, T), rel=1e-11)


@pytest.mark.parametrize("name", ["h2o2-10", "gri-53"])
def test_product_fits_match_numpy_lstsq(name):
    """Polynomials::<4>::new (transport.rs:125-136): Householder QR (product) vs SVD least squares (oracle)."""
    M = Mechanism(synth.mechanism_path(name))
    m = om.load(synth.mechanism_path(name))
    c, v, d = M.transport_fits(4, 1000.0)
    ks = range(M.K) if M.K <= 10 else [0, 7, 23, 52]
    oc, ov, od = ofits.fits({**m, "K": M.K}, 4, 1000.0) if M.K <= 10 else (None, None, None)
    if oc is not None:
        np.testing.assert_allclose(c, oc, rtol=1e-9, atol=1e-12 * np.abs(oc).max())
        np.testing.assert_allclose(v, ov, rtol=1e-9, atol=1e-12 * np.abs(ov).max())
        np.testing.assert_allclose(d, od, rtol=1e-9, atol=1e-12 * np.abs(od).max())
    # the fits follow the physics to the accuracy a cubic in ln T reaches (the synthetic species' sigmoid cp makes λ the
    # least polynomial: up to ~10 % at 3000 K); this guards against gross errors, the QR/SVD agreement above pins the fit
    ph = ofits.Physics(m)
    T = np.array([300.0, 1000.0, 2000.0, 3000.0])
    x = np.log(T / 1000.0)
    for k in ks:
        lam = np.polyval(c[k][::-1], x) * np.sqrt(T)
        ref = np.array([ph.conductivity(k, t) for t in T])
        assert np.max(np.abs(lam / ref - 1)) < 0.15
        eta = np.polyval(v[k][::-1], x) ** 2 * np.sqrt(T)
        ref = np.array([ph.viscosity(k, t) for t in T])
        assert np.max(np.abs(eta / ref - 1)) < 0.05


def binary_state(m, T, P, X1, k1, k2):
    X = np.zeros(m["K"])
    X[k1], X[k2] = X1, 1 - X1
    Y = X * m["W"] / np.dot(X, m["W"])
    return T, P, Y, X


def test_mixture_rules_reduce_to_closed_forms():
    """Pure species: μ = η_k, λ = λ_k. Binary: Wilke and D_1m = (1 - Y_1) D_12 / X_2."""
    path = synth.mechanism_path("h2o2-10")
    M = Mechanism(path)
    m = om.load(path)
    c, v, d = M.transport_fits(4, 1000.0)
    o = oracle.Oracle(m, (c, v, d), 1000.0)
    K = m["K"]
    T, P = 1500.0, 2e5
    l = math.log(T / 1000.0)
    pw = np.array([1, l, l * l, l ** 3])
    eta = (v @ pw) ** 2 * math.sqrt(T)
    lam = (c @ pw) * math.sqrt(T)
    # pure bulk species (last)
    Y = np.zeros(K)
    Y[-1] = 1.0
    out, _ = o.batch([T], [P], Y[:, None], rates=False)
    assert out[0, 0] == pytest.approx(lam[-1], rel=1e-13)
    assert out[1, 0] == pytest.approx(eta[-1], rel=1e-13)
    # binary mixture of species 0 and the bulk
    T_, P_, Y, X = binary_state(m, T, P, 0.3, 0, K - 1)
    out, _ = o.batch([T], [P], Y[:, None], rates=False)
    W = m["W"]
    phi = lambda a, b: (1 + math.sqrt(eta[a] / eta[b]) * (W[b] / W[a]) ** 0.25) ** 2 / math.sqrt(8 * (1 + W[a] / W[b]))
    a, b = 0, K - 1
    mu = X[a] * eta[a] / (X[a] + X[b] * phi(a, b)) + X[b] * eta[b] / (X[a] * phi(b, a) + X[b])
    assert out[1, 0] == pytest.approx(mu, rel=1e-13)
    assert out[0, 0] == pytest.approx(0.5 * (X[a] * lam[a] + X[b] * lam[b] + 1 / (X[a] / lam[a] + X[b] / lam[b])), rel=1e-13)
    D12 = (d[b, a] @ pw) * T ** 1.5 / P
    assert out[2 + a, 0] == pytest.approx((1 - Y[a]) * D12 / X[b], rel=1e-12)


def test_collision_table_data_matches_header_and_reference():
    """combustion_amd/data/collision_integrals.json, the C++ header and (when present) the reference table agree."""
    data = json.load(open(os.path.join(ROOT, "combustion_amd", "data", "collision_integrals.json")))
    header = open(os.path.join(ROOT, "combustion_amd", "csrc", "collision_integrals.hpp")).read()
    nums = lambda s: [float(x) for x in re.findall(r"[-+]?\d*\.?\d+(?:e[-+]?\d+)?", s)]
    om22 = header[header.index("omega22[37][8] = {") + 18:header.index("};", header.index("omega22[37][8]"))]
    assert nums(om22) == [x for row in data["omega22"]["rows"] for x in row]
    ast = header[header.index("a_star[39][8] = {") + 17:header.index("};", header.index("a_star[39][8]"))]
    assert nums(ast) == [x for row in data["a_star"]["rows"] for x in row]
    ref = "/root/reference/transport/collision_integrals.rs"
    if not os.path.exists(ref):
        pytest.skip("reference tree not mounted")
    src = open(ref, encoding="utf-8").read()
    body = src[src.index("=", src.index("Ω⃰22: [[f64")) + 1:src.index("\n];", src.index("Ω⃰22: [[f64"))]
    assert nums(body) == [x for row in data["omega22"]["rows"] for x in row]
    body = src[src.index("=", src.index("A⃰: [[f64")) + 1:src.index("\n];", src.index("A⃰: [[f64"))]
    assert nums(body) == [x for row in data["a_star"]["rows"] for x in row]

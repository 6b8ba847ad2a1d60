"""Cantera-YAML front end (yaml.rs:40-171, lib.rs:22-140): the product's C++ parser against the Python
restatement in oracle/mechanism.py, on the drawn mechanisms and on a hand-written file exercising the
supported features; plus the error behaviour of the C ABI."""
import math
import os
import re

import numpy as np
import pytest

import combustion_amd
from combustion_amd import Error, Mechanism, native, synth
from oracle import mechanism as om

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

HANDWRITTEN = """\
description: |-
  A hand-written mechanism: block scalars, comments # not a comment inside a block scalar
  flow and block collections, multi-line flow sequences.
generator: test  # trailing comment
units: {length: cm, time: s, quantity: mol, activation-energy: cal/mol}
phases:
- name: gas
  elements: [O, H, N]
  species: [H2, O2, OH, "H2O", N2, AR]
species:
- name: H2
  composition: {H: 2}
  thermo:
    model: NASA7
    temperature-ranges: [200.0, 1000.0, 3500.0]
    data:
    - [2.34433112, 7.98052075e-03, -1.9478151e-05, 2.01572094e-08, -7.37611761e-12,
      -917.935173, 0.683010238]
    - [3.3372792, -4.94024731e-05, 4.99456778e-07, -1.79566394e-10, 2.00255376e-14,
      -950.158922, -3.20502331]
    note: TPIS78
  transport:
    model: gas
    geometry: linear
    well-depth: 38.0
    diameter: 2.92
    polarizability: 0.79
    rotational-relaxation: 280.0
- name: O2
  composition: {O: 2}
  thermo:
    model: NASA7
    temperature-ranges: [200.0, 1000.0, 3500.0]
    data:
    - [3.78245636, -2.99673416e-03, 9.84730201e-06, -9.68129509e-09, 3.24372837e-12, -1063.94356, 3.65767573]
    - [3.28253784, 1.48308754e-03, -7.57966669e-07, 2.09470555e-10, -2.16717794e-14, -1088.45772, 5.45323129]
  transport: {model: gas, geometry: linear, well-depth: 107.4, diameter: 3.458, polarizability: 1.6, rotational-relaxation: 3.8}
- name: OH
  composition: {O: 1, H: 1}
  thermo:
    model: NASA7
    temperature-ranges: [200.0, 3500.0]
    data:
    - [3.99201543, -2.40131752e-03, 4.61793841e-06, -3.88113333e-09, 1.3641147e-12, 3615.08056, -0.103925458]
  transport:
    model: gas
    geometry: linear
    well-depth: 80.0
    diameter: 2.75
- name: H2O
  composition: {H: 2, O: 1}
  thermo:
    model: NASA7
    temperature-ranges: [200.0, 1000.0, 3500.0]
    data:
    - [4.19864056, -2.0364341e-03, 6.52040211e-06, -5.48797062e-09, 1.77197817e-12, -3.02937267e+04, -0.849032208]
    - [4.19864056, -2.0364341e-03, 6.52040211e-06, -5.48797062e-09, 1.77197817e-12, -3.02937267e+04, -0.849032209]
  transport:
    model: gas
    geometry: nonlinear
    well-depth: 572.4
    diameter: 2.605
    dipole: 1.844
    rotational-relaxation: 4.0
- name: AR
  composition: {Ar: 1}
  thermo:
    model: NASA7
    temperature-ranges: [200.0, 1000.0, 6000.0]
    data:
    - [2.5, 0.0, 0.0, 0.0, 0.0, -745.375, 4.366]
    - [2.5, 0.0, 0.0, 0.0, 0.0, -745.375, 4.366]
  transport: {model: gas, geometry: atom, well-depth: 136.5, diameter: 3.33}
- name: N2
  composition: {N: 2}
  thermo:
    model: NASA7
    temperature-ranges: [300.0, 1000.0, 5000.0]
    data:
    - [3.298677, 1.4082404e-03, -3.963222e-06, 5.641515e-09, -2.444854e-12, -1020.8999, 3.950372]
    - [2.92664, 1.4879768e-03, -5.68476e-07, 1.0097038e-10, -6.753351e-15, -922.7977, 5.980528]
  transport:
    model: gas
    geometry: linear
    well-depth: 97.53
    diameter: 3.621
reactions:
- equation: 2 OH (+M) <=> H2O + O2 (+M)  # a Lindemann falloff without a Troe block
  type: falloff
  low-P-rate-constant: {A: 2.3e+18, b: -0.9, Ea: -1700.0}
  high-P-rate-constant: {A: 7.4e+13, b: -0.37, Ea: 0.0}
  efficiencies: {AR: 0.7, H2: 2.0, H2O: 6.0}
- equation: H2 + O2 => 2 OH
  rate-constant: {A: 1.7e+13, b: 0.0, Ea: 47780.0}
- equation: 2 OH + M <=> H2 + O2 + M
  type: three-body
  rate-constant: {A: 1.2e+17, b: -1.0, Ea: 0.0}
  efficiencies: {AR: 0.83, H2: 2.4, H2O: 15.4}
- equation: OH + H2 (+M) <=> H2O + 0.5 H2 (+M)
  type: falloff
  low-P-rate-constant: {A: 1.0e+20, b: 0.0, Ea: 0.0}
  high-P-rate-constant: {A: 1.0e+13, b: 0.0, Ea: 0.0}
  Troe: {A: 0.7346, T3: 94.0, T1: 1756.0}
"""


def without_fractional(text):
    return text.replace("- equation: OH + H2 (+M) <=> H2O + 0.5 H2 (+M)", "- equation: 2 OH + H2 (+M) <=> H2O + H2O2 (+M)").replace(
        "species: [H2, O2, OH, \"H2O\", N2, AR]", "species: [H2, O2, OH, H2O, N2, AR]")


def test_fractional_coefficients_are_rejected():
    with pytest.raises(Error) as e:
        Mechanism(text=HANDWRITTEN)
    assert e.value.kind == "parse"


def handwritten():
    text = without_fractional(HANDWRITTEN).replace("2 OH + H2 (+M) <=> H2O + H2O2 (+M)", "OH + H2 (+M) <=> H2O + H (+M)")
    # H is not declared: use a declared species instead
    return text.replace("OH + H2 (+M) <=> H2O + H (+M)", "2 OH + H2 (+M) <=> 2 H2O (+M)")


def compare(product, restated):
    t = product.tables
    assert [s["name"] for s in t["species"]] == restated["names"]
    assert t["active"] == restated["Ka"]
    np.testing.assert_array_equal(np.array([s["molar_mass"] for s in t["species"]]), restated["W"])
    split = np.array([s["temperature_split"] for s in t["species"]])
    np.testing.assert_array_equal(np.isnan(split), np.isnan(restated["split"]))
    np.testing.assert_array_equal(split[~np.isnan(split)], restated["split"][~np.isnan(restated["split"])])
    np.testing.assert_array_equal(np.array([s["pieces"] for s in t["species"]]), restated["nasa"])
    for key, mine in [("diameter", "diameter"), ("well_depth_J", "well_depth_J"), ("polarizability", "polarizability"),
                      ("permanent_dipole_moment", "dipole"), ("rotational_relaxation", "rotational_relaxation"),
                      ("internal_degrees_of_freedom", "internal_dof")]:
        np.testing.assert_allclose([s[key] for s in t["species"]], restated[mine], rtol=1e-15, atol=0)
    models = {"elementary": 0, "irreversible": 1, "three-body": 2, "pressure-modification": 3, "troe": 4, "sri": 5}
    for r, x in enumerate(t["reactions"]):
        assert models[x["model"]] == restated["model"][r]
        np.testing.assert_array_equal(x["reactants"], restated["reactants"][r])
        np.testing.assert_array_equal(x["products"], restated["products"][r])
        np.testing.assert_allclose(x["rate_constant"], restated["rate"][r], rtol=1e-15)
        if x["model"] in ("pressure-modification", "troe"):
            np.testing.assert_allclose(x["k0"], restated["k0"][r], rtol=1e-15)
        if x["model"] == "troe":
            np.testing.assert_array_equal(x["troe"], restated["troe"][r])
        if x["efficiencies"]:
            np.testing.assert_array_equal(x["efficiencies"], restated["eff"][r])


@pytest.mark.parametrize("name", ["h2o2-10", "gri-53", "large-160"])
def test_parser_matches_restatement_on_drawn_mechanisms(name):
    path = synth.mechanism_path(name)
    compare(Mechanism(path), om.load(path))


def test_parser_matches_restatement_on_handwritten_file():
    text = handwritten()
    M = Mechanism(text=text)
    m = om.load(text=text)
    compare(M, m)
    t = M.tables
    names = [s["name"] for s in t["species"]]
    # active species first (yaml.rs:159-162): AR and N2 never react
    assert names[-2:] == ["AR", "N2"] and M.active == 4
    # duplicate NASA pieces merged into one range (yaml.rs:64-77), single range -> NaN split (lib.rs:45)
    assert math.isnan(t["species"][names.index("H2O")]["temperature_split"])
    assert math.isnan(t["species"][names.index("OH")]["temperature_split"])
    # Ea in cal/mol -> K; A in cm-mol-s with the reaction order (+1 for three-body and low-pressure limits)
    r = t["reactions"]
    assert r[1]["model"] == "irreversible"
    assert r[1]["rate_constant"][2] == pytest.approx(47780.0 * 4.184 / (1.380649e-23 * 6.02214076e23), rel=1e-15)
    assert r[0]["model"] == "pressure-modification" and r[0]["k0"][0] == pytest.approx(2.3e18 * 1e-12, rel=1e-15)
    assert r[2]["model"] == "three-body" and r[2]["rate_constant"][0] == pytest.approx(1.2e17 * 1e-12, rel=1e-15)
    assert r[3]["model"] == "troe" and r[3]["troe"][3] == math.inf
    # molar masses from atomic weights (lib.rs:22-24)
    assert t["species"][names.index("H2O")]["molar_mass"] == pytest.approx((2 * 1.008 + 15.999) / 1e3, rel=1e-15)


def test_handwritten_program_matches_oracle():
    """The hand-written mechanism (Lindemann, irreversible, three-body, Troe without T2, single-range and merged NASA,
    a polar species) through the product's program and the closed-form oracle."""
    import oracle
    from helpers import relative_error
    text = handwritten()
    M = Mechanism(text=text)
    m = om.load(text=text)
    fits = M.transport_fits(4, 1000.0)
    o = oracle.Oracle(m, fits, 1000.0)
    T, P, Y = synth.states(M.K, 256, 12, (300.0, 3000.0), (0.1, 50.0), 0.5)
    ref, floor = o.batch(T, P, Y)
    got = oracle.interpret(M.chemistry().linear(), np.vstack([T, P, Y]))
    assert relative_error(got, ref, floor).max() <= 1e-12


@pytest.mark.parametrize("text,kind", [
    ("units: {length: m, time: s, quantity: mol, activation-energy: K}\nspecies: []\nreactions: []\n", "parse"),
    ("units: {length: cm, time: s, quantity: mol, activation-energy: K}\nspecies: [\n", "parse"),
    ("units: {length: cm, time: s, quantity: mol, activation-energy: K}\nspecies:\n- name: A\n  composition: {Xx: 1}\n", "parse"),
])
def test_malformed_files_report_parse_errors(text, kind):
    with pytest.raises(Error) as e:
        Mechanism(text=text)
    assert e.value.kind == kind


def test_sri_falloff_is_parsed_and_built():
    """reaction.rs:194-203 leaves SRI unimplemented; this build parses it (yaml.rs:140-146) and builds it
    (tests/test_sri.py checks the values)."""
    text = handwritten().replace("  Troe: {A: 0.7346, T3: 94.0, T1: 1756.0}", "  SRI: {A: 1.1, B: 700.0, C: 1234.0}")
    M = Mechanism(text=text)
    assert M.tables["reactions"][3]["model"] == "sri"
    assert M.chemistry().n_outputs == M.active + 2 + M.K


def test_unknown_species_in_equation():
    text = handwritten().replace("H2 + O2 => 2 OH", "H2 + O2 => 2 OX")
    with pytest.raises(Error) as e:
        Mechanism(text=text)
    assert e.value.kind == "parse"


def test_header_symbols_are_exported():
    header = open(os.path.join(ROOT, "include", "combustion_amd.h")).read()
    declared = set(re.findall(r"\b(cb_[a-z0-9_]+)\s*\(", header))
    lib = native.lib()
    missing = [s for s in declared if not hasattr(lib, s)]
    assert declared and not missing
    assert set(native.SIGNATURES) == declared


def test_null_arguments_return_status_not_abort():
    lib = native.lib()
    import ctypes as C
    assert lib.cb_mechanism_load_yaml(None, None) == -1
    assert b"null" in lib.cb_last_error()
    assert lib.cb_mechanism_load_yaml(b"/nonexistent.yaml", C.byref(C.c_void_p())) == -1
    assert lib.cb_hip_call(None, None, 0, 0, None, 0, 0, None, None) == -1


def test_program_sizes_and_op_counts():
    M = Mechanism(synth.mechanism_path("gri-53"))
    p = M.chemistry()
    assert (p.n_inputs, p.n_outputs) == (2 + 53, 52 + 2 + 53)
    c = p.op_counts()
    assert c["exp"] > 325 and c["div"] > 1378 and c["weighted"] == c["alu"] + 20 * (c["exp"] + c["ln"]) + 8 * (c["div"] + c["sqrt"])
    r = M.rates()
    assert (r.n_inputs, r.n_outputs) == (4 + 52, 2 + 52)

"""Pressure-dependent Arrhenius tables with several entries (PLOG; the reference asserts a single entry,
yaml.rs:120,149): parsing, the oracle against the closed forms of log-linear interpolation in ln P, and the
product's program against the oracle."""
import re

import numpy as np
import pytest

import oracle
from combustion_amd import Error, Mechanism, synth
from helpers import plog_variant, relative_error, setup_text

ATM = 101325.0


@pytest.mark.parametrize("name,n", [("h2o2-10", 2048), ("gri-53", 256)])
def test_plog_program_matches_oracle(name, n):
    text = plog_variant(name)
    M, mech, o = setup_text(text)
    assert sum(len(p) > 1 for p in mech["plog"]) >= 5
    T, _, Y = synth.config_states(name, n, 25)
    P = ATM * np.exp(np.random.default_rng(25).uniform(np.log(0.5), np.log(60.0), n))  # below, inside and above 2-30 atm
    P[:3] = [2 * ATM, 8 * ATM, 30 * ATM]  # on the table's pressures
    ref, floor = o.batch(T, P, Y, threads=4)
    got = oracle.interpret(M.chemistry().linear(), np.vstack([T, P, Y]))
    assert relative_error(got, ref, floor).max() <= 1e-12


def first_elementary(text):
    m = re.search(r"(- equation: [^\n]*<=>[^\n(]*)\n  rate-constant: \{A: ([^,]*), b: ([^,]*), Ea: ([^}]*)\}", text)
    return m, float(m.group(2)), float(m.group(3)), float(m.group(4))


def test_plog_interpolation_closed_forms():
    """Two entries k1 (1 atm) and k2 (10 atm): the rate of progress equals k1's at and below 1 atm, k2's at and
    above 10 atm, and the geometric mean of the two at √10 atm (ln k linear in ln P)."""
    base = open(synth.mechanism_path("h2o2-10")).read()
    m, A, b, Ea = first_elementary(base)
    k1, k2 = (A, b, Ea), (3.0 * A, b + 0.2, 0.9 * Ea)

    def with_rate(entries):
        if len(entries) == 1:
            block = f"  rate-constant: {{A: {entries[0][1][0]!r}, b: {entries[0][1][1]!r}, Ea: {entries[0][1][2]!r}}}"
        else:
            block = "  type: pressure-dependent-Arrhenius\n  rate-constants:\n" + "\n".join(
                f"  - {{P: {P}, A: {k[0]!r}, b: {k[1]!r}, Ea: {k[2]!r}}}" for P, k in entries)
        return base[:m.start()] + m.group(1) + "\n" + block + base[m.end():]

    texts = {"plog": with_rate([("1.0 atm", k1), ("10.0 atm", k2)]), "k1": with_rate([(None, k1)]), "k2": with_rate([(None, k2)])}
    n = 8
    T, _, Y = synth.config_states("h2o2-10", n, 26)
    r = next(i for i, e in enumerate(setup_text(texts["plog"])[1]["plog"]) if e)
    q = {}
    for key, text in texts.items():
        _, mech, o = setup_text(text)
        for P in (0.5, 1.0, np.sqrt(10.0), 10.0, 20.0):
            q[key, P] = np.array([o.reaction_rates(T[s], P * ATM, Y[:, s])[r] for s in range(n)])
    for P in (0.5, 1.0):
        np.testing.assert_allclose(q["plog", P], q["k1", P], rtol=1e-13)
    for P in (10.0, 20.0):
        np.testing.assert_allclose(q["plog", P], q["k2", P], rtol=1e-13)
    P = np.sqrt(10.0)
    np.testing.assert_allclose(q["plog", P], np.sign(q["k1", P]) * np.sqrt(q["k1", P] * q["k2", P]), rtol=1e-12)


def test_plog_parsing_units_order_and_sums():
    """Pressures in Pa (bare number), atm and torr; entries sorted by pressure; equal pressures kept (summed by the
    rate); the MFMA rates kernel reports the table unsupported."""
    base = open(synth.mechanism_path("h2o2-10")).read()
    m, A, b, Ea = first_elementary(base)
    block = ("  type: pressure-dependent-Arrhenius\n  rate-constants:\n"
             f"  - {{P: 760 torr, A: {A!r}, b: {b!r}, Ea: {Ea!r}}}\n"
             f"  - {{P: 10132.5, A: {2 * A!r}, b: {b!r}, Ea: {Ea!r}}}\n"
             f"  - {{P: 1.0 atm, A: {3 * A!r}, b: {b!r}, Ea: {Ea!r}}}")
    text = base[:m.start()] + m.group(1) + "\n" + block + base[m.end():]
    M = Mechanism(text=text)
    plog = next(x["plog"] for x in M.tables["reactions"] if x["plog"])
    assert [row[0] for row in plog] == pytest.approx([10132.5, 101325.0, 101325.0], rel=1e-15)
    assert [row[1] / plog[0][1] for row in plog] == pytest.approx([1.0, 0.5, 1.5], rel=1e-15)  # stable order at equal P
    _, mech, _ = setup_text(text)
    assert [row[0] for row in next(p for p in mech["plog"] if p)] == pytest.approx([10132.5, 101325.0, 101325.0], rel=1e-15)
    with pytest.raises(Error) as e:
        M.chemistry().hip_source(form="mixed")
    assert e.value.kind == "unsupported"

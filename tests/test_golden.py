"""Golden fixtures (tests/golden/*.npz, written by tests/golden/make_golden.py): the oracle reproduces them bit
for bit, and the product's program, interpreted in FP64 (ast/interpret.rs semantics), matches them to 1e-12."""
import os

import numpy as np
import pytest

import oracle
from combustion_amd import synth
from helpers import relative_error, setup

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


@pytest.mark.parametrize("name", ["h2o2-10", "gri-53", "large-160"])
def test_oracle_reproduces_fixture(name):
    g = np.load(os.path.join(HERE, f"{name}.npz"))
    M, mech, o = setup(name)
    out, floor = o.batch(g["T"], g["P"], g["Y"])
    np.testing.assert_array_equal(out, g["out"])
    np.testing.assert_array_equal(floor, g["floor"])


@pytest.mark.parametrize("name", ["h2o2-10", "gri-53", "large-160"])
def test_program_matches_fixture(name):
    g = np.load(os.path.join(HERE, f"{name}.npz"))
    M, mech, o = setup(name)
    got = oracle.interpret(M.chemistry().linear(), np.vstack([g["T"], g["P"], g["Y"]]))
    assert relative_error(got, g["out"], g["floor"]).max() <= 1e-12


def test_fixture_inputs_follow_the_generator():
    g = np.load(os.path.join(HERE, "h2o2-10.npz"))
    T, P, Y = synth.config_states("h2o2-10", len(g["T"]), int(g["seed"]))
    np.testing.assert_array_equal(T, g["T"])
    np.testing.assert_array_equal(Y, g["Y"])

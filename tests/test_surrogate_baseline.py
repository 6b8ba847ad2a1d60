"""The surrogate's O(K·M) baselines (combustion_amd.surrogate.log_baseline) on the CPU: with every species major
they are the mixture-averaged rules themselves and must equal the oracle's transport; with fewer major species the
error vanishes for states whose other species are absent."""
import numpy as np

from combustion_amd import surrogate, synth
from helpers import relative_error, setup


def test_all_species_major_is_the_oracle_transport():
    M, mech, o = setup("gri-53")
    T, P, Y = synth.config_states("gri-53", 512, 8)
    ref, floor = o.batch(T, P, Y, rates=False)
    major = surrogate.choose_major(M, Y, M.K)
    assert sorted(major.tolist()) == list(range(M.K))
    got = np.exp(surrogate.log_baseline(M, T, P, Y, major)).T
    assert relative_error(got, ref, floor).max() < 1e-12


def test_absent_minor_species_cost_nothing():
    """States made of 6 species only: the baseline over those 6 (plus the bulk) is exact, and a species set that
    misses one of them is not."""
    M, mech, o = setup("gri-53")
    T, P, Y = synth.config_states("gri-53", 256, 9)
    keep = np.array([M.K - 1, 3, 7, 12, 20, 31])
    Y = np.where(np.isin(np.arange(M.K), keep)[:, None], Y, 0.0)
    Y /= Y.sum(0)
    ref, floor = o.batch(T, P, Y, rates=False)
    exact = np.exp(surrogate.log_baseline(M, T, P, Y, keep.astype(np.int32))).T
    assert relative_error(exact, ref, floor).max() < 1e-12
    partial = np.exp(surrogate.log_baseline(M, T, P, Y, keep[:-1].astype(np.int32))).T
    assert relative_error(partial, ref, floor).max() > 1e-6
    assert surrogate.choose_major(M, Y, 6).tolist() != [] and set(surrogate.choose_major(M, Y, 6).tolist()) == set(keep.tolist())

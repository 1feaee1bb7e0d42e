"""Integer sampling work must be bit-exact: the product's mask/ordering code
against the oracle's restatement of buildHarmonicSampling and the test
helper's independent restatement of the random-checkered mask."""
import numpy as np

from disorder_amd import sampling
from oracle import disorder_oracle as O
from tests.helpers import checkered_sampling


def test_random_checkered_matches_restatement():
    for (n1, n2, t1, t2) in ((16, 12, 4, 4), (256, 192, 8, 8), (20, 15, 5, 3)):
        S = t1 * t2
        rng_a = np.random.default_rng(42)
        rng_b = np.random.default_rng(42)
        perms = sampling.permutation_table(rng_a, (n1 // t1) * (n2 // t2), S)
        H = sampling.random_checkered(n1, n2, t1, t2, perms)
        Href = checkered_sampling(rng_b, n1, n2, S, t1, t2)
        assert np.array_equal(H, Href)


def test_harmonic_sampling_matches_oracle_bit_exact():
    rng = np.random.default_rng(3)
    for S in (1, 4, 16):
        H = rng.integers(0, S + 1, size=(24, 18))
        k1, k2, off = sampling.harmonic_sampling(H, S, with_outer=True)
        rk1, rk2, roff, _ = O.build_harmonic_sampling(H, S)
        assert np.array_equal(k1, rk1) and np.array_equal(k2, rk2) and np.array_equal(off, roff)
        k1b, k2b, offb = sampling.harmonic_sampling(H, S)
        assert np.array_equal(offb, roff[:S + 1]) and np.array_equal(k1b, rk1[:roff[S]])


def test_config_sampling_counts():
    rng = np.random.default_rng(0)
    perms = sampling.permutation_table(rng, 32 * 24, 64)
    H = sampling.random_checkered(256, 192, 8, 8, perms)
    k1, k2, off = sampling.harmonic_sampling(H, 64)
    assert off[-1] == 256 * 192
    assert np.all(np.diff(off) == 768)
    # every phase-encode position exactly once
    assert len(set(zip(k1.tolist(), k2.tolist()))) == 256 * 192

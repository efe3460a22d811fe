"""The counter-based generator (COUNTER mode) as the oracle restates it:
Philox4x32-10 from Salmon et al., "Parallel random numbers: as easy as 1, 2,
3" (SC'11), and the sampler's streams built on it.

* Known answers: the three Philox4x32-10 vectors the authors publish with
  their Random123 library (counter, key -> output block).
* Statistics on 10^7 draws of the streams the sampler consumes: the first
  uniform of consecutive draw sites (one per SNP and sweep) and the run of
  uniforms within one site, each tested for uniformity (Kolmogorov-Smirnov
  and a 1,000-bin chi-square) and for lag-1 correlation; streams of
  different chains and iterations are uncorrelated.

The device's uniforms are compared with these bit for bit in
tests/test_gpu_philox.py.
"""
import numpy as np
import pytest
import scipy.stats

import orc

# Random123 kat_vectors, philox4x32 with 10 rounds: ctr[4], key[2] -> out[4]
KAT = [
    ((0x00000000, 0x00000000, 0x00000000, 0x00000000), (0x00000000, 0x00000000),
     (0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8)),
    ((0xffffffff, 0xffffffff, 0xffffffff, 0xffffffff), (0xffffffff, 0xffffffff),
     (0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd)),
    ((0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344), (0xa4093822, 0x299f31d0),
     (0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1)),
]


@pytest.mark.parametrize("ctr,key,want", KAT)
def test_philox_known_answers(ctr, key, want):
    assert orc.philox4x32_10(ctr, key) == want


def test_stream_words_map_to_unit_interval():
    """A block's words (w0, w1) and (w2, w3) become two uniforms as 52-bit
    fractions: ((w0 << 32 | w1) >> 12) * 2^-52."""
    key = orc.chain_key(123, 4)
    blk = orc.philox4x32_10((7, 9, 2, 0), (key & 0xFFFFFFFF, key >> 32))
    u = orc.counter_uniforms(123, 4, 7, 9, 2, 4)
    want = [((blk[0] << 32 | blk[1]) >> 12) / 2.0 ** 52, ((blk[2] << 32 | blk[3]) >> 12) / 2.0 ** 52]
    assert u[0] == want[0] and u[1] == want[1]
    blk1 = orc.philox4x32_10((7, 9, 2, 1), (key & 0xFFFFFFFF, key >> 32))
    assert u[2] == ((blk1[0] << 32 | blk1[1]) >> 12) / 2.0 ** 52


def check_uniform(u: np.ndarray, label: str) -> None:
    n = len(u)
    assert u.min() >= 0.0 and u.max() < 1.0
    ks = scipy.stats.kstest(u, "uniform")
    counts = np.bincount((u * 1000).astype(np.int64), minlength=1000)
    chi2 = scipy.stats.chisquare(counts)
    r = np.corrcoef(u[:-1], u[1:])[0, 1]
    print(f"{label}: n={n} KS p={ks.pvalue:.3g} chi2 p={chi2.pvalue:.3g} lag-1 r={r:.2e}")
    assert ks.pvalue > 1e-4 and chi2.pvalue > 1e-4
    assert abs(r) < 5.0 / np.sqrt(n)


def test_first_draws_of_consecutive_sites_are_uniform_and_uncorrelated():
    """One sweep's per-SNP draws: the first uniform of sites 0 .. 10^7 - 1."""
    check_uniform(orc.counter_first(20240607, 0, 1, 0, 2, 10_000_000), "site-first")


def test_draws_within_one_site_are_uniform_and_uncorrelated():
    """The run of uniforms one site consumes (rejection loops draw many)."""
    check_uniform(orc.counter_uniforms(-42, 3, 17, 5, 3, 10_000_000), "within-site")


@pytest.mark.parametrize("other", [dict(chain=1), dict(iteration=2), dict(seed=20240608)])
def test_streams_of_other_chains_and_iterations_are_independent(other):
    n = 1_000_000
    base = dict(seed=20240607, chain=0, iteration=1)
    a = orc.counter_first(base["seed"], base["chain"], base["iteration"], 0, 2, n)
    kw = dict(base, **other)
    b = orc.counter_first(kw["seed"], kw["chain"], kw["iteration"], 0, 2, n)
    r = np.corrcoef(a, b)[0, 1]
    assert abs(r) < 5.0 / np.sqrt(n), (other, r)
    assert np.mean(a == b) < 1e-5

"""The device's counter-based generator against the oracle's Philox4x32-10
(Salmon et al., SC'11; pinned by the Random123 known answers in
tests/test_philox.py): every uniform bit for bit, over a grid of seeds,
chains, iterations, draw sites and tags, including the extreme counter
values the sampler uses (setup and per-iteration scalar sites near 2^32)."""
import ctypes as C
import itertools

import numpy as np
import pytest

import orc
from sodabayes_amd import _lib

pytestmark = pytest.mark.gpu

SEEDS = [1, -42, 20240607, 2 ** 62 + 12345]
CHAINS = [0, 1, 15, 63]
ITERATIONS = [0, 1, 999, 0x7FFFFFFF, 0xFFFFFFFE]
SITES = [0, 1, 4095, 499_999, 0xFFFFFFF0, 0xFFFFFFF2]
TAGS = [1, 2, 3, 4, 5, 6]
PER = 9   # odd: the second word pair of the last block is left unused


def device_uniforms(seed, chain, ctr, per):
    ctr = np.ascontiguousarray(np.asarray(ctr, dtype=np.uint32).reshape(-1, 3))
    out = np.zeros((len(ctr), per))
    _lib.check(_lib.load().sb_counter_uniforms(seed, chain, ctr.ctypes.data_as(C.POINTER(C.c_uint32)),
                                               len(ctr), per, out.ctypes.data_as(C.POINTER(C.c_double))))
    return out


@pytest.mark.parametrize("seed", SEEDS)
def test_device_uniforms_equal_the_oracle(seed):
    sites = list(itertools.product(ITERATIONS, SITES, TAGS))
    for chain in CHAINS:
        got = device_uniforms(seed, chain, sites, PER)
        for k, (it, site, tag) in enumerate(sites):
            want = orc.counter_uniforms(seed, chain, it, site, tag, PER)
            assert np.array_equal(got[k], want), (seed, chain, it, site, tag, got[k], want)
    print(f"seed {seed}: {len(sites) * len(CHAINS) * PER} uniforms bit-equal")


def test_device_sweep_draws_equal_the_oracle():
    """The first uniform of 500,000 consecutive SNP sites of one sweep."""
    n = 500_000
    ctr = np.stack([np.full(n, 7), np.arange(n), np.full(n, 2)], axis=1)
    got = device_uniforms(20240607, 5, ctr, 1)[:, 0]
    assert np.array_equal(got, orc.counter_first(20240607, 5, 7, 0, 2, n))

"""sb_predict on the GPU against the oracle and the reference's .gv values.

The GPU evaluates pred_i = (mu - sum_j 2 f_j c_j) + sum_j x_ij c_j with
c_j = g_j / s_j; the oracle sums z_ij g_j in SNP order.  The two differ only
by rounding, bounded here by 1e-12 of sum_j |x_ij c_j| + sum_j 2 f_j |c_j| +
|mu| per individual (FP64 sums over at most a few 10^4 terms).  Individuals
are independent, so at the large size the oracle checks slices of the
genotype bytes (first, middle and the ragged last one).
"""
import os

import numpy as np
import pytest

import orc
from common import GOLDEN
from sodabayes_amd import plink
from sodabayes_amd import predict as sp

pytestmark = pytest.mark.gpu

EPS = np.finfo(np.float64).eps


def scale(bed, n, f, eff, mu):
    c = np.where((f < EPS) | (1.0 - f < EPS), 0.0, eff / np.sqrt(np.maximum(2.0 * f * (1.0 - f), 1e-300)))
    x = plink.decode_bed(bed, n).astype(np.float64)
    return np.abs(x) @ np.abs(c) + np.sum(2.0 * f * np.abs(c)) + abs(mu)


def random_case(rng, n, p, mono=True):
    counts = rng.integers(0, 3, size=(n, p)).astype(np.int8)
    miss = rng.random((n, p)) < 0.02
    if mono and p >= 3:
        counts[:, 0] = 0
        counts[:, p - 1] = 2
        miss[:, [0, p - 1]] = False
    bed = plink.encode_bed(counts, miss)
    f = plink.allele_freq(bed, n)
    eff = rng.normal(0, 0.05, p)
    return bed, f, eff


@pytest.mark.parametrize("run", ["example1", "example3", "example4"])
def test_predict_matches_reference_examples(run):
    z = np.load(os.path.join(GOLDEN, "predict_kat.npz"))
    n = int(z["nind"])
    bed, f, eff, mu = z["bed"], z[f"{run}_freq"], z[f"{run}_eff"], float(z[f"{run}_mu"])
    got = sp.predict(bed, n, f, eff, mu)
    ref = orc.predict(bed, n, f, eff, mu)
    assert np.all(np.abs(got - ref) <= 1e-12 * scale(bed, n, f, eff, mu))
    want = z[f"{run}_pred"]
    ex = np.floor(np.log10(np.abs(want))) + 1
    assert np.all(np.abs(got - want) <= 0.5 * 10.0 ** (ex - 7) * 1.0001 + 1e-12)


@pytest.mark.parametrize("n,p", [(1, 1), (3, 7), (16, 8), (17, 9), (200, 5000), (1001, 333), (4099, 2500)])
def test_predict_matches_oracle(n, p):
    rng = np.random.default_rng(1000 * n + p)
    bed, f, eff = random_case(rng, n, p)
    got = sp.predict(bed, n, f, eff, -0.75)
    ref = orc.predict(bed, n, f, eff, -0.75)
    assert np.all(np.abs(got - ref) <= 1e-12 * scale(bed, n, f, eff, -0.75))


def test_predict_reads_device_genotypes_in_place():
    import torch
    rng = np.random.default_rng(5)
    n, p = 777, 1234
    bed, f, eff = random_case(rng, n, p)
    host = sp.predict(bed, n, f, eff, 0.5)
    dev = sp.predict(torch.from_numpy(bed).cuda(), n, f, eff, 0.5)
    assert np.array_equal(host, dev)


def test_predict_large_slices_match_oracle():
    rng = np.random.default_rng(11)
    n, p = 50003, 20000
    nb = plink.bytes_per_snp(n)
    bed = rng.integers(0, 256, size=(p, nb), dtype=np.uint8)
    f = rng.uniform(0.02, 0.98, p)
    f[:50] = 0.0
    eff = rng.normal(0, 0.01, p)
    got = sp.predict(bed, n, f, eff, 3.0)
    for b0, b1 in ((0, 128), (6000, 6128), (nb - 5, nb)):
        i0, i1 = 4 * b0, min(n, 4 * b1)
        sub = np.ascontiguousarray(bed[:, b0:b1])
        ref = orc.predict(sub, i1 - i0, f, eff, 3.0)
        tol = 1e-12 * scale(sub, i1 - i0, f, eff, 3.0)
        assert np.all(np.abs(got[i0:i1] - ref) <= tol), (b0, b1)

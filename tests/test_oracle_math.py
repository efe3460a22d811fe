"""The oracle against what the mathematics fixes, independent of the
reference: brute-force integer Gram, eigenvalues, Cholesky identities, LAPACK
rank, distribution moments."""
import numpy as np
import pytest
import scipy.linalg
import scipy.stats

import orc
from sodabayes_amd import plink, synth


def small_data(n=133, p=40, seed=3, missing=0.03):
    x, _ = synth.genotypes(n, p, seed)
    x[:, 5] = 0   # monomorphic
    x[:, 9] = 2   # monomorphic the other way
    miss = np.random.default_rng(seed).random(x.shape) < missing
    miss[:, [5, 9]] = False
    return plink.encode_bed(x, miss), x, miss


def test_plink_roundtrip_and_missing_as_het():
    bed, x, miss = small_data()
    dec = plink.decode_bed(bed, x.shape[0])
    want = np.where(miss, 1, x)
    assert np.array_equal(dec, want)


def test_gram_count_is_integer_xtx():
    bed, x, miss = small_data()
    n = x.shape[0]
    counts = plink.decode_bed(bed, n).astype(np.int64)
    gb = orc.pack_genbits(bed, n)
    XtX = counts.T @ counts
    for j in range(counts.shape[1]):
        for k in range(j + 1):
            assert orc.gram_count(gb, j, k) == XtX[j, k]


def test_xxproduct_is_standardised_dot():
    bed, x, _ = small_data()
    n = x.shape[0]
    f = orc.freq(bed, n)
    Z = orc.standardize(bed, n, f)
    G = orc.gram(bed, n, list(range(bed.shape[0])), f)
    np.testing.assert_allclose(G, Z.T @ Z, rtol=1e-11, atol=1e-9)
    assert np.all(G[5] == 0) and np.all(G[:, 9] == 0)


def test_freq_and_standardisation():
    bed, x, miss = small_data()
    n = x.shape[0]
    f = orc.freq(bed, n)
    c = np.where(miss, 1, x).astype(float)
    np.testing.assert_array_equal(f, c.sum(0) / (2 * n))
    Z = orc.standardize(bed, n, f)
    ok = (f > 0) & (f < 1)
    np.testing.assert_allclose(Z[:, ok], (c[:, ok] - 2 * f[ok]) / np.sqrt(2 * f[ok] * (1 - f[ok])), rtol=1e-15)
    assert np.all(Z[:, ~ok] == 0)


def test_power_method_finds_lambda_max():
    rng = np.random.default_rng(1)
    B = rng.standard_normal((60, 30))
    A = B.T @ B
    lam, it = orc.power_method(A, orc.Rng(9))
    ev = np.linalg.eigvalsh(A)[-1]
    assert abs(lam - ev) / ev < 1e-9
    assert it < 1000 * 30


def test_cholesky_pd_identities_and_failure():
    rng = np.random.default_rng(2)
    B = rng.standard_normal((40, 40))
    M = B @ B.T + 40 * np.eye(40)
    st, U = orc.cholesky_pd(M)
    assert st == 0
    assert np.allclose(np.tril(U, -1), 0)
    np.testing.assert_allclose(U.T @ U, M, rtol=1e-13, atol=1e-11)
    np.testing.assert_allclose(U, scipy.linalg.cholesky(M, lower=False), rtol=1e-11, atol=1e-12)
    M2 = M.copy()
    M2[7, 7] = -1.0
    st, _ = orc.cholesky_pd(M2)
    assert st == -1                      # non-positive diagonal: status -1
    v = rng.standard_normal(40)
    M3 = np.outer(v, v) + np.diag(np.r_[np.ones(20), np.zeros(20)]) * 1e-3
    st, _ = orc.cholesky_pd(M3 + 1e-300)
    assert st > 0                        # failing pivot reported 1-based


def test_cholesky_psd_rank_and_reconstruction():
    rng = np.random.default_rng(4)
    B = rng.standard_normal((30, 11))
    A = B @ B.T
    st, F, piv, rank = orc.cholesky_psd(A)
    assert st == 0 and rank == np.linalg.matrix_rank(A) == 11
    R = np.triu(F)[:rank, :]
    P = piv - 1
    np.testing.assert_allclose(R.T @ R, A[np.ix_(P, P)], rtol=1e-10, atol=1e-9)


def test_permutation_is_a_permutation():
    p = orc.permutate(1000, 64, orc.Rng(5))
    assert sorted(p) == list(range(1, 1001))


@pytest.mark.parametrize("seed", [1, 99])
def test_distribution_moments(seed):
    r = orc.Rng(seed)
    N = 20000
    x = np.array([r.normal(1.5, 2.0) for _ in range(N)])
    assert abs(x.mean() - 1.5) < 5 * 2 / np.sqrt(N)
    assert scipy.stats.kstest((x - 1.5) / 2.0, "norm").pvalue > 1e-4
    for shape in (0.6, 3.0):
        g = np.array([r.gamma(shape, 2.0) for _ in range(N)])
        assert scipy.stats.kstest(g, "gamma", args=(shape, 0, 2.0)).pvalue > 1e-4
    nu, s = 6.0, 0.5
    v = np.array([r.sinvchi2(nu, s) for _ in range(N)])
    # scaled inverse chi-square(nu, s) = InvGamma(nu/2, nu s / 2)
    assert scipy.stats.kstest(v, "invgamma", args=(nu / 2, 0, nu * s / 2)).pvalue > 1e-4
    a = np.array([4.0, 2.0, 1.0])
    d = np.array([r.dirichlet(a) for _ in range(5000)])
    np.testing.assert_allclose(d.mean(0), a / a.sum(), atol=0.01)
    np.testing.assert_allclose(d.sum(1), 1.0, rtol=1e-14)

"""Pivoted-Cholesky fallback and skipped-SNP updates on the GPU (FP64 path).

The fixture case psd_bayesa lowers block 1's d so that dI - G is indefinite
there; the library's test hook sb_test_shift_d applies the same shift.
The reference fixture pins BayesA (setup and chain: test_gpu_parity); here
the other methods are checked against the oracle in REPLAY mode (classes
equal, continuous state within 1e-9 relative), COUNTER mode through the
residual identity r = y - Z g - mu (1e-11) and the block products against
numpy FP64 (1e-12); the mixed batched path sweeps the pivoted-Cholesky
block with the FP64 block kernels (identity, products, posterior).
"""
import numpy as np
import pytest

import orc
from common import load_case, oracle_options, rel_err, sampler_options
from sodabayes_amd import _lib, soda

pytestmark = pytest.mark.gpu

R4 = "0.0,0.0001,0.001,0.01"
METHODS = {
    "BayesA": dict(method="BayesA"),
    "BayesB": dict(method="BayesB", pi=0.9),
    "BayesC": dict(method="BayesC", pi=0.9),
    "BayesR": dict(method="BayesR", nmix=4, mix=R4),
}


def case_opts(method, **extra):
    c = load_case("psd_bayesa")
    o = dict(c["opts"])
    o.update(METHODS[method])
    o.update(extra)
    return c, o


@pytest.fixture
def dshift():
    """The fixture's block and shift of d (sb_test_shift_d)."""
    o = load_case("psd_bayesa")["opts"]
    return int(o["dshift_block"]), float(o["dshift"])


def make(c, o, dshift, **kw):
    return soda.SodaSampler(sampler_options(o, **kw), int(c["nind"]), int(c["nloci"]), c["why"],
                            bed=c["bed"], test_dshift=dshift)


@pytest.mark.parametrize("method", ["BayesB", "BayesC", "BayesR"])
def test_replay_fallback_matches_oracle(dshift, method):
    c, o = case_opts(method)
    n, p = int(c["nind"]), int(c["nloci"])
    ref = orc.run(oracle_options(o, n, p), c["bed"], c["why"], None)
    s = make(c, o, dshift)
    s.setup()
    assert s.nskipped == int(ref["nskipped"][0]) and s.nskipped > 0
    assert s.naind == int(ref["naind"][0])
    assert np.array_equal(s.skipped(), ref["i_skipped"][:s.nskipped].astype(np.int64))
    s.init()
    hyp = s.run()
    res = s.results()
    assert np.array_equal(res["snptracker"], ref["snptracker"].astype(float))
    for k in ("g", "yadj", "yAadj", "gstore", "varindist"):
        assert rel_err(res[k], ref[k]) < 1e-9, (k, rel_err(res[k], ref[k]))
    assert rel_err(hyp[:, 2], ref["hyp_mu"]) < 1e-11
    assert rel_err(hyp[:, 4], ref["hyp_vare"]) < 1e-9
    s.close()


@pytest.mark.parametrize("method", ["BayesA", "BayesB", "BayesR"])
def test_counter_fallback_keeps_residuals(dshift, method):
    c, o = case_opts(method, numit=30)
    n, p = int(c["nind"]), int(c["nloci"])
    s = make(c, o, dshift, rng="counter")
    s.setup()
    assert s.nskipped > 0
    # every skipped SNP sits in the shifted block
    bounds = soda.block_bounds(p, s.blocksize, s.nblocks, s.blockt)
    a, e = bounds[int(o["dshift_block"])]
    assert set(s.skipped() - 1) <= set(int(v) for v in s.permvec[a:e])
    s.init()
    s.run(30, trace=False)
    f = orc.freq(c["bed"], n)
    Z = orc.standardize(c["bed"], n, f)
    g = s.get("G", p)
    mu = s.get("SCALARS", 4)[0]
    want = c["why"] - Z @ g - mu
    r = s.get("YADJ", n)
    assert np.linalg.norm(r - want) / np.linalg.norm(want) < 1e-11
    # block products through the dense augmenting block
    ra = s.get("YAADJ", s.blocksize)
    XA = s.get("XA", s.blocksize * p).reshape(s.blocksize, p, order="F")
    for b, (a, e) in enumerate(bounds):
        nb = e - a
        want_b = Z[:, s.permvec[a:e]].T @ r + XA[:nb, a:e].T @ ra[:nb]
        got_b = s.block_rhs(b)[0]
        assert np.linalg.norm(got_b - want_b) / np.linalg.norm(want_b) < 1e-12, b
    s.close()


@pytest.mark.parametrize("method", ["BayesA", "BayesR"])
def test_mixed_path_sweeps_fallback_blocks(dshift, method):
    """The mixed batched path takes the same fallback as the FP64 path (same
    skipped SNPs and naind), sweeps the pivoted-Cholesky block with the FP64
    block kernels chain by chain and the skipped SNPs one by one, and keeps
    r = y - Z g - mu for every chain (to the fixed-point remainders of its
    batched blocks); the products of every block, the dense one included,
    match numpy FP64 of the same state."""
    C = 4
    c, o = case_opts(method, numit=30)
    n, p = int(c["nind"]), int(c["nloci"])
    ref = make(c, o, dshift, rng="counter")
    ref.setup()
    s = make(c, o, dshift, rng="counter", precision="mixed", chains=C)
    s.setup()
    assert s.nskipped == ref.nskipped > 0 and s.naind == ref.naind
    assert np.array_equal(s.skipped(), ref.skipped())
    ref.close()
    s.init()
    s.run(30, trace=False)
    f = orc.freq(c["bed"], n)
    Z = orc.standardize(c["bed"], n, f)
    XA = s.get("XA", s.blocksize * p).reshape(s.blocksize, p, order="F")
    bounds = soda.block_bounds(p, s.blocksize, s.nblocks, s.blockt)
    worst_r, worst_p = 0.0, 0.0
    for ch in range(C):
        g, mu, r = s.get("G", p, ch), s.get("SCALARS", 4, ch)[0], s.get("YADJ", n, ch)
        want = c["why"] - Z @ g - mu
        worst_r = max(worst_r, np.linalg.norm(r - want) / np.linalg.norm(want))
        ra = s.get("YAADJ", s.blocksize, ch)
        for b, (a, e) in enumerate(bounds):
            nb = e - a
            want_b = Z[:, s.permvec[a:e]].T @ r + XA[:nb, a:e].T @ ra[:nb]
            got_b = s.block_rhs(b)[ch]
            worst_p = max(worst_p, np.linalg.norm(got_b - want_b) / np.linalg.norm(want_b))
    print(f"mixed x{C} {method} with a pivoted-Cholesky block: residual identity {worst_r:.2e}, "
          f"products {worst_p:.2e}")
    # r carries the fixed-point remainders of the batched blocks' updates (at
    # most half a quantum per SNP): the bar of the full-size mixed runs
    assert worst_r < 5e-6 and worst_p < 1e-6
    s.close()


def test_mixed_fallback_posterior_matches_fp64(dshift):
    """Posterior means of mu, sigma2_e, sigma2_g and model size from 16
    mixed chains with the fallback block against 16 FP64 counter chains of
    the same job: within 4 combined standard errors of the chain means."""
    c, o = case_opts("BayesR", numit=1500, burnin=300)
    n, p = int(c["nind"]), int(c["nloci"])
    keys = ("mu_store", "vare_store", "vara_store", "included_store")

    def means(rs):
        return np.array([[float(np.ravel(r[k])[0]) for k in keys] for r in rs])

    s = make(c, o, dshift, rng="counter", precision="mixed", chains=16, seed=101)
    s.setup()
    s.init()
    s.run(trace=False)
    mixed = means([s.results(k) for k in range(16)])
    s.close()
    fp = []
    for k in range(16):
        t = make(c, o, dshift, rng="counter", seed=201 + k)
        t.setup()
        t.init()
        t.run(trace=False)
        fp.append(t.results())
        t.close()
    fp64 = means(fp)
    z = (mixed.mean(0) - fp64.mean(0)) / np.sqrt(mixed.var(0, ddof=1) / 16 + fp64.var(0, ddof=1) / 16)
    for k, zk, a, b in zip(keys, z, mixed.mean(0), fp64.mean(0)):
        print(f"fallback {k}: mixed x16 {a:.6g} vs FP64 x16 {b:.6g}  z={zk:.2f}")
    assert np.all(np.abs(z) < 4.0)

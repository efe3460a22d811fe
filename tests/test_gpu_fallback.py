"""Pivoted-Cholesky fallback and skipped-SNP updates on the GPU (FP64 path).

The fixture case psd_bayesa lowers block 1's d so that dI - G is indefinite
there; the library's test hook sb_test_shift_d applies the same shift.
The reference fixture pins BayesA (setup and chain: test_gpu_parity); here
the other methods are checked against the oracle in REPLAY mode (classes
equal, continuous state within 1e-9 relative), COUNTER mode through the
residual identity r = y - Z g - mu (1e-11) and the block products against
numpy FP64 (1e-12), and the mixed path must refuse the fallback.
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


def test_mixed_path_refuses_fallback(dshift):
    c, o = case_opts("BayesR")
    s = make(c, o, dshift, rng="counter", precision="mixed", chains=2)
    with pytest.raises(_lib.SbError) as ei:
        s.setup()
    assert ei.value.code == _lib.SB_ERR_NOT_PD
    assert "FP64 path only" in str(ei.value)
    s.close()

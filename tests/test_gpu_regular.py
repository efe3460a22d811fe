"""The single-site sampler (-SODAOff, models.f90:607-962) through the Python host.

The reference's own files for -SODAOff jobs pin the replay chain draw for
draw (tests/test_fortran_host.py).  Here, on the counter-based generator:
* a chain's trajectory depends only on the seed and its chain index, so
  chain c of a 3-chain handle equals chain c of a handle with more chains;
* the single-site sampler and the SODA sampler target the same posterior:
  posterior means of mu, vare, vara and model size over 16 chains of each
  agree within 4 combined Monte Carlo standard errors, and the posterior
  effect means correlate;
* r = y - Z g - mu holds at the end of a run;
* a checkpoint resumes the chain bit for bit;
* with held-out rows (cross-validation folds) the held-out phenotypes never
  enter and the identity holds with their current draws.
"""
import numpy as np
import pytest

import orc
from sodabayes_amd import soda, synth

R4 = "0.0,0.0001,0.001,0.01"


def sampler(bed, n, p, y, chains, its, burnin=0, seed=7, **kw):
    o = soda.SodaOptions(method="BayesR", nmix=4, mix=R4, numit=its, burnin=burnin, seed=seed, rng="counter",
                         chains=chains, **kw)
    s = soda.SodaSampler(o, n, p, y, bed=bed)
    s.setup()
    s.init()
    s.run(trace=False)
    return s


def test_sodaoff_options_select_the_single_site_sampler():
    bed, y, _ = synth.dataset(100, 50, seed=80)
    o = soda.SodaOptions(method="BayesR", nmix=4, mix=R4, sodaoff=True, blocksize=60, rblocks=True, numit=2, seed=3)
    s = soda.SodaSampler(o, 100, 50, y, bed=bed)
    assert s.regular and (s.blocksize, s.nblocks, s.blockt) == (0, 0, 0)
    assert np.array_equal(s.permvec, np.arange(50))          # SNPs in id order (models.f90:672-675)
    o = soda.SodaOptions(method="BayesR", nmix=4, mix=R4, blocksize=1, numit=2, seed=3)
    assert soda.SodaSampler(o, 100, 50, y, bed=bed).regular
    o = soda.SodaOptions(method="BayesR", nmix=4, mix=R4, numit=2, seed=3)       # -nthreads 1 partition
    assert soda.SodaSampler(o, 100, 50, y, bed=bed).regular
    o = soda.SodaOptions(method="BayesR", nmix=4, mix=R4, blocksize=2, numit=2, seed=3)
    assert not soda.SodaSampler(o, 100, 50, y, bed=bed).regular


@pytest.mark.gpu
def test_single_site_chains_depend_only_on_their_index():
    n, p = 300, 120
    bed, y, _ = synth.dataset(n, p, seed=81, nqtl=10, h2=0.5)
    a = sampler(bed, n, p, y, 3, 12, sodaoff=True)
    b = sampler(bed, n, p, y, 5, 12, sodaoff=True)
    for c in range(3):
        for q, k in (("G", p), ("CLASS", p), ("SCALARS", 4), ("YADJ", n)):
            assert np.array_equal(a.get(q, k, c), b.get(q, k, c)), (c, q)
    assert not np.array_equal(a.get("G", p, 0), a.get("G", p, 1))
    Z = orc.standardize(bed, n, orc.freq(bed, n))
    for c in range(3):
        g, mu = a.get("G", p, c), a.get("SCALARS", 4, c)[0]
        want = y - Z @ g - mu
        assert np.linalg.norm(a.get("YADJ", n, c) - want) < 1e-9 * np.linalg.norm(want)
    a.close()
    b.close()


@pytest.mark.gpu
def test_single_site_and_soda_samplers_share_the_posterior():
    n, p, C, its, burn = 600, 200, 16, 1500, 300
    bed, y, _ = synth.dataset(n, p, seed=82, nqtl=10, h2=0.5)
    reg = sampler(bed, n, p, y, C, its, burnin=burn, seed=31, sodaoff=True)
    sod = sampler(bed, n, p, y, C, its, burnin=burn, seed=37, blocksize=100, precision="mixed")
    rr = [reg.results(c) for c in range(C)]
    rs = [sod.results(c) for c in range(C)]
    reg.close()
    sod.close()
    for key in ("mu_store", "vara_store", "vare_store", "included_store"):
        a = np.array([float(np.ravel(r[key])[0]) for r in rr])
        b = np.array([float(np.ravel(r[key])[0]) for r in rs])
        z = abs(a.mean() - b.mean()) / np.sqrt(a.var(ddof=1) / C + b.var(ddof=1) / C)
        print(f"{key}: single-site {a.mean():.6g} vs SODA {b.mean():.6g}  z={z:.2f}")
        assert z < 4.0, key
    ga = np.mean([r["gstore"] for r in rr], axis=0)
    gb = np.mean([r["gstore"] for r in rs], axis=0)
    corr = np.corrcoef(ga, gb)[0, 1]
    print(f"correlation of posterior effect means {corr:.4f}")
    assert corr > 0.95


@pytest.mark.gpu
def test_single_site_checkpoint_resumes_bit_for_bit():
    n, p = 400, 150
    bed, y, _ = synth.dataset(n, p, seed=83, nqtl=10, h2=0.5)
    o = soda.SodaOptions(method="BayesR", nmix=4, mix=R4, numit=30, burnin=2, seed=41, sodaoff=True)

    def make():
        s = soda.SodaSampler(o, n, p, y, bed=bed)
        s.setup()
        return s

    a = make()
    a.init()
    a.run(9, trace=False)
    blob = a.save_state()
    hyp_a = a.run(7)
    want = {q: a.get(q, k, 0) for q, k in (("G", p), ("CLASS", p), ("YADJ", n), ("SCALARS", 4), ("GSTORE", p))}
    a.close()
    b = make()
    b.load_state(blob)
    hyp_b = b.run(7)
    assert np.array_equal(hyp_a, hyp_b)
    for q, k in (("G", p), ("CLASS", p), ("YADJ", n), ("SCALARS", 4), ("GSTORE", p)):
        assert np.array_equal(want[q], b.get(q, k, 0)), q
    b.close()


@pytest.mark.gpu
def test_single_site_folds_keep_the_identity_and_never_read_heldout_phenotypes():
    n, p, C = 500, 120, 4
    bed, y, _ = synth.dataset(n, p, seed=84, nqtl=10, h2=0.5)
    held = np.zeros((C, n), bool)
    rs = np.random.default_rng(85)
    for c in range(C):
        held[c, rs.choice(n, size=n // 5, replace=False)] = True
    ys = np.tile(y, (C, 1))
    junk = ys.copy()
    junk[held] = 1.0e6

    def run(why):
        o = soda.SodaOptions(method="BayesCpi", numit=12, burnin=0, seed=43, rng="counter", chains=C, sodaoff=True)
        s = soda.SodaSampler(o, n, p, why, bed=bed, heldout=held)
        s.setup()
        s.init()
        s.run(trace=False)
        return s

    a, b = run(ys), run(junk)
    Z = orc.standardize(bed, n, orc.freq(bed, n))
    for c in range(C):
        ga, gb = a.get("G", p, c), b.get("G", p, c)
        assert np.array_equal(ga, gb), c
        r, mu = a.get("YADJ", n, c), a.get("SCALARS", 4, c)[0]
        want = a.get("WHY", n, c) - Z @ ga - mu
        assert np.linalg.norm(r - want) < 1e-9 * np.linalg.norm(want), c
    a.close()
    b.close()

"""Batched traits and cross-validation folds (sb_mcmc_init_chains) on the GPU.

Routing: chains whose phenotype columns are all y reproduce sb_mcmc_init's
chains bit for bit, and a chain's trajectory depends only on its own
phenotype (with the seed and its chain index), not on the other chains'.

Held-out rows (the validation individuals of a fold) are missing
phenotypes, redrawn every iteration from y_i ~ N(mu + z_i'g, vare):
* their phenotype values never enter (bit-equal runs when they change);
* at the end of an iteration their residuals are fresh N(0, vare) draws;
* r = y - Z g - mu holds with y at the current draws;
* the chains sample the posterior given the training rows: posterior means
  of mu, vare, vara and model size agree with a long FP64 chain run on the
  training rows alone (same allele-frequency standardisation) within 4
  combined Monte Carlo standard errors, and the effect means correlate.
"""
import numpy as np
import pytest

import orc
from sodabayes_amd import soda, synth

R4 = "0.0,0.0001,0.001,0.01"
OPTS = dict(method="BayesR", nmix=4, mix=R4, blocksize=100)


def data(n=801, p=300, seed=51):
    bed, y, _ = synth.dataset(n, p, seed=seed, nqtl=15, h2=0.5)
    _, y2, _ = synth.dataset(n, p, seed=seed + 1, nqtl=15, h2=0.3)   # a second trait (other QTL)
    return bed, y, y2


def run(bed, n, p, why, chains, its, heldout=None, seed=7, freq=None, precision="mixed", burnin=0):
    o = soda.SodaOptions(numit=its, burnin=burnin, seed=seed, rng="counter", precision=precision,
                         chains=chains, **OPTS)
    s = soda.SodaSampler(o, n, p, why, bed=bed, heldout=heldout, freq=freq)
    s.setup()
    s.init()
    hyp = s.run(trace=True)
    return s, hyp


def state(s, c):
    p, n = s.nloci, s.nind
    return {"G": s.get("G", p, c), "CLASS": s.get("CLASS", p, c), "YADJ": s.get("YADJ", n, c),
            "SCALARS": s.get("SCALARS", 4, c)}


def same(a, b):
    return all(np.array_equal(a[k], b[k]) for k in a)


def test_heldout_and_phenotype_shapes_are_checked():
    bed, y, _ = data(200, 60)
    o = soda.SodaOptions(numit=2, burnin=0, seed=3, rng="counter", precision="mixed", chains=3, **OPTS)
    s = soda.SodaSampler(o, 200, 60, np.stack([y, y]), bed=bed)
    with pytest.raises(ValueError):
        s.init()
    s = soda.SodaSampler(o, 200, 60, y, bed=bed, heldout=np.zeros((2, 200), bool))
    with pytest.raises(ValueError):
        s.init()


@pytest.mark.gpu
def test_identical_columns_reproduce_the_shared_phenotype_chains():
    n, p, C = 801, 300, 4
    bed, y, _ = data(n, p)
    s1, _ = run(bed, n, p, y, C, 12)
    s2, _ = run(bed, n, p, np.tile(y, (C, 1)), C, 12)
    for c in range(C):
        assert same(state(s1, c), state(s2, c)), c
    s1.close()
    s2.close()


@pytest.mark.gpu
def test_a_chain_depends_only_on_its_own_phenotype():
    n, p = 801, 300
    bed, ya, yb = data(n, p)
    sab, _ = run(bed, n, p, np.stack([ya, yb]), 2, 12)
    sbb, _ = run(bed, n, p, np.stack([yb, yb]), 2, 12)
    saa, _ = run(bed, n, p, np.stack([ya, ya]), 2, 12)
    assert same(state(sab, 1), state(sbb, 1))
    assert same(state(sab, 0), state(saa, 0))
    assert not np.array_equal(state(sab, 0)["G"], state(sab, 1)["G"])
    for s in (sab, sbb, saa):
        s.close()


def folds(n, C, seed, frac=(0.05, 0.4)):
    """Per chain, a random held-out fraction between 5% and 40% of the rows."""
    rs = np.random.default_rng(seed)
    h = np.zeros((C, n), bool)
    for c in range(C):
        k = int(rs.uniform(*frac) * n)
        h[c, rs.choice(n, size=k, replace=False)] = True
    return h


@pytest.mark.gpu
def test_heldout_phenotypes_never_enter():
    n, p, C = 801, 300, 6
    bed, y, _ = data(n, p)
    h = folds(n, C, 3)
    ys = np.tile(y, (C, 1))
    junk = ys.copy()
    junk[h] = 1.0e6 * np.random.default_rng(4).standard_normal(int(h.sum()))
    s1, _ = run(bed, n, p, ys, C, 10, heldout=h)
    s2, _ = run(bed, n, p, junk, C, 10, heldout=h)
    for c in range(C):
        assert same(state(s1, c), state(s2, c)), c
    s1.close()
    s2.close()


@pytest.mark.gpu
def test_heldout_residuals_are_fresh_draws_and_the_identity_holds():
    n, p, C = 1201, 300, 8
    bed, y, _ = data(n, p)
    h = folds(n, C, 5, frac=(0.3, 0.4))
    s, _ = run(bed, n, p, np.tile(y, (C, 1)), C, 15, heldout=h)
    Z = orc.standardize(bed, n, orc.freq(bed, n))
    for c in range(C):
        st = state(s, c)
        r, vare, mu = st["YADJ"], st["SCALARS"][1], st["SCALARS"][0]
        e = r[h[c]]
        m = e.size
        z_mean = e.mean() / np.sqrt(vare / m)
        ratio = e.var() / vare
        print(f"chain {c}: {m} held-out residuals, mean z {z_mean:.2f}, variance / vare {ratio:.3f}")
        assert abs(z_mean) < 4.5
        assert abs(ratio - 1.0) < 5.0 * np.sqrt(2.0 / m)
        yc = s.get("WHY", n, c)
        assert np.array_equal(yc[~h[c]], y[~h[c]])        # training phenotypes untouched
        want = yc - Z @ st["G"] - mu
        err = np.linalg.norm(r - want) / np.linalg.norm(want)
        assert err < 1e-5, (c, err)
    s.close()


def batch_means_se(x, nb=20):
    x = np.asarray(x, float)
    m = len(x) // nb
    b = x[: m * nb].reshape(nb, m).mean(1)
    return b.std(ddof=1) / np.sqrt(nb)


@pytest.mark.gpu
def test_folds_sample_the_posterior_given_the_training_rows():
    n, p, C = 800, 200, 16
    bed, y, x = synth.dataset(n, p, seed=61, nqtl=10, h2=0.5)
    freq = orc.freq(bed, n)
    held = np.zeros(n, bool)
    held[np.random.default_rng(62).choice(n, size=n // 4, replace=False)] = True
    # long FP64 chain on the training rows alone, standardised with the same frequencies
    from sodabayes_amd import plink
    keep = ~held
    bed_t = plink.encode_bed(x[keep])
    ref, ref_hyp = run(bed_t, int(keep.sum()), p, y[keep], 1, 8000, seed=11, freq=freq, precision="fp64",
                       burnin=1000)
    ref_res = ref.results(0)
    ref_hyp = ref_hyp[1000:]
    ref.close()
    # the imputed quarter of the rows slows the mixing of vare (data
    # augmentation), hence the longer burn-in than the shared-phenotype test's
    s, _ = run(bed, n, p, np.tile(y, (C, 1)), C, 3000, heldout=np.tile(held, (C, 1)), seed=29, freq=freq,
               burnin=1000)
    res = [s.results(c) for c in range(C)]
    s.close()
    cols = {"mu": (2, "mu_store"), "vara": (3, "vara_store"), "vare": (4, "vare_store"),
            "included": (1, "included_store")}
    for name, (col, key) in cols.items():
        m_ref, se_ref = ref_hyp[:, col].mean(), batch_means_se(ref_hyp[:, col])
        cm = np.array([float(np.ravel(r[key])[0]) for r in res])
        m_gpu, se_gpu = cm.mean(), cm.std(ddof=1) / np.sqrt(C)
        z = abs(m_gpu - m_ref) / np.sqrt(se_gpu ** 2 + se_ref ** 2)
        print(f"folds x{C} {name}: {m_gpu:.6g} +- {se_gpu:.2g} vs training-only {m_ref:.6g} +- {se_ref:.2g}"
              f"  z={z:.2f}")
        assert z < 4.0, name
    g = np.mean([r["gstore"] for r in res], axis=0)
    corr = np.corrcoef(g, ref_res["gstore"])[0, 1]
    print(f"folds x{C}: correlation of posterior effect means {corr:.4f}")
    assert corr > 0.95

"""Mixed-precision batched path and the counter-based RNG on the GPU.

Bars:
* block products X_s'r + A_s'r_a of every chain against an FP64 evaluation of
  the same state: norm-wise relative error <= 1e-6 (mixed, int8 limbs) and
  <= 1e-12 (FP64 path);
* posterior means of mu, sigma2_e, sigma2_g and of the effects from batched
  mixed chains against long oracle chains (the reference's algorithm) within
  4 combined Monte Carlo standard errors.
"""
import numpy as np
import pytest

import orc
from sodabayes_amd import soda, synth

pytestmark = pytest.mark.gpu

R4 = "0.0,0.0001,0.001,0.01"


def dataset(n=1001, p=300, seed=41):
    bed, why, _ = synth.dataset(n, p, seed=seed, nqtl=15, h2=0.5)
    return bed, why


def fp64_products(s, bed, chain):
    """X_s'r + U_s'r_a in numpy FP64 from the sampler's downloaded state."""
    n, p = s.nind, s.nloci
    f = orc.freq(bed, n)
    Z = orc.standardize(bed, n, f)
    r = s.get("YADJ", n, chain)
    ra = s.get("YAADJ", s.blocksize, chain)
    XA = s.get("XA", s.blocksize * p).reshape(s.blocksize, p, order="F")
    out = []
    for a, e in soda.block_bounds(p, s.blocksize, s.nblocks, s.blockt):
        cols = s.permvec[a:e]
        nb = e - a
        out.append(Z[:, cols].T @ r + XA[:nb, a:e].T @ ra[:nb])
    return out


METHOD_OPTS = {
    "BayesR": dict(method="BayesR", nmix=4, mix=R4),
    "BayesA": dict(method="BayesA"),
    "BayesB": dict(method="BayesB", pi=0.9),
    "BayesC": dict(method="BayesC", pi=0.9),
    "BayesCpi": dict(method="BayesCpi"),
    "BayesRc": dict(method="BayesRc", nmix=4, mix=R4),
}


def sampler(method, precision, chains, seed, blocksize=100, n=1001, p=300):
    bed, why = dataset(n, p)
    grp = (np.arange(p) % 3) + 1 if method == "BayesRc" else None
    o = soda.SodaOptions(blocksize=blocksize, numit=50, burnin=0, seed=seed, rng="counter",
                         precision=precision, chains=chains, **METHOD_OPTS[method])
    return soda.SodaSampler(o, n, p, why, bed=bed, grp=grp), bed, why


# chain counts cover every (limb tile, chain tile) kernel configuration:
# 1-16 chains (3 limb tiles, 1 chain tile), 20 (6, 2), 40 (9, 3), and 49-64
# as two chain groups of (6, 2): 64 (both full) and 50 (the second group
# holding 18 chains)
@pytest.mark.parametrize("method,precision,chains,bar", [
    ("BayesR", "mixed", 4, 1e-6), ("BayesR", "mixed", 1, 1e-6), ("BayesR", "fp64", 1, 1e-12),
    ("BayesA", "mixed", 4, 1e-6), ("BayesB", "mixed", 3, 1e-6), ("BayesC", "mixed", 5, 1e-6),
    ("BayesCpi", "mixed", 6, 1e-6), ("BayesRc", "mixed", 4, 1e-6),
    ("BayesR", "mixed", 20, 1e-6), ("BayesR", "mixed", 40, 1e-6), ("BayesR", "mixed", 64, 1e-6),
    ("BayesR", "mixed", 50, 1e-6)])
def test_block_products_against_fp64(method, precision, chains, bar):
    s, bed, why = sampler(method, precision, chains, 5)
    s.setup()
    s.init()
    s.run(3, trace=False)
    worst = 0.0
    for b in range(s.nblocks):
        got = s.block_rhs(b)
        for c in range(chains):
            ref = fp64_products(s, bed, c)[b]
            err = np.linalg.norm(got[c] - ref) / np.linalg.norm(ref)
            worst = max(worst, err)
    print(f"{method} {precision} x{chains}: worst norm-wise relative error of X'r + A'r_a = {worst:.3e}")
    assert worst <= bar
    s.close()


@pytest.mark.parametrize("precision,chains,bar", [("mixed", 16, 1e-6), ("fp64", 1, 1e-12)])
def test_block_products_over_several_chunks(precision, chains, bar):
    """n = 9,000 individuals: the products span five 2,048-individual
    genotype chunks and several X_s'Q chunks of 16 super-steps, whose exact
    partials are summed across workgroups; blocks of 512 SNPs."""
    n, p = 9000, 1024
    s, bed, why = sampler("BayesR", precision, chains, 6, blocksize=512, n=n, p=p)
    s.setup()
    s.init()
    s.run(2, trace=False)
    worst = 0.0
    for b in range(s.nblocks):
        got = s.block_rhs(b)
        for c in range(chains):
            ref = fp64_products(s, bed, c)[b]
            worst = max(worst, np.linalg.norm(got[c] - ref) / np.linalg.norm(ref))
    print(f"n={n} {precision} x{chains}: worst norm-wise relative error of X'r + A'r_a = {worst:.3e}")
    assert worst <= bar
    s.close()


def batch_means_se(x, nb=20):
    x = np.asarray(x)
    m = len(x) // nb
    bm = x[: m * nb].reshape(nb, m).mean(1)
    return bm.std(ddof=1) / np.sqrt(nb)


POST_DATA = dict(n=1000, p=300, seed=77)
POST_OPTS = dict(method="BayesR", nmix=4, mix=R4, blocksize=100, df_va=4.0, scale_va=2.0,
                 df_ve=4.0, scale_ve=0.5)


def posterior_trace(rng, precision, chains, numit, burnin, seed, opts=POST_OPTS):
    d = POST_DATA
    bed, why, _ = synth.dataset(d["n"], d["p"], seed=d["seed"], nqtl=15, h2=0.5)
    o = soda.SodaOptions(numit=numit, burnin=burnin, seed=seed, rng=rng, precision=precision,
                         chains=chains, **opts)
    s = soda.SodaSampler(o, d["n"], d["p"], why, bed=bed)
    s.setup()
    s.init()
    hyp = s.run()
    res = [s.results(c) for c in range(chains)]
    s.close()
    return hyp, res


@pytest.fixture(scope="module")
def replay_reference():
    """Long chain of the reference's own algorithm (REPLAY mode, which
    matches the reference bit for bit on the fixtures)."""
    hyp, res = posterior_trace("replay", "fp64", 1, 4000, 500, 11)
    return hyp[500:], res[0]


@pytest.mark.parametrize("precision,chains,numit,burnin", [("fp64", 1, 4000, 500),
                                                           ("mixed", 16, 700, 200)])
def test_counter_and_mixed_posteriors_match_replay(replay_reference, precision, chains, numit, burnin):
    """Posterior means of mu, sigma2_e, sigma2_g and model size from the
    production samplers (Philox draws; FP64 or mixed-precision batched
    products) agree with the reference algorithm within 4 combined Monte
    Carlo standard errors, and their effect means correlate with it."""
    hyp, res = posterior_trace("counter", precision, chains, numit, burnin, 29)
    compare_posteriors(replay_reference, hyp, res, chains, burnin, f"{precision} x{chains}")


def compare_posteriors(reference, hyp, res, chains, burnin, label):
    ref_hyp, ref_res = reference
    cols = {"mu": 2, "vara": 3, "vare": 4, "included": 1}
    for name, col in cols.items():
        m_ref, se_ref = ref_hyp[:, col].mean(), batch_means_se(ref_hyp[:, col])
        if chains == 1:
            tr = hyp[burnin:, col]
            m_gpu, se_gpu = tr.mean(), batch_means_se(tr)
        else:
            key = {"mu": "mu_store", "vara": "vara_store", "vare": "vare_store",
                   "included": "included_store"}[name]
            cm = np.array([float(np.ravel(r[key])[0]) for r in res])
            m_gpu, se_gpu = cm.mean(), cm.std(ddof=1) / np.sqrt(chains)
        z = abs(m_gpu - m_ref) / np.sqrt(se_gpu ** 2 + se_ref ** 2)
        print(f"{label} {name}: {m_gpu:.6g} +- {se_gpu:.2g} vs reference "
              f"{m_ref:.6g} +- {se_ref:.2g}  z={z:.2f}")
        assert z < 4.0, name
    g = np.mean([r["gstore"] for r in res], axis=0)
    corr = np.corrcoef(g, ref_res["gstore"])[0, 1]
    print(f"{label}: correlation of posterior effect means {corr:.4f}")
    assert corr > 0.95


# Six classes: the draw code compiled for up to eight.  REPLAY of that code
# is pinned by the reference's 5- and 9-class jobs (tests/test_fortran_host.py);
# here the two Philox samplers, the batched SODA path (class probabilities
# from per-workgroup constants, draw_mix_fast) and the single-site sampler
# (draw_mix_wave), 16 chains each, must agree on the posterior.  Their
# standard errors come from the spread of the chain means: a single long
# chain mixes too slowly over six classes to serve as the reference.
WIDE_OPTS = dict(POST_OPTS, nmix=6, mix="0.0,0.00003,0.0001,0.0003,0.001,0.01")


def test_wide_mixture_samplers_agree():
    C = 16
    _, soda_res = posterior_trace("counter", "mixed", C, 1500, 500, 31, WIDE_OPTS)
    _, site_res = posterior_trace("counter", "fp64", C, 1500, 500, 37, dict(WIDE_OPTS, sodaoff=True))
    for key in ("mu_store", "vara_store", "vare_store", "included_store"):
        a = np.array([float(np.ravel(r[key])[0]) for r in soda_res])
        b = np.array([float(np.ravel(r[key])[0]) for r in site_res])
        z = abs(a.mean() - b.mean()) / np.sqrt(a.var(ddof=1) / C + b.var(ddof=1) / C)
        print(f"6 classes {key}: SODA mixed x{C} {a.mean():.6g} vs single-site x{C} {b.mean():.6g}  z={z:.2f}")
        assert z < 4.0, key
    ga = np.mean([r["gstore"] for r in soda_res], axis=0)
    gb = np.mean([r["gstore"] for r in site_res], axis=0)
    corr = np.corrcoef(ga, gb)[0, 1]
    print(f"6 classes: correlation of posterior effect means {corr:.4f}")
    assert corr > 0.95


@pytest.mark.parametrize("case", ["bayesr", "nblocks", "ldblock", "ragged"])
def test_counter_mode_bound_matches_reference_d(case):
    """COUNTER mode takes d from Lanczos instead of the power iteration: on
    blocks of <= 96 SNPs the Ritz value is lambda_max, so d matches the
    reference's power-method d."""
    from common import load_case, rel_err, sampler_options
    c = load_case(case)
    o = sampler_options(c["opts"], rng="counter")
    s = soda.SodaSampler(o, int(c["nind"]), int(c["nloci"]), c["why"], bed=c["bed"])
    xxd = s.setup()
    assert rel_err(xxd, c["out"]["xxd"]) < 1e-10
    s.close()


@pytest.mark.parametrize("method,precision,chains,bar", [
    ("BayesR", "fp64", 1, 1e-11), ("BayesR", "mixed", 8, 1e-5), ("BayesA", "mixed", 4, 1e-5),
    ("BayesB", "mixed", 4, 1e-5), ("BayesRc", "mixed", 4, 1e-5), ("BayesR", "mixed", 64, 1e-5),
    ("BayesR", "mixed", 40, 1e-5), ("BayesR", "mixed", 50, 1e-5)])
def test_residuals_stay_consistent(method, precision, chains, bar):
    """After the sweeps' incremental updates r_c = y - X beta_c - mu_c still
    holds for every chain (before the every-1000-iterations resync)."""
    n, p = 1001, 300
    s, bed, why = sampler(method, precision, chains, 8)
    s.setup()
    s.init()
    s.run(20, trace=False)
    f = orc.freq(bed, n)
    Z = orc.standardize(bed, n, f)
    worst = 0.0
    for c in range(chains):
        g = s.get("G", p, c)
        mu = s.get("SCALARS", 4, c)[0]
        r = s.get("YADJ", n, c)
        want = why - Z @ g - mu
        worst = max(worst, np.linalg.norm(r - want) / np.linalg.norm(want))
    print(f"{method} {precision} x{chains}: residual identity error {worst:.3e}")
    assert worst < bar
    s.close()


@pytest.mark.parametrize("precision,slots,absent", [
    ("mixed", ("rquant", "xq", "drawb", "dquant", "xupd", "rared"), ("rhs", "draw", "list", "update")),
    ("fp64", ("rhs", "draw", "list", "update"), ("rquant", "xq", "drawb", "dquant", "xupd", "rared"))])
def test_kernel_timing_slots(monkeypatch, precision, slots, absent):
    """SODABAYES_KERNEL_TIMING=1 times every block kernel of the path in use
    (one launch per block and iteration; the r limbs of the batched path
    once per iteration, the later blocks' come with the r_a reduction) and
    nothing of the other path."""
    s, bed, why = sampler("BayesR", precision, 4 if precision == "mixed" else 1, 3)
    s.setup()
    s.init()
    s.run(1, trace=False)
    assert all(v[1] == 0 for v in s.kernel_times().values())   # off unless asked for
    monkeypatch.setenv("SODABAYES_KERNEL_TIMING", "1")
    s.run(2, trace=False)
    kt = s.kernel_times()
    for k in slots:
        want = 2 if k == "rquant" else 2 * s.nblocks
        assert kt[k][1] == want and kt[k][0] > 0 and kt[k][2] > 0, k
    for k in absent:
        assert kt[k][1] == 0, k
    s.close()


def _chain_states(s, chains):
    p, n = s.nloci, s.nind
    return [(s.get("G", p, c), s.get("CLASS", p, c), s.get("YADJ", n, c)) for c in range(chains)]


@pytest.mark.parametrize("chains", [4, 40, 64])
def test_tile_shapes_give_identical_chains(monkeypatch, chains):
    """The X_s'Q chunk length and the row tiles per wave only regroup exact
    int32 limb sums, so every shape - fixed ones (four and two chunks of the
    8 super-steps at n = 1,001) and the one the tuner picks before the first
    sweep - gives bit-identical chains; the tuner reports a shape it tried."""
    states, infos = [], []
    for kss, rw in ((2, 4), (5, 2), (None, None)):
        for k, v in (("SODABAYES_KSS", kss), ("SODABAYES_ROWS_PER_WAVE", rw)):
            if v is None:
                monkeypatch.delenv(k, raising=False)
            else:
                monkeypatch.setenv(k, str(v))
        s, _, _ = sampler("BayesR", "mixed", chains, 13)
        s.setup()
        s.init()
        s.run(8, trace=False)
        states.append(_chain_states(s, chains))
        infos.append(s.info())
        s.close()
    for other in states[1:]:
        for a, b in zip(states[0], other):
            assert all(np.array_equal(x, y) for x, y in zip(a, b))
    tuned = infos[-1]
    print(f"x{chains}: tuned shape kss={tuned['chunk_supersteps']} rows/wave={tuned['rows_per_wave']} "
          f"({tuned['tuned_block_us']:.1f} us for block 0)")
    assert tuned["tuned_block_us"] > 0
    assert 1 <= tuned["chunk_supersteps"] <= 8
    assert tuned["rows_per_wave"] in ((2, 4) if chains <= 32 else (2,))
    assert infos[0]["chunk_supersteps"] == 2 and infos[0]["rows_per_wave"] == (4 if chains <= 32 else 2)
    assert infos[1]["chunk_supersteps"] == 4 and infos[1]["rows_per_wave"] == 2
    assert infos[0]["tuned_block_us"] == 0

"""The workloads at their full sizes, on the GPU, through the C ABI.

Genotypes are generated in device memory (sodabayes_amd.synth.device_*),
so nothing of their size passes through the host.  The checks are the ones
that do not depend on size, each against an evaluation independent of the
path under test:

* products: X_s'r + U_s'r_a of the sampler's current state, for every
  chain, against an FP64 evaluation of that same state in torch from the
  .bed bytes (standardised as the reference's xscale does) and the block's
  factor (SB_Q_XA_BLOCK): norm-wise error <= 1e-6 for the mixed-precision
  path (models.f90:358), <= 1e-12 for the FP64 path;
* Gram: the device's standardised X_s'X_s bit for bit against the oracle's
  popcount restatement of xxproduct (soda.f90:520-556) on sampled blocks;
* residual identity: the incrementally updated r against y - Z beta - mu
  (on all individuals at 20,000 x 50,000; on 4,096 sampled individuals at
  the larger sizes, where Z has up to 6e10 entries).  The mixed path's r
  differs from it by the remainders its fixed-point update carries to each
  SNP's next update (at most half a quantum per SNP), a bounded difference
  that does not grow with the iterations: below 5e-6 relative even when all
  500,000 SNPs move every iteration (BayesA);
* held-out rows are never read: folds whose held-out phenotypes are
  replaced by 1e6-scale noise give bit-identical chains;
* traits are independent: changing every other trait's phenotype leaves a
  trait's chain bit-identical.

Workloads (the reference runs one job at a time; batched chains are this
build's):
  c2  BayesR, n = 20,000, p = 50,000, 50-SNP LD blocks (r^2 about 0.3),
      1,000 QTL, h2 = 0.3, subsets of 2,048: FP64 and mixed (one chain);
      identity after 1,000 iterations with the resync switched off
  c3  BayesA and BayesB, n = 50,000, p = 500,000, t(4) QTL, h2 = 0.4,
      -pi 0.99 -df_va 4 -df_ve 4, subsets of 4,096, 16 chains (mixed)
  c4  BayesCpi, n = 40,000, p = 100,000, 64 folds holding out 5-40% of the
      individuals, subsets of 2,048, 64 chains (mixed)
  c5  BayesRc, n = 100,000, p = 600,000, 4 annotation classes, 16 traits,
      subsets of 4,096 (mixed)
"""
from __future__ import annotations

import numpy as np
import pytest

import orc
from sodabayes_amd import check, soda, synth

pytestmark = pytest.mark.gpu

R4 = "0.0,0.0001,0.001,0.01"
PROD_BAR = {"mixed": 1e-6, "fp64": 1e-12}


@pytest.fixture(scope="module")
def torch():
    import torch as t
    if not t.cuda.is_available():
        pytest.fail("no GPU visible to torch")
    return t


def check_products(torch, s, bed, freq, blocks, chains, precision, label):
    worst = check.products_error(torch, s, bed, freq, blocks, chains)
    print(f"{label}: worst norm-wise error of X'r + U'r_a over {len(blocks)} blocks x {len(chains)} chains "
          f"= {worst:.3e}")
    assert worst <= PROD_BAR[precision], label
    return worst


block_cols = check.block_cols
residual_identity = check.residual_identity
sample_blocks = check.sample_blocks


# -------------------------------------------------------------------- c2 --

@pytest.fixture(scope="module")
def c2_data(torch):
    bed, freq, y, _ = synth.device_dataset(torch, 20000, 50000, 2002, fmin=0.05, fmax=0.5, nqtl=1000, h2=0.3,
                                           effect_df=0.0, ld_block=50)
    return bed, freq, y


def c2_sampler(c2_data, precision, numit, resync=0):
    bed, freq, y = c2_data
    o = soda.SodaOptions(method="BayesR", nmix=4, mix=R4, blocksize=2048, numit=numit, burnin=0, seed=29,
                         rng="counter", precision=precision, chains=1, resync=resync)
    s = soda.SodaSampler(o, 20000, 50000, y, freq=freq, bed_device_ptr=bed.data_ptr())
    s.setup()
    s.init()
    return s


def test_c2_gram_bit_exact_on_sampled_blocks(torch, c2_data):
    bed, freq, _ = c2_data
    s = c2_sampler(c2_data, "mixed", 1)
    for b in (0, s.nblocks // 2, s.nblocks - 1):
        cols = block_cols(s, b)
        nb = len(cols)
        got = s.get("GRAM", nb * nb, b).reshape(nb, nb)
        rows = bed[torch.as_tensor(np.asarray(cols, dtype=np.int64), device=bed.device)].cpu().numpy()
        want = orc.gram_block(rows, s.nind, freq[cols])
        assert np.array_equal(got, want), f"block {b}"
        print(f"c2 block {b}: {nb} x {nb} Gram bit-exact")
    s.close()


@pytest.mark.parametrize("precision", ["mixed", "fp64"])
def test_c2_products_every_block_and_identity_after_1000_iterations(torch, c2_data, precision):
    """1,000 iterations with the every-1,000-iterations resync switched off,
    then the products of every block and the residual identity."""
    bed, freq, _ = c2_data
    s = c2_sampler(c2_data, precision, 1000, resync=10 ** 9)
    s.run(1000, trace=False)
    check_products(torch, s, bed, freq, range(s.nblocks), [0], precision, f"c2 {precision} x1")
    err = residual_identity(torch, s, bed, freq, [0])
    print(f"c2 {precision} x1: residual identity after 1,000 iterations without resync: {err:.3e}")
    assert err < (1e-7 if precision == "mixed" else 1e-10)
    s.close()


# -------------------------------------------------------------------- c3 --

@pytest.fixture(scope="module")
def c3_data(torch):
    bed, freq, y, _ = synth.device_dataset(torch, 50000, 500000, 3003, nqtl=5000, h2=0.4, effect_df=4.0)
    return bed, freq, y


@pytest.mark.parametrize("method", ["BayesA", "BayesB"])
def test_c3_products_and_identity(torch, c3_data, method):
    bed, freq, y = c3_data
    n, p, C = 50000, 500000, 16
    vary = float(np.var(y))
    share = p * (1.0 if method == "BayesA" else 0.01)
    kw = dict(df_va=4.0, df_ve=4.0, scale_va=0.4 * vary / share * 1.5, scale_ve=0.6 * vary * 1.5)
    if method == "BayesB":
        kw["pi"] = 0.99
    o = soda.SodaOptions(method=method, blocksize=4096, numit=200, burnin=0, seed=31, rng="counter",
                         precision="mixed", chains=C, **kw)
    s = soda.SodaSampler(o, n, p, y, freq=freq, bed_device_ptr=bed.data_ptr())
    s.setup()
    s.init()
    s.run(20, trace=False)
    check_products(torch, s, bed, freq, sample_blocks(s.nblocks), list(range(C)), "mixed", f"c3 {method} x{C}")
    rows = np.random.default_rng(3).choice(n, size=4096, replace=False)
    err20 = residual_identity(torch, s, bed, freq, list(range(C)), rows=rows)
    s.run(180, trace=False)
    err200 = residual_identity(torch, s, bed, freq, list(range(C)), rows=rows)
    print(f"c3 {method} x{C}: residual identity (4,096 sampled individuals, every chain) after 20 iterations "
          f"{err20:.3e}, after 200 {err200:.3e}")
    # r differs from y - Z beta - mu by the remainders the update's fixed
    # point carries (at most half a quantum per SNP, each quantum 2^-23 of
    # its block's largest |delta / s|): a sum over all p SNPs when every SNP
    # moves (BayesA), bounded and not growing with the iterations
    assert err200 < 5e-6 and err200 < 3.0 * max(err20, 1e-9)
    s.close()


# -------------------------------------------------------------------- c4 --

@pytest.fixture(scope="module")
def c4_data(torch):
    n, p, C = 40000, 100000, 64
    bed, freq, y, _ = synth.device_dataset(torch, n, p, 4004, nqtl=1000, h2=0.3, effect_df=0.0)
    rs = np.random.default_rng(4)
    held = np.zeros((C, n), bool)
    for c in range(C):
        held[c, rs.choice(n, size=int(rs.uniform(0.05, 0.4) * n), replace=False)] = True
    return bed, freq, y, held


def c4_run(c4_data, why, iters=10):
    bed, freq, _, held = c4_data
    o = soda.SodaOptions(method="BayesCpi", blocksize=2048, numit=iters, burnin=0, seed=41, rng="counter",
                         precision="mixed", chains=64)
    s = soda.SodaSampler(o, 40000, 100000, why, freq=freq, bed_device_ptr=bed.data_ptr(), heldout=held)
    s.setup()
    s.init()
    s.run(iters, trace=False)
    return s


def test_c4_folds_never_read_held_out_rows_and_keep_identities(torch, c4_data):
    bed, freq, y, held = c4_data
    C = 64
    s = c4_run(c4_data, np.tile(y, (C, 1)))
    state = [(s.get("G", 100000, c), s.get("CLASS", 100000, c), s.get("YADJ", 40000, c)) for c in range(C)]
    chains = [0, 17, 42, 63]
    check_products(torch, s, bed, freq, sample_blocks(s.nblocks, 4), chains, "mixed", "c4 folds x64")
    rows = np.random.default_rng(5).choice(40000, size=4096, replace=False)
    err = residual_identity(torch, s, bed, freq, chains, rows=rows)
    print(f"c4 folds x64: residual identity with the drawn held-out phenotypes: {err:.3e}")
    assert err < 5e-6
    s.close()
    junk = np.tile(y, (C, 1))
    junk[held] = np.random.default_rng(6).standard_normal(int(held.sum())) * 1e6
    s = c4_run(c4_data, junk)
    for c in range(C):
        g, cl, r = s.get("G", 100000, c), s.get("CLASS", 100000, c), s.get("YADJ", 40000, c)
        assert np.array_equal(g, state[c][0]) and np.array_equal(cl, state[c][1]), c
        assert np.array_equal(r, state[c][2]), c
    print("c4 folds x64: held-out phenotypes replaced by 1e6-scale noise give bit-identical chains")
    s.close()


# -------------------------------------------------------------------- c5 --

@pytest.fixture(scope="module")
def c5_data(torch):
    n, p, T = 100000, 600000, 16
    bed, freq, Y = synth.device_traits(torch, n, p, 5005, ntraits=T)
    grp = np.random.default_rng(5).choice(4, size=p, p=[0.6, 0.2, 0.15, 0.05]) + 1
    return bed, freq, Y, grp


def c5_run(c5_data, Y, iters):
    bed, freq, _, grp = c5_data
    o = soda.SodaOptions(method="BayesRc", nmix=4, mix=R4, blocksize=4096, numit=iters, burnin=0, seed=51,
                         rng="counter", precision="mixed", chains=Y.shape[0])
    s = soda.SodaSampler(o, 100000, 600000, Y, freq=freq, bed_device_ptr=bed.data_ptr(), grp=grp)
    s.setup()
    s.init()
    s.run(iters, trace=False)
    return s


def test_c5_traits_products_identity_and_independence(torch, c5_data):
    bed, freq, Y, _ = c5_data
    n, p = 100000, 600000
    s = c5_run(c5_data, Y, 8)
    keep = (s.get("G", p, 3), s.get("CLASS", p, 3), s.get("YADJ", n, 3))
    chains = [0, 3, 9, 15]
    check_products(torch, s, bed, freq, sample_blocks(s.nblocks), chains, "mixed", "c5 traits x16")
    rows = np.random.default_rng(7).choice(n, size=4096, replace=False)
    err = residual_identity(torch, s, bed, freq, chains, rows=rows)
    print(f"c5 traits x16: residual identity (4,096 sampled individuals): {err:.3e}")
    assert err < 5e-6
    s.close()
    # trait 3's chain does not depend on the other traits' phenotypes
    Y2 = Y[::-1].copy()
    Y2[3] = Y[3]
    s = c5_run(c5_data, Y2, 8)
    assert np.array_equal(s.get("G", p, 3), keep[0]) and np.array_equal(s.get("CLASS", p, 3), keep[1])
    assert np.array_equal(s.get("YADJ", n, 3), keep[2])
    print("c5 traits x16: trait 3's chain bit-identical with every other trait's phenotype changed")
    s.close()

"""Device-state checkpoints (sb_state_size / sb_state_save / sb_state_load).

A chain saved after k iterations and continued in a fresh handle (set up on
the same inputs) gives, bit for bit, the state and the trace records of the
uninterrupted run, on the FP64 replay path (generator state in the
checkpoint) and on the mixed batched path with per-chain phenotypes and
held-out rows.  A checkpoint of another job shape, or of other genotypes,
is refused.
"""
import numpy as np
import pytest

from sodabayes_amd import soda, synth

pytestmark = pytest.mark.gpu

R4 = "0.0,0.0001,0.001,0.01"
QUERIES = ["G", "CLASS", "YADJ", "YAADJ", "SCALARS", "GSTORE", "STORES", "P", "WHY", "INDISTSTORE"]


def make(opts, n, p, bed, why, **kw):
    s = soda.SodaSampler(opts, n, p, why, bed=bed, **kw)
    s.setup()
    return s


def snapshot(s, chains):
    sizes = {"G": s.nloci, "CLASS": s.nloci, "YADJ": s.nind, "YAADJ": s.blocksize, "SCALARS": 4,
             "GSTORE": s.nloci, "STORES": 5, "P": s.ngroup * s.nmix, "WHY": s.nind,
             "INDISTSTORE": s.nloci * s.nmix}
    return [{q: s.get(q, sizes[q], c) for q in QUERIES} for c in range(chains)]


def check_resume(opts, n, p, bed, why, chains, k1, k2, **kw):
    a = make(opts, n, p, bed, why, **kw)
    a.init()
    a.run(k1, trace=False)
    blob = a.save_state()
    hyp_a = a.run(k2)
    want = snapshot(a, chains)
    a.close()
    b = make(opts, n, p, bed, why, **kw)
    b.load_state(blob)
    assert b.rep == k1
    hyp_b = b.run(k2)
    got = snapshot(b, chains)
    b.close()
    assert np.array_equal(hyp_a, hyp_b)
    for c in range(chains):
        for q in QUERIES:
            assert np.array_equal(want[c][q], got[c][q]), (c, q)
        # every sample after burn-in counts one class of each SNP (the counts
        # were read out once mid-run, by the save, and again at the end)
        counts = want[c]["INDISTSTORE"].reshape(-1, p)
        assert np.all(counts >= 0) and np.all(counts == np.round(counts)), c
        assert np.all(counts.sum(axis=0) == k1 + k2 - opts.burnin), c
    return blob


def test_replay_chain_resumes_bit_for_bit():
    n, p = 600, 240
    bed, why, _ = synth.dataset(n, p, seed=71, nqtl=12, h2=0.5)
    o = soda.SodaOptions(method="BayesR", nmix=4, mix=R4, blocksize=60, numit=40, burnin=5, seed=19)
    check_resume(o, n, p, bed, why, 1, 15, 12)


def test_batched_traits_and_folds_resume_bit_for_bit():
    n, p, C = 700, 300, 4
    bed, y, _ = synth.dataset(n, p, seed=72, nqtl=12, h2=0.5)
    _, y2, _ = synth.dataset(n, p, seed=73, nqtl=12, h2=0.3)
    why = np.stack([y, y2, y, y2])
    held = np.zeros((C, n), bool)
    held[1, ::5] = True
    held[3, 1::3] = True
    grp = (np.arange(p) % 4) + 1
    o = soda.SodaOptions(method="BayesRc", nmix=4, mix=R4, blocksize=100, numit=40, burnin=5, seed=23,
                         rng="counter", precision="mixed", chains=C)
    check_resume(o, n, p, bed, why, C, 11, 9, heldout=held, grp=grp)


def test_checkpoint_of_another_job_is_refused():
    n, p = 500, 200
    bed, why, _ = synth.dataset(n, p, seed=74, nqtl=10, h2=0.5)
    o = soda.SodaOptions(method="BayesCpi", blocksize=50, numit=10, burnin=0, seed=5, rng="counter")
    a = make(o, n, p, bed, why)
    a.init()
    a.run(3, trace=False)
    blob = a.save_state()
    a.close()
    other = soda.SodaOptions(method="BayesCpi", blocksize=40, numit=10, burnin=0, seed=5, rng="counter")
    b = make(other, n, p, bed, why)
    with pytest.raises(RuntimeError, match="another job shape"):
        b.load_state(blob)
    b.close()
    bed2, _, _ = synth.dataset(n, p, seed=75, nqtl=10, h2=0.5)
    c = make(o, n, p, bed2, why)
    with pytest.raises(RuntimeError, match="augmenting matrices differ"):
        c.load_state(blob)
    c.close()
    d = make(o, n, p, bed, why)
    with pytest.raises(RuntimeError, match="not a checkpoint"):
        d.load_state(b"x" * 64)
    d.close()

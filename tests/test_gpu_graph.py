"""The sweep captured as a HIP graph (SODABAYES_GRAPH=1) against eager
launches (the default): bit-identical chains and .hyp traces, for the replayed
FP64 path, the counter-based FP64 path and the mixed batched path.  The
graph is captured once per handle and replayed for every iteration of
several sb_mcmc_run calls, with the host synchronising between calls, a
resync (every 7 iterations here) run eagerly between replays, and a
checkpoint download in between; every replay's result is checked through
the final state and the per-iteration trace."""
import numpy as np
import pytest

from sodabayes_amd import soda, synth

pytestmark = pytest.mark.gpu

R4 = "0.0,0.0001,0.001,0.01"


def run(monkeypatch, graph, rng, precision, chains):
    monkeypatch.setenv("SODABAYES_GRAPH", "1" if graph else "0")
    n, p = 1500, 600
    bed, why, _ = synth.dataset(n, p, seed=61, nqtl=20, h2=0.5)
    o = soda.SodaOptions(method="BayesR", nmix=4, mix=R4, blocksize=150, numit=40, burnin=10, thin=3, seed=67,
                         rng=rng, precision=precision, chains=chains, resync=7)
    s = soda.SodaSampler(o, n, p, why, bed=bed)
    s.setup()
    s.init()
    hyps = [s.run(13), s.run(1), s.run(26)]
    blob = s.save_state()
    out = [s.results(c) for c in range(chains)]
    s.close()
    return np.concatenate(hyps), out, blob


@pytest.mark.parametrize("rng,precision,chains", [("replay", "fp64", 1), ("counter", "fp64", 1),
                                                  ("counter", "mixed", 6), ("counter", "mixed", 64)])
def test_graph_replays_equal_eager_launches(monkeypatch, rng, precision, chains):
    he, re_, be = run(monkeypatch, False, rng, precision, chains)
    hg, rg, bg = run(monkeypatch, True, rng, precision, chains)
    assert np.array_equal(he, hg)
    assert be == bg
    for a, b in zip(re_, rg):
        for k in ("g", "yadj", "yAadj", "gstore", "snptracker", "mu_store", "vare_store", "indiststore"):
            assert np.array_equal(np.asarray(a[k]), np.asarray(b[k])), k
    print(f"{rng} {precision} x{chains}: 40 iterations over 3 calls, graph == eager bit for bit")

"""The plain-SODA workload at full size: BayesCpi, n = 2,000 individuals,
p = 10,000 SNPs, SODA subsets of 1,000 SNPs (f ~ U(0.05, 0.5), 200 QTL,
h2 = 0.5), against reference runs of that exact job
(tests/golden/c1_bayescpi.npz, made by tests/golden/make_c1_golden.py with
the reference program built from its sources).

* Trajectory: the reference program with its SODA branch on the GPU
  (REPLAY draws, FP64) runs the reference's 100-iteration job and writes the
  reference's files: every .hyp line (model size and class counts exactly;
  mu, Vara, Ve to the printed digit), the .eff effects and PIPs, .mod, .frq.
* Posterior: the production samplers (counter-based Philox draws; FP64 one
  chain per handle, and the mixed-precision batched path with 16 chains)
  against 16 reference chains of 4,000 iterations (2,000 burn-in).  For mu,
  Ve, Vara, model size and pi the difference of the means lies within 3
  combined Monte Carlo standard errors (standard errors from the spread of
  independent chain means).  For the 10,000 effect means and the 2,000 GEBVs,
  whose 3-SE test is taken element by element, at most 1.5% of the elements
  may exceed 3 SE and the mean squared z-score must stay below 1.6: the
  split-half check of the reference chains against themselves
  (test_reference_halves_calibrate_the_vector_bar) fixes what chance alone
  gives for the same statistics.
"""
from __future__ import annotations

import hashlib
import json
import os
import subprocess

import numpy as np
import pytest

import orc
from sodabayes_amd import plink, soda, synth

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
FIXTURE = os.path.join(ROOT, "tests", "golden", "c1_bayescpi.npz")
EXE = os.path.join(ROOT, "sodabayes_amd", "fortran", "build", "sodabayes_gpu")
Z3, VEC_FRAC, VEC_MSZ = 3.0, 0.015, 1.6


def fixture():
    z = np.load(FIXTURE)
    return z, json.loads(bytes(z["meta"]).decode())


def c1_data(meta):
    bed, y, _ = synth.dataset(meta["n"], meta["p"], seed=meta["data_seed"], nqtl=meta["nqtl"], h2=meta["h2"],
                              fmin=0.05, fmax=0.5)
    return bed, y


def write_data(d, meta):
    bed, y = c1_data(meta)
    plink.write_plink(os.path.join(d, "d"), bed, y)
    md5 = hashlib.md5(open(os.path.join(d, "d.bed"), "rb").read()).hexdigest()
    assert md5 == meta["bed_md5"], "the synthetic genotypes differ from the fixture's"
    return bed, y


def reference_chains(z, meta):
    """Per reference chain: means of (mu, Ve, Vara, model size, pi_2) and the
    effect means."""
    hyp = z["long/hyp"].astype(np.float64)        # chains x iterations x (size, mu, Vara, Ve)
    mod = z["long/mod"]                           # chains x (mu, Ve, Vara, NSNP, P1, P2)
    scal = np.stack([hyp[:, :, 1].mean(1), hyp[:, :, 3].mean(1), hyp[:, :, 2].mean(1),
                     hyp[:, :, 0].mean(1), mod[:, 5]], axis=1)
    return scal, z["long/eff"].astype(np.float64)


def z_scores(a, b):
    """Element-wise z of the difference of two sets of chain means
    (rows = chains), standard errors from each set's spread."""
    se2 = a.var(0, ddof=1) / len(a) + b.var(0, ddof=1) / len(b)
    return (a.mean(0) - b.mean(0)) / np.sqrt(np.maximum(se2, 1e-300))


def check_vector(name, za):
    frac = float(np.mean(np.abs(za) > Z3))
    msz = float(np.mean(za ** 2))
    print(f"  {name}: {len(za)} elements, {100 * frac:.2f}% beyond 3 SE, mean z^2 {msz:.3f}")
    assert frac <= VEC_FRAC and msz < VEC_MSZ, name


SCALARS = ("mu", "Ve", "Vara", "model size", "pi_2")


def test_fixture_data_and_shape():
    z, meta = fixture()
    assert meta["n"] == 2000 and meta["p"] == 10000 and meta["nqtl"] == 200
    assert "-blocksize" in meta["traj_args"] and meta["traj_args"][meta["traj_args"].index("-blocksize") + 1] == "1000"
    bed, _ = c1_data(meta)
    import tempfile
    with tempfile.TemporaryDirectory() as d:
        write_data(d, meta)
    scal, eff = reference_chains(z, meta)
    assert scal.shape == (16, 5) and eff.shape == (16, 10000)
    hyp = z["long/hyp"]
    assert hyp.shape[1] == meta["numit"] - meta["burnin"]


def test_reference_halves_calibrate_the_vector_bar():
    """Two halves of the reference's own chains (8 against 8) pass the
    vector bar: it is what chance gives, not a loosened test."""
    z, meta = fixture()
    scal, eff = reference_chains(z, meta)
    bed, _ = c1_data(meta)
    Z = orc.standardize(bed, meta["n"], orc.freq(bed, meta["n"]))
    gebv = eff @ Z.T
    for name, v in (("effects", eff), ("GEBV", gebv)):
        za = z_scores(v[0::2], v[1::2])
        check_vector(f"reference halves, {name}", za)
    zs = z_scores(scal[0::2], scal[1::2])
    print("  reference halves, scalars:", dict(zip(SCALARS, np.round(zs, 2))))


@pytest.mark.gpu
def test_replay_job_writes_the_reference_files(tmp_path):
    """The reference's 100-iteration job on the GPU program (REPLAY, FP64):
    the same .hyp/.eff/.mod/.frq/.log as the reference program."""
    from test_fortran_host import compare_file
    if not os.path.exists(EXE):
        pytest.fail(f"{EXE} is not built")
    z, meta = fixture()
    write_data(str(tmp_path), meta)
    r = subprocess.run([EXE, *meta["traj_args"]], cwd=tmp_path, env=dict(os.environ, OMP_NUM_THREADS="1"),
                       capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stdout[-3000:] + r.stderr[-3000:]
    assert "SODA on the GPU: rng replay , precision fp64 , chains 1" in r.stdout
    for sfx in (".hyp", ".eff", ".mod", ".frq", ".log"):
        compare_file("t" + sfx, bytes(z[f"traj/t{sfx}"]).decode(), (tmp_path / ("t" + sfx)).read_text())
    hyp = (tmp_path / "t.hyp").read_text().splitlines()
    assert len(hyp) == 101


def gpu_chains(meta, precision, chains, seed):
    """Counter-based chains of the reference's job: (scalars per chain,
    effect means per chain)."""
    bed, y = c1_data(meta)
    n, p = meta["n"], meta["p"]
    handles = [(chains, seed)] if precision == "mixed" else [(1, seed + k) for k in range(chains)]
    scal, eff = [], []
    for c, sd in handles:
        o = soda.SodaOptions(method="BayesCpi", blocksize=1000, numit=meta["numit"], burnin=meta["burnin"],
                             seed=sd, rng="counter", precision=precision, chains=c)
        s = soda.SodaSampler(o, n, p, y, bed=bed)
        s.setup()
        s.init()
        s.run(trace=False)
        for k in range(c):
            r = s.results(k)
            scal.append([r["mu_store"], r["vare_store"], r["vara_store"], r["included_store"],
                         r["pstore"][1]])
            eff.append(r["gstore"])
        s.close()
    return np.array(scal), np.array(eff)


@pytest.mark.gpu
@pytest.mark.parametrize("precision,chains", [("fp64", 8), ("mixed", 16)])
def test_posterior_within_3_mcse_of_reference_chains(precision, chains):
    z, meta = fixture()
    ref_s, ref_e = reference_chains(z, meta)
    got_s, got_e = gpu_chains(meta, precision, chains, 777)
    zs = z_scores(got_s, ref_s)
    for name, zv, a, b in zip(SCALARS, zs, got_s.mean(0), ref_s.mean(0)):
        print(f"  {precision} x{chains} {name}: {a:.6g} vs reference {b:.6g}  z={zv:.2f}")
    assert np.all(np.abs(zs) < Z3), dict(zip(SCALARS, zs))
    bed, _ = c1_data(meta)
    Z = orc.standardize(bed, meta["n"], orc.freq(bed, meta["n"]))
    check_vector(f"{precision} effects", z_scores(got_e, ref_e))
    check_vector(f"{precision} GEBV", z_scores(got_e @ Z.T, ref_e @ Z.T))

"""Writes tests/golden/c1_bayescpi.npz: reference runs at the plain-SODA
workload of n = 2,000 individuals and p = 10,000 SNPs.

Data: allele frequencies f_j ~ U(0.05, 0.5), genotypes x_ij ~ Bin(2, f_j),
200 QTL with N(0, 1) effects scaled to h2 = 0.5 (sodabayes_amd.synth, seed
DATA_SEED).  BayesCpi, SODA subsets of 1,000 consecutive SNPs (-blocksize
1000).  The reference's BayesCpi starts pi at 0.5 whatever -pi says
(models.f90:72-77, useroptions.f90:561-562), so the chains start there.

Two kinds of reference runs (oracle/_ref/sodabayes_ref, built from
/root/reference/src by ``make -C oracle ref``, -nthreads 1):

* trajectory: 100 iterations, 20 burn-in, seed TRAJ_SEED.  Every output
  file is kept; the GPU program must reproduce them (replayed draws).
* long chains: NCHAINS independent chains of 4,000 iterations with 2,000
  burn-in (seeds LONG_SEEDS).  From pi = 0.5 some chains need about 1,500
  iterations to reach the posterior's model size (about 200 of 10,000
  SNPs), so the 500-iteration burn-in of a plain 2,000-iteration run would
  mix the descent into the means.  Each chain's posterior means are kept: the
  .mod summary (mu, Ve, Vara, NSNP, P1, P2), the per-iteration trace
  columns of its .hyp after burn-in (for batch means), and the effect means
  of its .eff.  The spread of the chain means gives the Monte Carlo
  standard errors of the posterior checks.

The genotypes are not stored: the test regenerates them from DATA_SEED and
checks them against the stored MD5 of the .bed payload.

Run from the repository root:  python tests/golden/make_c1_golden.py
(about 10 minutes on 8 cores).
"""
from __future__ import annotations

import hashlib
import json
import os
import subprocess
import sys
import tempfile

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
from sodabayes_amd import plink, synth  # noqa: E402

REF = os.path.join(ROOT, "oracle", "_ref", "sodabayes_ref")
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "c1_bayescpi.npz")
N, P, NQTL, H2, DATA_SEED = 2000, 10000, 200, 0.5, 101
TRAJ_SEED = 11
TRAJ_ARGS = ["-bfile", "d", "-method", "BayesCpi", "-blocksize", "1000", "-numit", "100", "-burnin", "20",
             "-seed", str(TRAJ_SEED), "-out", "t"]
NCHAINS = 16
LONG_SEEDS = [1001 + k for k in range(NCHAINS)]
NUMIT, BURNIN = 4000, 2000
SUFFIXES = (".hyp", ".eff", ".mod", ".frq", ".log")


def long_args(seed: int) -> list[str]:
    return ["-bfile", "d", "-method", "BayesCpi", "-blocksize", "1000", "-numit", str(NUMIT), "-burnin",
            str(BURNIN), "-seed", str(seed), "-out", f"l{seed}"]


def data():
    bed, y, _ = synth.dataset(N, P, seed=DATA_SEED, nqtl=NQTL, h2=H2, fmin=0.05, fmax=0.5)
    return bed, y


def read_hyp(path: str) -> np.ndarray:
    lines = open(path).read().splitlines()[1:]
    return np.array([[float(v) for v in ln.split()] for ln in lines])


def read_mod(path: str) -> dict[str, float]:
    return {ln.split()[0]: float(ln.split()[1]) for ln in open(path).read().splitlines() if ln.strip()}


def read_eff(path: str) -> np.ndarray:
    return np.array([float(ln.split()[1]) for ln in open(path).read().splitlines()[1:]])


def main() -> None:
    if not os.path.exists(REF):
        sys.exit("build the reference first: make -C oracle ref")
    bed, y = data()
    store: dict[str, np.ndarray] = {}
    env = dict(os.environ, OMP_NUM_THREADS="1")
    with tempfile.TemporaryDirectory() as td:
        plink.write_plink(os.path.join(td, "d"), bed, y)
        md5 = hashlib.md5(open(os.path.join(td, "d.bed"), "rb").read()).hexdigest()
        jobs = [(TRAJ_ARGS, "t")] + [(long_args(s), f"l{s}") for s in LONG_SEEDS]
        running = []
        for args, pre in jobs:
            while len(running) >= (os.cpu_count() or 8):
                running.pop(0).wait()
            running.append(subprocess.Popen([REF, *args], cwd=td, env=env, stdout=subprocess.DEVNULL))
        for pr in running:
            pr.wait()
        for sfx in SUFFIXES:
            path = os.path.join(td, "t" + sfx)
            store[f"traj/t{sfx}"] = np.frombuffer(open(path).read().encode(), dtype=np.uint8)
        mods, hyps, effs = [], [], []
        for s in LONG_SEEDS:
            m = read_mod(os.path.join(td, f"l{s}.mod"))
            mods.append([m["Mean"], m["Ve"], m["Vara"], m["NSNP"], m["P1"], m["P2"]])
            h = read_hyp(os.path.join(td, f"l{s}.hyp"))
            hyps.append(h[BURNIN:, 1:5])   # ModelSize, mu, Vara, Ve after burn-in
            effs.append(read_eff(os.path.join(td, f"l{s}.eff")))
        store["long/mod"] = np.array(mods)                     # chains x (mu, Ve, Vara, NSNP, P1, P2)
        store["long/hyp"] = np.array(hyps, dtype=np.float32)   # chains x iterations x 4
        store["long/eff"] = np.array(effs, dtype=np.float32)   # chains x p
    meta = {"n": N, "p": P, "nqtl": NQTL, "h2": H2, "data_seed": DATA_SEED, "bed_md5": md5,
            "traj_args": TRAJ_ARGS, "long_seeds": LONG_SEEDS, "numit": NUMIT, "burnin": BURNIN}
    store["meta"] = np.frombuffer(json.dumps(meta).encode(), dtype=np.uint8)
    np.savez_compressed(OUT, **store)
    mods = store["long/mod"]
    print("chain means (mu, Ve, Vara, NSNP, P1, P2):", mods.mean(0))
    print("chain SE:", mods.std(0, ddof=1) / np.sqrt(NCHAINS))
    print(f"wrote {OUT} ({os.path.getsize(OUT)} bytes)")


if __name__ == "__main__":
    main()

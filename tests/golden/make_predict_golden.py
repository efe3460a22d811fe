"""Writes the prediction fixture from the data files the reference ships.

The reference's examples hold three prediction runs (``-predict``) on the
same genotypes, examples/data/simdata1_pred.{bed,bim,fam}, with the model
files (.frq, .eff, .mod) of three training runs and the resulting genomic
values (.gv, written by write_dgv, utilities.f90:625-633):

  example1: BayesR   (bayesr_mcmc_example_1.*  -> pred_bayesr_mcmc_example_1.gv)
  example3: BayesR   (bayesr_mcmc_example_2.*  -> pred_bayesr__example_3.gv)
  example4: BayesRc  (bayesrc_mcmc_example_4.* -> pred_bayesrc_mcmc_example_4.gv)

predict_kat.npz keeps the first NKEEP individuals (every prediction depends
only on its own genotypes): the SNP-major .bed payload cut to their bytes,
and per run the frequencies, effects and mean as the reference reads them
(list-directed reads of the text files, utilities.f90:543, 647-656) and the
expected values as printed (E15.7).

Run from the repository root:  python tests/golden/make_predict_golden.py
"""
from __future__ import annotations

import os

import numpy as np

REF = "/root/reference/examples"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "predict_kat.npz")
NKEEP = 200   # multiple of 4: whole .bed bytes

RUNS = {
    "example1": ("example1", "bayesr_mcmc_example_1", "pred_bayesr_mcmc_example_1.gv"),
    "example3": ("example3", "bayesr_mcmc_example_2", "pred_bayesr__example_3.gv"),
    "example4": ("example4", "bayesrc_mcmc_example_4", "pred_bayesrc_mcmc_example_4.gv"),
}


def second_column(path, skip=0):
    vals = []
    with open(path) as f:
        for k, line in enumerate(f):
            if k < skip or not line.strip():
                continue
            vals.append(float(line.split()[1].replace("D", "E")))
    return np.array(vals)


def main():
    data = os.path.join(REF, "data")
    nind = sum(1 for _ in open(os.path.join(data, "simdata1_pred.fam")))
    nloci = sum(1 for _ in open(os.path.join(data, "simdata1_pred.bim")))
    raw = np.fromfile(os.path.join(data, "simdata1_pred.bed"), dtype=np.uint8)
    assert bytes(raw[:3]) == bytes([0x6C, 0x1B, 0x01])
    nb = (nind + 3) // 4
    bed = raw[3:].reshape(nloci, nb)[:, :NKEEP // 4].copy()
    out = dict(nind=np.int64(NKEEP), nind_full=np.int64(nind), nloci=np.int64(nloci), bed=bed,
               runs=np.array(list(RUNS)))
    for name, (folder, stem, gv) in RUNS.items():
        d = os.path.join(REF, folder)
        freq = second_column(os.path.join(d, stem + ".frq"))
        eff = second_column(os.path.join(d, stem + ".eff"), skip=1)
        mu = second_column(os.path.join(d, stem + ".mod"))[0]
        pred = second_column(os.path.join(d, gv))
        assert freq.size == nloci and eff.size == nloci and pred.size == nind
        out[f"{name}_freq"] = freq
        out[f"{name}_eff"] = eff
        out[f"{name}_mu"] = np.float64(mu)
        out[f"{name}_pred"] = pred[:NKEEP]
    np.savez_compressed(OUT, **out)
    print(f"wrote {OUT}: n={NKEEP} of {nind}, p={nloci}")


if __name__ == "__main__":
    main()

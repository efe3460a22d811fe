"""Writes tests/golden/fortran_jobs.npz: whole reference jobs and their files.

Runs the reference program (oracle/_ref/sodabayes_ref, built from
/root/reference/src by ``make -C oracle ref``) with its default -nthreads 1
on small seeded synthetic PLINK data, from the command line and from a .par
file with two jobs, plus a -predict job on the first job's model files, and
-SODAOff jobs (the single-site sampler) for every method, and mixtures of 5
and 9 classes.  For
each job the fixture keeps the input files, the arguments and every output
file (.hyp, .eff, .mod, .frq, .gv, .log).  The GPU test runs the Fortran
program with its SODA branch on the GPU (sodabayes_amd/fortran) on the same
inputs and compares the files.

Run from the repository root:  python tests/golden/make_fortran_golden.py
"""
from __future__ import annotations

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
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "fortran_jobs.npz")
R4 = ["-nmix", "4", "-mix", "0.0,0.0001,0.001,0.01"]
R5 = ["-nmix", "5", "-mix", "0.0,0.0001,0.001,0.01,0.1"]
R9 = ["-nmix", "9", "-mix", "0.0,0.00001,0.0001,0.0003,0.001,0.003,0.01,0.03,0.1"]

JOB_C = """begin jobC
-bfile d
-method BayesC
-pi 0.95        ! only a .par file can set -pi
-numit 24
-burnin 6
-thin 2
-blocksize 50
-seed 31
-out c
end
"""
JOB_B = """begin jobB
-bfile d
-method BayesB
-pi 0.9
-df_va 4
-scale_va 0.002
-nblocks 4
-numit 16
-burnin 4
-seed 32
-out b
end
"""
# Both jobs in one file.  The reference stops at the second job of a .par
# file (load_geno leaves the .bim connected to unit 40, utilities.f90:414,
# and get_size opens it again, :268), so its files come from one run per
# job; no state carries from job C to job B (df_va > 0, fresh seed).
PAR_TWO_JOBS = "# two jobs in one parameter file\n" + JOB_C + JOB_B

# name: (nind, nloci, data seed, groups, arguments (None: the .par file), output prefixes)
CASES = {
    "bayesr_cmdline": (200, 240, 41, 0, ["-bfile", "d", "-method", "BayesR", *R4, "-blocksize", "60",
                                         "-numit", "30", "-burnin", "10", "-seed", "11", "-out", "r"], ["r"]),
    "bayesa_rblocks": (150, 180, 42, 0, ["-bfile", "d", "-method", "BayesA", "-rblocks", "-blocksize", "45",
                                         "-numit", "20", "-burnin", "5", "-df_va", "4", "-scale_va", "0.002",
                                         "-seed", "15", "-out", "a"], ["a"]),
    "bayesrc_groups": (180, 220, 43, 4, ["-bfile", "d", "-method", "BayesRc", *R4, "-grpfile", "d.grp",
                                         "-blocksize", "55", "-numit", "25", "-burnin", "5", "-seed", "16",
                                         "-out", "rc"], ["rc"]),
    "bayescpi_nblocks": (170, 233, 44, 0, ["-bfile", "d", "-method", "BayesCpi", "-nblocks", "7",
                                           "-numit", "15", "-burnin", "5", "-seed", "17", "-out", "cp"], ["cp"]),
    "par_two_jobs": (160, 200, 45, 0, None, ["c", "b"]),
    # -SODAOff: the single-site sampler (models.f90:607-962)
    "bayesr_sodaoff": (150, 130, 46, 0, ["-bfile", "d", "-method", "BayesR", *R4, "-SODAOff", "-numit", "20",
                                         "-burnin", "5", "-seed", "21", "-out", "ro"], ["ro"]),
    "bayesa_sodaoff": (140, 120, 47, 0, ["-bfile", "d", "-method", "BayesA", "-SODAOff", "-df_va", "4",
                                         "-scale_va", "0.002", "-numit", "15", "-burnin", "3", "-seed", "22",
                                         "-out", "ao"], ["ao"]),
    "bayesb_sodaoff": (130, 110, 48, 0, ["-bfile", "d", "-method", "BayesB", "-SODAOff", "-df_va", "4",
                                         "-scale_va", "0.002", "-numit", "15", "-burnin", "3", "-seed", "23",
                                         "-out", "bo"], ["bo"]),
    "bayesrc_sodaoff": (140, 150, 49, 3, ["-bfile", "d", "-method", "BayesRc", *R4, "-grpfile", "d.grp",
                                          "-SODAOff", "-numit", "15", "-burnin", "3", "-seed", "24",
                                          "-out", "rco"], ["rco"]),
    "bayescpi_sodaoff": (150, 140, 50, 0, ["-bfile", "d", "-method", "BayesCpi", "-SODAOff", "-numit", "20",
                                           "-burnin", "5", "-thin", "2", "-seed", "25", "-out", "cpo"], ["cpo"]),
    # mixtures of more than four classes (5 and 9), with SODA blocks and single-site
    "bayesr5_blocks": (160, 150, 51, 0, ["-bfile", "d", "-method", "BayesR", *R5, "-blocksize", "50", "-numit", "20",
                                         "-burnin", "5", "-seed", "26", "-out", "r5"], ["r5"]),
    "bayesr5_sodaoff": (140, 120, 52, 0, ["-bfile", "d", "-method", "BayesR", *R5, "-SODAOff", "-numit", "20",
                                          "-burnin", "5", "-seed", "27", "-out", "r5o"], ["r5o"]),
    "bayesr9_blocks": (150, 140, 53, 0, ["-bfile", "d", "-method", "BayesR", *R9, "-blocksize", "40", "-numit", "15",
                                         "-burnin", "3", "-seed", "28", "-out", "r9"], ["r9"]),
    "bayesr9_sodaoff": (130, 100, 54, 0, ["-bfile", "d", "-method", "BayesR", *R9, "-SODAOff", "-numit", "15",
                                          "-burnin", "3", "-seed", "29", "-out", "r9o"], ["r9o"]),
    # more individuals than one pass of the GPU single-site sweep covers with
    # one per thread (3 per thread; 8 per thread over two passes)
    "bayesr_sodaoff_n1100": (1100, 90, 55, 0, ["-bfile", "d", "-method", "BayesR", *R4, "-SODAOff", "-numit", "10",
                                                "-burnin", "2", "-seed", "30", "-out", "rn"], ["rn"]),
    "bayescpi_sodaoff_n4500": (4500, 70, 56, 0, ["-bfile", "d", "-method", "BayesCpi", "-SODAOff", "-numit", "8",
                                                  "-burnin", "2", "-seed", "31", "-out", "cn"], ["cn"]),
    # SODA blocks over more individuals than one 2,048-individual genotype chunk
    "bayesr_blocks_n4500": (4500, 240, 57, 0, ["-bfile", "d", "-method", "BayesR", *R4, "-blocksize", "80",
                                               "-numit", "8", "-burnin", "2", "-seed", "32", "-out", "rb"], ["rb"]),
}
PREDICT_ARGS = ["-predict", "-bfile", "d", "-model", "r.mod", "-freq", "r.frq", "-effects", "r.eff",
                "-out", "p"]
SUFFIXES = (".hyp", ".eff", ".mod", ".frq", ".gv", ".log")


def run_ref(cwd: str, args: list[str]) -> None:
    env = dict(os.environ, OMP_NUM_THREADS="1")
    r = subprocess.run([REF, *args], cwd=cwd, env=env, capture_output=True, text=True, timeout=600)
    if r.returncode != 0:
        raise RuntimeError(f"reference failed: {args}\n{r.stdout[-2000:]}\n{r.stderr[-2000:]}")


def outputs(cwd: str, prefixes: list[str]) -> dict[str, str]:
    got = {}
    for pre in prefixes:
        for sfx in SUFFIXES:
            path = os.path.join(cwd, pre + sfx)
            if os.path.exists(path):
                got[pre + sfx] = open(path).read()
    return got


def main() -> None:
    if not os.path.exists(REF):
        sys.exit("build the reference first: make -C oracle ref")
    store: dict[str, np.ndarray] = {}
    manifest = {}
    for name, (n, p, seed, ngrp, args, prefixes) in CASES.items():
        bed, y, _ = synth.dataset(n, p, seed=seed, nqtl=15, h2=0.5)
        with tempfile.TemporaryDirectory() as td:
            plink.write_plink(os.path.join(td, "d"), bed, y)
            inputs = {f: open(os.path.join(td, f), "rb").read() for f in ("d.bed", "d.bim", "d.fam")}
            if ngrp:
                plink.write_groups(os.path.join(td, "d.grp"), synth.snp_groups(p, seed))
                inputs["d.grp"] = open(os.path.join(td, "d.grp"), "rb").read()
            if args is None:
                for k, job in enumerate((JOB_C, JOB_B)):
                    with open(os.path.join(td, f"job{k}.par"), "w") as f:
                        f.write(job)
                    run_ref(td, [f"job{k}.par"])
                with open(os.path.join(td, "jobs.par"), "w") as f:
                    f.write(PAR_TWO_JOBS)
                inputs["jobs.par"] = PAR_TWO_JOBS.encode()
                args = ["jobs.par"]
            else:
                run_ref(td, args)
            outs = outputs(td, prefixes)
            case = {"args": args, "inputs": sorted(inputs), "outputs": sorted(outs)}
            if name == "bayesr_cmdline":
                # -predict on this job's model files (SODABayes.f90:145-148)
                run_ref(td, PREDICT_ARGS)
                pred = outputs(td, ["p"])
                case["predict_args"] = PREDICT_ARGS
                case["predict_outputs"] = sorted(pred)
                outs.update({"predict/" + k: v for k, v in pred.items()})
        for f, b in inputs.items():
            store[f"{name}/in/{f}"] = np.frombuffer(b, dtype=np.uint8)
        for f, t in outs.items():
            store[f"{name}/out/{f}"] = np.frombuffer(t.encode(), dtype=np.uint8)
        manifest[name] = case
        print(f"{name}: {len(outs)} output files")
    store["manifest"] = np.frombuffer(json.dumps(manifest).encode(), dtype=np.uint8)
    np.savez_compressed(OUT, **store)
    print(f"wrote {OUT} ({os.path.getsize(OUT)} bytes)")


if __name__ == "__main__":
    main()

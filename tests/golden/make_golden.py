"""Writes the golden fixtures in this folder from the reference's own code.

Runs oracle/_ref/golden_driver (the reference's Fortran modules compiled from
/root/reference/src by ``make -C oracle ref``) on small seeded synthetic
PLINK data and stores inputs + full-precision outputs as .npz fixtures:

  soda_<case>.npz : bed payload, phenotype, SNP groups, options (JSON), and
                    every record of the driver's dump (setup products and the
                    final chain state after ``numit`` iterations)
  rng_<seed>.npz  : stream-0 variates of parallelrng.f90
  mat_<case>.npz  : choleskyPD / choleskyPSD / power_method on given matrices

Run from the repository root:  python tests/golden/make_golden.py
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
from sodabayes_amd.dump import read_dump  # noqa: E402

DRIVER = os.path.join(ROOT, "oracle", "_ref", "golden_driver")
OUT = os.path.dirname(os.path.abspath(__file__))

R4 = dict(nmix=4, mix="0.0,0.0001,0.001,0.01")

# name: (nind, nloci, data seed, dataset kwargs, reference options)
CASES = {
    "bayesr": (200, 240, 5, {}, dict(method="BayesR", blocksize=60, numit=30, burnin=10, seed=11, **R4)),
    "bayescpi": (160, 200, 6, {}, dict(method="BayesCpi", blocksize=50, numit=30, burnin=10, seed=12)),
    "bayesc": (160, 200, 7, {}, dict(method="BayesC", pi=0.95, blocksize=50, numit=30, burnin=10, seed=13)),
    "bayesb": (150, 180, 8, {}, dict(method="BayesB", pi=0.9, blocksize=45, numit=20, burnin=5,
                                     df_va=4.0, scale_va=0.002, seed=14)),
    "bayesa": (150, 180, 9, {}, dict(method="BayesA", blocksize=45, numit=20, burnin=5,
                                     df_va=4.0, scale_va=0.002, seed=15)),
    "bayesrc": (180, 220, 10, {"groups": 4}, dict(method="BayesRc", blocksize=55, numit=25, burnin=5,
                                                   seed=16, **R4)),
    "nblocks": (170, 233, 11, {}, dict(method="BayesR", nblocks=7, numit=15, burnin=5, seed=17, **R4)),
    "rblocks": (160, 190, 12, {}, dict(method="BayesCpi", blocksize=64, rblocks=True, numit=15,
                                       burnin=5, seed=18)),
    "dfva0": (160, 150, 13, {}, dict(method="BayesR", blocksize=40, numit=15, burnin=5, df_va=0.0,
                                     scale_va=0.5, seed=19, **R4)),
    "thin": (140, 160, 14, {}, dict(method="BayesCpi", blocksize=40, numit=24, burnin=6, thin=3,
                                    seed=20)),
    "ragged": (203, 150, 15, {"missing": 0.02, "mono": (3, 77)},
               dict(method="BayesR", blocksize=50, numit=15, burnin=5, seed=21, **R4)),
    "resync": (96, 80, 16, {}, dict(method="BayesCpi", blocksize=20, numit=1002, burnin=900, seed=22)),
    "ldblock": (220, 200, 17, {"ld_block": 25},
                dict(method="BayesR", blocksize=50, numit=20, burnin=5, seed=23, **R4)),
    # d of block 1 lowered by 30% (test hook): dI - G is indefinite there, so
    # the reference's pivoted-Cholesky fallback, skip lists and skipped-SNP
    # updates run (BayesA: the one method whose skipped-SNP update reads no
    # per-thread-only state, models.f90:178-193)
    "psd_bayesa": (150, 180, 24, {"dshift": (1, 0.3)},
                   dict(method="BayesA", blocksize=45, numit=20, burnin=5, df_va=4.0, scale_va=0.002,
                        seed=25)),
}


def make_data(nind, nloci, seed, kw):
    x, _ = synth.genotypes(nind, nloci, seed, ld_block=kw.get("ld_block", 0))
    for j in kw.get("mono", ()):
        x[:, j] = 0
    y, _, _ = synth.phenotype(x, seed, max(4, nloci // 20), 0.5)
    missing = None
    if kw.get("missing"):
        rng = np.random.default_rng(seed + 1)
        missing = rng.random(x.shape) < kw["missing"]
    bed = plink.encode_bed(x, missing)
    grp = synth.snp_groups(nloci, seed) if kw.get("groups") else None
    return bed, y, grp


def ref_args(opts, prefix, grpfile):
    args = ["-bfile", prefix, "-out", prefix + "_out", "-nthreads", "1"]
    for k, v in opts.items():
        if isinstance(v, bool):
            if v:
                args.append("-" + k)
        else:
            args += ["-" + k, str(v)]
    if grpfile:
        args += ["-grpfile", grpfile]
    return args


def run_driver(mode, args=(), env_extra=None):
    with tempfile.TemporaryDirectory() as td:
        out = os.path.join(td, "dump.bin")
        env = dict(os.environ, GOLDEN_MODE=mode, GOLDEN_OUT=out, OMP_NUM_THREADS="1")
        env.update(env_extra or {})
        if mode == "soda":
            optfile = os.path.join(td, "opts.txt")
            with open(optfile, "w") as f:
                f.write("\n".join(args) + "\n")
            env["GOLDEN_OPTS"] = optfile
            # one harmless flag on the command line: with none the
            # reference prints its help and stops (useroptions.f90:210-217)
            args = ("-SODAOn",)
        r = subprocess.run([DRIVER, *args], env=env, cwd=td, capture_output=True, text=True)
        dump = read_dump(out) if os.path.exists(out) else {}
        if r.returncode != 0 or "STOP" in r.stderr or not dump:
            raise RuntimeError(f"golden driver failed ({mode}): {r.stderr[-500:]}")
        return dump


def soda_case(name, spec):
    nind, nloci, seed, kw, opts = spec
    bed, y, grp = make_data(nind, nloci, seed, kw)
    with tempfile.TemporaryDirectory() as td:
        prefix = os.path.join(td, "d")
        plink.write_plink(prefix, bed, y)
        gfile = None
        if grp is not None:
            gfile = os.path.join(td, "grp.txt")
            plink.write_groups(gfile, grp)
        env = None
        if "dshift" in kw:
            blk, frac = kw["dshift"]
            d0 = run_driver("soda", ref_args(opts, prefix, gfile))
            shift = float(np.round(frac * float(np.ravel(d0["xxd"])[blk]), 3))
            opts = dict(opts, dshift_block=blk, dshift=shift)
            env = {"GOLDEN_DSHIFT_BLOCK": str(blk), "GOLDEN_DSHIFT": repr(shift)}
        dump = run_driver("soda", ref_args({k: v for k, v in opts.items() if not k.startswith("dshift")},
                                            prefix, gfile), env)
    rec = {f"out_{k}": v for k, v in dump.items()}
    rec.update(bed=bed, why=y, options=json.dumps(opts), nind=nind, nloci=nloci)
    if grp is not None:
        rec["grp"] = grp
    np.savez_compressed(os.path.join(OUT, f"soda_{name}.npz"), **rec)


def rng_case(seed):
    dump = run_driver("rng", env_extra={"GOLDEN_SEED": str(seed)})
    np.savez_compressed(os.path.join(OUT, f"rng_{seed}.npz"), seed=seed, **dump)


def mat_case(name, A, seed=0):
    n = A.shape[0]
    rec = {"A": A}
    with tempfile.TemporaryDirectory() as td:
        fin = os.path.join(td, "in.bin")
        with open(fin, "wb") as f:
            f.write(np.array([n], dtype="<i8").tobytes())
            f.write(np.asfortranarray(A).tobytes(order="F"))
        for mode in ("pd", "psd", "power"):
            d = run_driver(mode, env_extra={"GOLDEN_IN": fin, "GOLDEN_SEED": str(seed)})
            for k, v in d.items():
                rec[f"{mode}_{k}"] = v
    rec["seed"] = seed
    np.savez_compressed(os.path.join(OUT, f"mat_{name}.npz"), **rec)


def main():
    if not os.path.exists(DRIVER):
        sys.exit("build the reference driver first: make -C oracle ref")
    for name, spec in CASES.items():
        soda_case(name, spec)
        print("soda", name)
    for seed in (123456789, 7, -42):
        rng_case(seed)
        print("rng", seed)
    rng = np.random.default_rng(3)
    B = rng.standard_normal((40, 12))
    G = B @ B.T                                  # PSD, rank 12
    mat_case("psd_rank12", G + 0.0, seed=5)
    C = rng.standard_normal((30, 30))
    mat_case("spd30", C @ C.T + 30 * np.eye(30), seed=6)
    D = rng.standard_normal((25, 25))
    M = D @ D.T
    lam = np.linalg.eigvalsh(M)[-1]
    mat_case("dI_minus_G", (lam + 0.001) * np.eye(25) - M, seed=7)
    print("matrices done")


if __name__ == "__main__":
    main()

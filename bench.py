#!/usr/bin/env python3
"""SODA Gibbs sweep throughput on one MI355X.

Workload (the metric's configuration): BayesR, n = 50,000 individuals,
p = 500,000 SNPs, SODA subsets of 4,096 SNPs, 4-class mixture
{0, 1e-4, 1e-3, 1e-2} sigma2_g.  Synthetic genotypes f_j ~ U(0.01, 0.5),
x_ij ~ Bin(2, f_j), generated directly in HBM as a SNP-major PLINK payload;
phenotype from 5,000 t(4) QTL at h2 = 0.4.  One step = one Gibbs iteration
(p SNP-effect updates).  Inputs are resident in HBM before timing starts.

Two passes share the genotypes: the plain FP64 single-chain path, then the
mixed-precision batched path (``--chains`` chains per launch, int8-limb MFMA
products).  ``value`` is the batched path's SNP-effect updates/s summed over
chains; ``fp64_gpu`` and ``speedup_vs_fp64_gpu`` give the FP64 path beside
it.  The roofline is that of the dominant kernel group (the block products),
timed with HIP events on the library's stream in a separate pass; the
reference CPU baseline is timed on this host.
"""
from __future__ import annotations

import argparse
import json
import os
import re
import subprocess
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E
INT8_PEAK_TOPS = 5000.0  # MI355X dense int8 MFMA

# kernel timing slots (sodabayes_amd._lib.KT_SLOTS) -> kernel names
KERNEL_NAMES = {
    "rhs": "k_rhs (X_s'r + U_s'r_a, FP64)", "draw": "k_draw_counter", "list": "k_list",
    "update": "k_update (r += Z delta, r_a += U delta)", "rquant": "k_rstat + k_quant (r limbs, first block)",
    "xq": "k_prod (X_s'Q + U_s'r_a, int8 MFMA)", "ut0": "U_s'r_a (inside k_prod)",
    "drawb": "k_draw_b (limb sums + draw)", "dquant": "k_dquant (delta limbs)",
    "xupd": "k_upd (R += X_s C + U_s diag(s) c, int8 MFMA)", "ut1": "U_s diag(s) c (inside k_upd)",
    "rared": "k_rared_quant (r_a reduction + next block's r limbs)",
}

# HBM bytes per launch from rocprofv3 --pmc FETCH_SIZE (x2, the factor a 1 GiB
# device copy shows on this GPU) + WRITE_SIZE (x1), per (path, kernel slot),
# measured on 4,096-SNP blocks at n = 50,000 (64 chains, two chain groups,
# for the batched path), the per-launch shape of this workload
# (profiles/soda_bayesr_n50k_bs4096_x64_pmc_traffic.md); None where not measured
TRAFFIC_CHAINS = 64
TRAFFIC = {
    ("mixed", "xupd"): int((87088 * 2 + 30922) * 1024),  # k_upd<6, 2> with the U_s diag(s) c pairs
    ("mixed", "xq"): int((100614 * 2 + 19565) * 1024),    # k_prod<6, 2> with the U_s'r_a pairs
    ("mixed", "drawb"): int((12982 * 2 + 7456) * 1024),  # k_draw_b<4>
    ("fp64", "rhs"): int((60338 * 2 + 920) * 1024),      # k_rhs
    ("fp64", "update"): int((50642 * 2 + 604) * 1024),   # k_update (k_ra_sum not counted)
}


def parse_args():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=30)
    ap.add_argument("--nind", type=int, default=50000)
    ap.add_argument("--nloci", type=int, default=500000)
    ap.add_argument("--blocksize", type=int, default=4096)
    ap.add_argument("--seed", type=int, default=20240607)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-only", action="store_true", help="only the reference CPU baseline (prints its JSON)")
    ap.add_argument("--cpu-nloci", type=int, default=8192, help="SNPs of the CPU baseline's sample")
    ap.add_argument("--cpu-blocksize", type=int, default=1024, help="SODA subset size of the CPU baseline's sample")
    ap.add_argument("--cpu-iters", type=int, default=100, help="iterations of the CPU baseline's sample")
    ap.add_argument("--no-accuracy", action="store_true", help="skip the mixed-vs-FP64 accuracy report")
    ap.add_argument("--no-roofline", action="store_true")
    ap.add_argument("--chains", type=int, default=64,
                    help="chains batched on the mixed path (64: the best of 16, 32, 48, 64)")
    ap.add_argument("--chains-only", action="store_true", help="skip the FP64 single-chain pass")
    ap.add_argument("--fp64-only", action="store_true", help="only the FP64 single-chain pass (no JSON line)")
    ap.add_argument("--no-folds", action="store_true", help="skip the cross-validation folds pass")
    ap.add_argument("--no-traits", action="store_true", help="skip the batched traits pass")
    ap.add_argument("--no-single-site", action="store_true", help="skip the single-site (-SODAOff) pass")
    ap.add_argument("--no-c3", action="store_true", help="skip the BayesA / BayesB pass on the headline genotypes")
    ap.add_argument("--copy-probe", action="store_true",
                    help="first run one plain 1 GiB device-to-device copy (a known-size reference for "
                         "rocprofv3 FETCH_SIZE / WRITE_SIZE counter runs)")
    return ap.parse_args()


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def synth_on_device(torch, n, p, seed, fmin=0.01, fmax=0.5, nqtl=5000, h2=0.4):
    """SNP-major PLINK payload (p x ceil(n/4) bytes) generated in HBM, with a
    phenotype from nqtl t(4) QTL at heritability h2 (synth.device_dataset)."""
    from sodabayes_amd import synth

    bed, freq, y, _ = synth.device_dataset(torch, n, p, seed, fmin=fmin, fmax=fmax, nqtl=nqtl, h2=h2)
    return bed, freq, y


def traits_on_device(torch, n, p, seed, ntraits=16, nqtl=3000, shared=0.5):
    """PLINK payload in HBM with ntraits correlated phenotypes (synth.device_traits)."""
    from sodabayes_amd import synth

    return synth.device_traits(torch, n, p, seed, ntraits=ntraits, nqtl=nqtl, shared=shared)


def cpu_cores() -> int:
    """Host threads the baseline may use: the CPU share the environment
    grants this job (OMP_NUM_THREADS, e.g. 16 per GPU), else every core."""
    try:
        share = int(os.environ.get("OMP_NUM_THREADS", "0"))
    except ValueError:
        share = 0
    n = os.cpu_count() or 1
    return min(share, n) if share > 0 else n


def cpu_baseline(seed: int, nloci: int = 8192, blocksize: int = 1024, numit: int = 100) -> dict | None:
    """The reference program (oracle/_ref/o3/sodabayes_ref: the reference
    sources built -O3 -fopenmp by oracle/Makefile) on a bounded sample of
    the workload: the same n = 50,000 individuals and BayesR mixture, nloci
    SNPs in subsets of blocksize, numit iterations, -nthreads = every core
    the job may use.  SNP-updates/s = nloci * numit / the MCMC seconds of its
    .log; the setup seconds (serial power method and Cholesky) are reported
    apart."""
    exe = os.path.join(ROOT, "oracle", "_ref", "o3", "sodabayes_ref")
    if not os.path.exists(exe):
        log("cpu baseline skipped: oracle/_ref/o3/sodabayes_ref not built")
        return None
    from sodabayes_amd import plink, synth

    n, p, numit, bs = 50000, nloci, numit, blocksize
    cores = cpu_cores()
    with tempfile.TemporaryDirectory() as td:
        bed, y, _ = synth.dataset(n, p, seed=seed % 100000, nqtl=40, h2=0.4, fmin=0.01, fmax=0.5)
        plink.write_plink(os.path.join(td, "d"), bed, y)
        cmd = [exe, "-bfile", "d", "-method", "BayesR", "-nmix", "4", "-mix", "0.0,0.0001,0.001,0.01",
               "-numit", str(numit), "-burnin", "0", "-blocksize", str(bs), "-nthreads", str(cores),
               "-seed", "7", "-out", "r"]
        env = dict(os.environ, OMP_NUM_THREADS=str(cores))
        log(f"cpu baseline: reference program (-O3), n={n} p={p} blocksize={bs} {numit} its, {cores} threads")
        r = subprocess.run(cmd, cwd=td, env=env, capture_output=True, text=True, timeout=3000)
        path = os.path.join(td, "r.log")
        txt = open(path, errors="replace").read() if os.path.exists(path) else ""
        m = re.search(r"Setup:\s*([\d.]+)\s+MCMC:\s*([\d.]+)", txt)
        if r.returncode != 0 or not m:
            log(f"cpu baseline failed: {r.stderr[-300:]}")
            return None
        setup, mcmc = float(m.group(1)), float(m.group(2))
    return {"value": p * numit / mcmc, "unit": "SNP-effect updates/s", "cores": cores, "kind": "reference",
            "setup_seconds": setup, "mcmc_seconds": mcmc,
            "sample": f"reference sodabayes (amdflang -O3 -fopenmp) BayesR n={n} p={p} blocksize={bs} "
                      f"{numit} iterations, -nthreads {cores}; MCMC {mcmc:.2f} s, setup {setup:.2f} s (not counted)"}


def accuracy(torch, s, bed, freq, chains: int) -> dict:
    """After the timed sweeps, outside the timed region: the worst norm-wise
    error of the path's block products X_s'r + U_s'r_a against an FP64
    evaluation of the same state over 8 blocks and every chain, and the
    residual identity r = y - Z beta - mu over 4,096 sampled individuals and
    every chain (sodabayes_amd.check)."""
    import numpy as np

    from sodabayes_amd import check

    blocks = check.sample_blocks(s.nblocks, 8)
    rows = np.random.default_rng(12345).choice(s.nind, size=min(4096, s.nind), replace=False)
    t0 = time.time()
    out = {"products_rel_err": check.products_error(torch, s, bed, freq, blocks, list(range(chains))),
           "residual_identity_rel_err": check.residual_identity(torch, s, bed, freq, list(range(chains)), rows=rows),
           "blocks": blocks, "chains": chains, "sampled_individuals": int(len(rows)),
           "iterations_before": int(s.rep)}
    log(f"accuracy ({time.time() - t0:.1f} s): products {out['products_rel_err']:.3e}, residual identity "
        f"{out['residual_identity_rel_err']:.3e} after {s.rep} iterations")
    return out


def run_path(soda, torch, a, why, freq, bed_ptr, precision: str, chains: int, bed=None):
    """Setup + timed sweeps of one product path; returns its metrics."""
    n, p, bs = a.nind, a.nloci, a.blocksize
    troof = 0 if a.no_roofline else max(3, a.steps // 3)
    opts = soda.SodaOptions(method="BayesR", nmix=4, mix="0.0,0.0001,0.001,0.01", blocksize=bs,
                            numit=a.warmup + a.steps + troof, burnin=0, seed=a.seed, rng="counter",
                            precision=precision, chains=chains)
    s = soda.SodaSampler(opts, n, p, why, freq=freq, bed_device_ptr=bed_ptr)
    t0 = time.time()
    s.setup()
    setup_s = time.time() - t0
    info = s.info_raw()
    log(f"[{precision} x{chains}] setup (gram + power method + cholesky, {s.nblocks} blocks): "
        f"{setup_s:.1f} s, power iterations <= {int(info[11])}")
    s.init()
    s.run(a.warmup, trace=False)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    s.run(a.steps, trace=False)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    out = {"setup_s": setup_s, "power_iters": int(info[11]), "dt": dt, "ms": dt * 1e3 / a.steps,
           "value": p * chains * a.steps / dt}
    log(f"[{precision} x{chains}] {out['ms']:.2f} ms/iteration, {out['value']:.4g} SNP-updates/s")
    if precision == "mixed":
        inf = s.info()
        out["tile_shape"] = {"chunk_supersteps": inf["chunk_supersteps"], "rows_per_wave": inf["rows_per_wave"]}
        log(f"[{precision} x{chains}] tile shape: {inf['chunk_supersteps']} super-steps per X_s'Q chunk, "
            f"{inf['rows_per_wave']} row tiles per wave (tuned: {inf['tuned_block_us']:.1f} us for block 0)")
    if troof:
        # separate pass with HIP events around every block kernel
        os.environ["SODABAYES_KERNEL_TIMING"] = "1"
        before = s.kernel_times()
        s.run(troof, trace=False)
        after = s.kernel_times()
        os.environ.pop("SODABAYES_KERNEL_TIMING", None)
        kernels = {}
        for k, (ms, n, b) in after.items():
            dn = n - before[k][1]
            if dn <= 0:
                continue
            avg_us = (ms - before[k][0]) / dn * 1e3
            bpl = (b - before[k][2]) / dn
            kernels[k] = {"name": KERNEL_NAMES[k], "avg_us": round(avg_us, 2), "launches": int(dn),
                          "algorithmic_bytes": int(bpl), "gbs": round(bpl / (avg_us * 1e-6) / 1e9, 1),
                          "share": 0.0}
        total = sum(v["avg_us"] * v["launches"] for v in kernels.values())
        for v in kernels.values():
            v["share"] = round(v["avg_us"] * v["launches"] / total, 4)
        dom = max(kernels, key=lambda k: kernels[k]["avg_us"] * kernels[k]["launches"])
        d = kernels[dom]
        out["kernels"] = kernels
        # The int8 product launches are bound by whichever of HBM and MFMA
        # needs more time at peak for their algorithmic work: bytes as above,
        # int8 operations 2 n n_b (3 C) for the genotype half and
        # 2 (n_b (n_b + 1) / 2) (4 x 3 C) for the U half (four U limbs, three
        # chain limbs)
        ops = None
        if precision == "mixed" and dom in ("xq", "xupd"):
            nb = a.blocksize
            ops = 2.0 * a.nind * nb * 3 * chains + 2.0 * (nb * (nb + 1) / 2) * 12 * chains
        traffic = TRAFFIC.get((precision, dom)) if (chains == (TRAFFIC_CHAINS if precision == "mixed" else 1)
                                                    and a.nind == 50000 and a.blocksize == 4096) else None
        if ops is not None and ops / (INT8_PEAK_TOPS * 1e12) > d["algorithmic_bytes"] / (HBM_PEAK_GBS * 1e9):
            tops = ops / (d["avg_us"] * 1e-6) / 1e12
            out["roofline"] = {
                "bound": "mfma", "achieved": round(tops, 1), "peak": INT8_PEAK_TOPS, "unit": "TOP/s (int8)",
                "frac": round(tops / INT8_PEAK_TOPS, 4), "traffic": traffic, "kernel": d["name"],
                "avg_launch_us": d["avg_us"], "algorithmic_ops_per_launch": int(ops),
                "algorithmic_bytes_per_launch": d["algorithmic_bytes"],
                "hbm_gbs": d["gbs"], "hbm_frac": round(d["gbs"] / HBM_PEAK_GBS, 4)}
        else:
            out["roofline"] = {
                "bound": "hbm", "achieved": d["gbs"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(d["gbs"] / HBM_PEAK_GBS, 4), "traffic": traffic,
                "kernel": d["name"], "avg_launch_us": d["avg_us"],
                "algorithmic_bytes_per_launch": d["algorithmic_bytes"]}
        for k, v in sorted(kernels.items(), key=lambda kv: -kv[1]["share"]):
            log(f"[{precision} x{chains}]   {v['name']:<34s} {v['avg_us']:8.2f} us/launch  {v['gbs']:7.0f} GB/s"
                f"  {v['share']:6.1%} of kernel time")
    if troof:
        # the operand layouts alone: one read-only pass per block over every
        # block (sb_layout_stream), the bandwidth the product kernels' loads
        # could reach with nothing else to do
        ls = s.layout_stream()
        out["layout_stream"] = {k: {"bytes_per_block": int(v[0]), "us_per_block": round(v[1], 2),
                                    "gbs": round(v[2], 1)} for k, v in ls.items()}
        log(f"[{precision} x{chains}] read-only layout passes: "
            + ", ".join(f"{k} {v[1]:.1f} us ({v[2]:.0f} GB/s)" for k, v in ls.items()))
    if bed is not None and not a.no_accuracy:
        out["accuracy"] = accuracy(torch, s, bed, freq, chains)
    s.close()
    return out


def run_folds(soda, torch, a) -> dict:
    """Cross-validation folds in one launch: BayesCpi, n = 40,000, p = 100,000,
    64 folds, each holding out a random 5-40% of the individuals (imputed
    every iteration, sb_mcmc_init_chains), on the mixed batched path.  The
    SODA subset size is chosen from {1024, 2048, 4096} by timing a few sweeps
    of each (soda.autotune_blocksize); then ``steps`` iterations are timed."""
    import numpy as np

    n, p, C = 40000, 100000, 64
    bed, freq, why = synth_on_device(torch, n, p, a.seed + 1, nqtl=1000, h2=0.3)
    rs = np.random.default_rng(a.seed)
    held = np.zeros((C, n), bool)
    for c in range(C):
        held[c, rs.choice(n, size=int(rs.uniform(0.05, 0.4) * n), replace=False)] = True
    opts = soda.SodaOptions(method="BayesCpi", seed=a.seed, rng="counter", precision="mixed", chains=C)
    kw = dict(freq=freq, bed_device_ptr=bed.data_ptr(), heldout=held)
    best, timings = soda.autotune_blocksize(opts, n, p, why, candidates=(1024, 2048, 4096), iters=3, warmup=2,
                                            **kw)
    log(f"[folds x{C}] subset size {best} (ms per iteration: "
        + ", ".join(f"{k}: {v:.1f}" for k, v in timings.items()) + ")")
    warm = 5
    o = soda.SodaOptions(method="BayesCpi", seed=a.seed, rng="counter", precision="mixed", chains=C,
                         blocksize=best, numit=warm + a.steps, burnin=0)
    s = soda.SodaSampler(o, n, p, why, **kw)
    s.setup()
    s.init()
    s.run(warm, trace=False)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    s.run(a.steps, trace=False)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    s.close()
    del bed
    out = {"workload": f"SODA BayesCpi n={n} p={p}, {C} cross-validation folds (5-40% of the individuals held "
                       "out and imputed each iteration), mixed batched path, counter-based RNG",
           "value": p * C * a.steps / dt, "unit": "SNP-effect updates/s", "ms_per_step": dt * 1e3 / a.steps,
           "chains": C, "blocksize": best, "blocksize_ms_per_step": timings}
    log(f"[folds x{C}] {out['ms_per_step']:.2f} ms/iteration, {out['value']:.4g} SNP-updates/s")
    return out


def run_traits(soda, torch, a) -> dict:
    """Batched traits in one launch: BayesRc, n = 100,000, p = 600,000, SNPs in
    4 annotation classes (proportions 0.6 / 0.2 / 0.15 / 0.05) with their own
    mixture proportions, 16 traits (traits_on_device), subsets of 4,096 SNPs,
    mixed batched path; ``steps`` iterations timed."""
    import numpy as np

    n, p, T = 100000, 600000, 16
    bed, freq, why = traits_on_device(torch, n, p, a.seed + 2, ntraits=T)
    grp = np.random.default_rng(a.seed).choice(4, size=p, p=[0.6, 0.2, 0.15, 0.05]) + 1
    warm = 5
    o = soda.SodaOptions(method="BayesRc", nmix=4, mix="0.0,0.0001,0.001,0.01", seed=a.seed, rng="counter",
                         precision="mixed", chains=T, blocksize=4096, numit=warm + a.steps, burnin=0)
    s = soda.SodaSampler(o, n, p, why, freq=freq, bed_device_ptr=bed.data_ptr(), grp=grp)
    t0 = time.time()
    s.setup()
    setup_s = time.time() - t0
    s.init()
    s.run(warm, trace=False)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    s.run(a.steps, trace=False)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    s.close()
    del bed
    out = {"workload": f"SODA BayesRc n={n} p={p} n_s=4096, 4 annotation classes, {T} traits (3,000 QTL each, "
                       "half shared, h2 ~ U(0.1, 0.6)), mixed batched path, counter-based RNG",
           "value": p * T * a.steps / dt, "unit": "SNP-effect updates/s", "ms_per_step": dt * 1e3 / a.steps,
           "chains": T, "setup_seconds": round(setup_s, 2)}
    log(f"[traits x{T}] setup {setup_s:.1f} s, {out['ms_per_step']:.2f} ms/iteration, "
        f"{out['value']:.4g} SNP-updates/s")
    return out


def run_c3(soda, torch, a, why, freq, bed_ptr) -> dict:
    """BayesA and BayesB on the headline genotypes (n = 50,000, p = 500,000,
    subsets of 4,096): -pi 0.99 -df_va 4 -df_ve 4, prior scales set from
    h2 = 0.4 of the phenotype's variance, mixed batched path, ``--chains``
    counter-based chains; a few warm-up iterations, then ``steps`` timed."""
    import numpy as np

    n, p, bs, C = a.nind, a.nloci, a.blocksize, a.chains
    vary = float(np.var(why))
    df = 4.0
    out = {"workload": f"SODA BayesA and BayesB n={n} p={p} n_s={bs}, -pi 0.99 -df_va 4 -df_ve 4 (scales from "
                       f"h2 = 0.4), {C} chains batched, mixed batched path, counter-based RNG",
           "unit": "SNP-effect updates/s"}
    for method in ("BayesA", "BayesB"):
        # the prior modes of the per-SNP and residual variances: h2 vary spread
        # over the SNPs (over the 1 - pi included ones for BayesB), (1 - h2) vary
        share = p * (1.0 if method == "BayesA" else 0.01)
        kw = dict(df_va=df, df_ve=df, scale_va=0.4 * vary / share * (df + 2.0) / df,
                  scale_ve=0.6 * vary * (df + 2.0) / df)
        if method == "BayesB":
            kw["pi"] = 0.99
        warm = 5
        o = soda.SodaOptions(method=method, blocksize=bs, numit=warm + a.steps, burnin=0, seed=a.seed,
                             rng="counter", precision="mixed", chains=C, **kw)
        s = soda.SodaSampler(o, n, p, why, freq=freq, bed_device_ptr=bed_ptr)
        s.setup()
        s.init()
        s.run(warm, trace=False)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        s.run(a.steps, trace=False)
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        s.close()
        out[method] = {"value": p * C * a.steps / dt, "ms_per_step": dt * 1e3 / a.steps, "chains": C}
        log(f"[c3 {method} x{C}] {out[method]['ms_per_step']:.2f} ms/iteration, "
            f"{out[method]['value']:.4g} SNP-updates/s")
    return out


def run_single_site(soda, torch, a) -> dict:
    """The single-site sampler (-SODAOff, regular_bayes) on the reference's
    example shape: BayesR, n = 2,326, p = 9,346 (examples/example2), 4-class
    mixture, synthetic genotypes of that shape.  REPLAY, one chain (the
    reference's -nthreads 1 stream), then 64 counter-based chains, one
    workgroup each; iterations timed after a short warm-up.  The chain is
    serial in SNPs, so this is a latency figure, not a bandwidth one."""
    n, p = 2326, 9346
    bed, freq, why = synth_on_device(torch, n, p, a.seed + 3, nqtl=200, h2=0.5)
    out = {"workload": f"single-site sampler (-SODAOff) BayesR n={n} p={p}, 4-class mixture, FP64",
           "unit": "SNP-effect updates/s"}
    for key, kw, its in (("replay", dict(rng="replay"), 20), ("counter_64_chains", dict(rng="counter", chains=64), 10)):
        warm = 3
        o = soda.SodaOptions(method="BayesR", nmix=4, mix="0.0,0.0001,0.001,0.01", seed=a.seed, sodaoff=True,
                             numit=warm + its, burnin=0, **kw)
        s = soda.SodaSampler(o, n, p, why, freq=freq, bed_device_ptr=bed.data_ptr())
        s.setup()
        s.init()
        s.run(warm, trace=False)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        s.run(its, trace=False)
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        s.close()
        C = kw.get("chains", 1)
        out[key] = {"value": p * C * its / dt, "ms_per_step": dt * 1e3 / its, "chains": C}
        log(f"[single-site {key}] {out[key]['ms_per_step']:.2f} ms/iteration, {out[key]['value']:.4g} SNP-updates/s")
    del bed
    return out


def main():
    a = parse_args()
    if a.cpu_only:
        print(json.dumps({"cpu_baseline": cpu_baseline(a.seed, a.cpu_nloci, a.cpu_blocksize, a.cpu_iters)}),
              flush=True)
        return
    if a.gpus != 1:
        print(f"bench.py runs on one GPU: this build has no multi-GPU path (asked for {a.gpus}); "
              "the sharding design is in DESIGN.md", file=sys.stderr)
        sys.exit(2)
    import torch

    from sodabayes_amd import soda

    n, p, bs = a.nind, a.nloci, a.blocksize
    if a.copy_probe:
        src = torch.ones(2 ** 27, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
        log("copy probe: 1 GiB read + 1 GiB written")
    t0 = time.time()
    bed, freq, why = synth_on_device(torch, n, p, a.seed)
    torch.cuda.synchronize()
    log(f"genotypes generated in HBM: n={n} p={p} ({bed.numel() / 1e9:.2f} GB) in {time.time() - t0:.1f} s")
    fp64 = None if a.chains_only else run_path(soda, torch, a, why, freq, bed.data_ptr(), "fp64", 1, bed=bed)
    if a.fp64_only:
        return
    mixed = run_path(soda, torch, a, why, freq, bed.data_ptr(), "mixed", a.chains, bed=bed)
    c3 = None if a.no_c3 else run_c3(soda, torch, a, why, freq, bed.data_ptr())
    del bed
    torch.cuda.empty_cache()
    folds = None if a.no_folds else run_folds(soda, torch, a)
    torch.cuda.empty_cache()
    traits = None if a.no_traits else run_traits(soda, torch, a)
    torch.cuda.empty_cache()
    single = None if a.no_single_site else run_single_site(soda, torch, a)
    torch.cuda.empty_cache()

    cpu = None if a.no_cpu_baseline else cpu_baseline(a.seed, a.cpu_nloci, a.cpu_blocksize, a.cpu_iters)
    out = {
        "metric": "SNP-effect updates/s (BayesR n=50k,p=500k): mixed+batched vs FP64 GPU, same acc.",
        "value": mixed["value"], "unit": "SNP-effect updates/s", "n_gpus": 1, "steps": a.steps,
        "warmup": a.warmup, "ms_per_step": mixed["ms"], "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "int8 MFMA limbs (int32 sums) + f64", "data": "synthetic",
        "config": {"workload": f"SODA BayesR n={n} p={p} n_s={bs}, 4-class mixture, {a.chains} chains "
                               "batched, counter-based RNG, mixed-precision products",
                   "nind": n, "nloci": p, "blocksize": bs, "chains": a.chains, "limbs": 3,
                   "parallelism": "single GPU"},
        "fp64_gpu": None if fp64 is None else {
            "value": fp64["value"], "ms_per_step": fp64["ms"], "chains": 1,
            "roofline": fp64.get("roofline"), "layout_stream": fp64.get("layout_stream"),
            "setup_seconds": round(fp64["setup_s"], 2),
            "accuracy": fp64.get("accuracy")},
        "speedup_vs_fp64_gpu": None if fp64 is None else mixed["value"] / fp64["value"],
        "setup_seconds": round(mixed["setup_s"], 2),
        "power_method_iterations": mixed["power_iters"],
        "iterations_per_s": a.steps / mixed["dt"],
        "tile_shape": mixed.get("tile_shape"),
        "accuracy": mixed.get("accuracy"),
        "layout_stream": mixed.get("layout_stream"),
        "folds": folds,
        "traits": traits,
        "bayesa_bayesb": c3,
        "single_site": single,
        "roofline": mixed.get("roofline"),
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()

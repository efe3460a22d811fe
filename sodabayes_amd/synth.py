"""Seeded synthetic genotype/phenotype generator for the SODA workloads.

Genotypes: allele frequency f_j ~ U(fmin, fmax), x_ij ~ Binomial(2, f_j).
With ``ld_block > 0`` SNPs in consecutive blocks share a latent haplotype:
each of the two alleles of x_ij copies the block's latent allele with
probability ``ld_copy`` and is drawn independently otherwise, which gives
r^2 of roughly ``ld_copy**2`` inside a block.
Phenotype: ``nqtl`` QTL at random loci with N(0,1) (or t(df)) effects on the
centred genotypes, scaled so that var(g) / var(y) = h2, plus N(0, 1-h2)
noise, y = g + e + 100.
"""
from __future__ import annotations

import numpy as np

from .plink import encode_bed


def genotypes(nind: int, nloci: int, seed: int, fmin: float = 0.05, fmax: float = 0.5,
              ld_block: int = 0, ld_copy: float = 0.55):
    rng = np.random.default_rng(seed)
    f = rng.uniform(fmin, fmax, size=nloci)
    if ld_block <= 0:
        x = rng.binomial(2, f[None, :], size=(nind, nloci)).astype(np.int8)
        return x, f
    x = np.empty((nind, nloci), dtype=np.int8)
    for b0 in range(0, nloci, ld_block):
        b1 = min(nloci, b0 + ld_block)
        fb = f[b0:b1]
        fl = float(fb.mean())
        lat = rng.random((nind, 2)) < fl
        cnt = np.zeros((nind, b1 - b0), dtype=np.int8)
        for a in range(2):
            copy = rng.random((nind, b1 - b0)) < ld_copy
            own = rng.random((nind, b1 - b0)) < fb[None, :]
            cnt += np.where(copy, lat[:, a:a + 1], own).astype(np.int8)
        x[:, b0:b1] = cnt
    return x, f


def phenotype(x: np.ndarray, seed: int, nqtl: int, h2: float, effect_df: float = 0.0):
    rng = np.random.default_rng(seed + 7919)
    nind, nloci = x.shape
    nqtl = min(nqtl, nloci)
    loci = np.sort(rng.choice(nloci, size=nqtl, replace=False))
    if effect_df > 0:
        eff = rng.standard_t(effect_df, size=nqtl)
    else:
        eff = rng.standard_normal(nqtl)
    xq = x[:, loci].astype(np.float64)
    xq -= xq.mean(axis=0, keepdims=True)
    g = xq @ eff
    vg = g.var()
    if vg > 0:
        g *= np.sqrt(h2 / vg)
    e = rng.standard_normal(nind) * np.sqrt(1.0 - h2)
    y = 100.0 + g + e
    # round-trip through the text form written to .fam so every reader sees
    # the same doubles
    y = np.array([float(repr(float(v))) for v in y])
    return y, loci, eff


def dataset(nind: int, nloci: int, seed: int, nqtl: int = 20, h2: float = 0.5,
            fmin: float = 0.05, fmax: float = 0.5, ld_block: int = 0,
            effect_df: float = 0.0):
    """Return (bed payload (nloci, ceil(n/4)) uint8, phenotype, counts)."""
    x, _ = genotypes(nind, nloci, seed, fmin, fmax, ld_block)
    y, _, _ = phenotype(x, seed, nqtl, h2, effect_df)
    return encode_bed(x), y, x


def snp_groups(nloci: int, seed: int, props=(0.6, 0.2, 0.15, 0.05)) -> np.ndarray:
    """Annotation class (1-based) per SNP with the given proportions."""
    rng = np.random.default_rng(seed + 104729)
    p = np.asarray(props, dtype=np.float64)
    p = p / p.sum()
    return (rng.choice(len(p), size=nloci, p=p) + 1).astype(np.int32)

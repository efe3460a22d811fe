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


# ---------------------------------------------------------------- device --
# The full-size workloads (p up to 600,000 SNPs) are generated in device
# memory with torch as plumbing: the genotypes never exist on the host.


def _device_bed_chunks(torch, n, p, seed, fmin, fmax, ld_block=0, ld_copy=0.55, chunk=2048):
    """SNP-major PLINK payload (p x ceil(n/4) uint8, on cuda:0) filled chunk
    by chunk; yields (j0, m, x, f) per chunk with x the (npad x m) uint8
    counts (rows >= n zero) and f the chunk's allele frequencies.  With
    ld_block > 0, consecutive blocks of ld_block SNPs share a latent
    haplotype: each allele copies the block's latent allele with
    probability ld_copy, else it is drawn from f_j (r^2 about ld_copy^2)."""
    dev = torch.device("cuda:0")
    g = torch.Generator(device=dev)
    g.manual_seed(seed)
    f = torch.rand(p, generator=g, device=dev, dtype=torch.float64) * (fmax - fmin) + fmin
    nb = (n + 3) // 4
    npad = nb * 4
    bed = torch.empty((p, nb), dtype=torch.uint8, device=dev)
    counts = torch.empty(p, dtype=torch.float64, device=dev)
    if ld_block > 0:
        chunk = max(ld_block, (chunk // ld_block) * ld_block)

    def gen():
        for j0 in range(0, p, chunk):
            m = min(chunk, p - j0)
            fc = f[j0:j0 + m].float()
            if ld_block <= 0:
                u = torch.rand((2, npad, m), generator=g, device=dev)
                x = (u[0] < fc).to(torch.uint8) + (u[1] < fc).to(torch.uint8)
                del u
            else:
                blk = torch.arange(m, device=dev) // ld_block
                nblk = int(blk[-1]) + 1
                fl = torch.zeros(nblk, device=dev).index_add_(0, blk, fc) / torch.bincount(blk, minlength=nblk)
                lat = torch.rand((2, npad, nblk), generator=g, device=dev) < fl
                x = torch.zeros((npad, m), dtype=torch.uint8, device=dev)
                for a in range(2):
                    copy = torch.rand((npad, m), generator=g, device=dev) < ld_copy
                    own = torch.rand((npad, m), generator=g, device=dev) < fc
                    x += torch.where(copy, lat[a][:, blk], own).to(torch.uint8)
                del lat
            if npad > n:
                x[n:] = 0
            counts[j0:j0 + m] = x.sum(0, dtype=torch.int64).to(torch.float64)
            code = x + (x > 0).to(torch.uint8)          # 0 -> 00, 1 -> 10, 2 -> 11
            c4 = code.view(nb, 4, m)
            bed[j0:j0 + m] = (c4[:, 0] | (c4[:, 1] << 2) | (c4[:, 2] << 4) | (c4[:, 3] << 6)).t()
            del code, c4
            yield j0, m, x, f
    return bed, counts, gen


def device_dataset(torch, n, p, seed, fmin=0.01, fmax=0.5, nqtl=5000, h2=0.4, effect_df=4.0, ld_block=0,
                   chunk=2048):
    """One phenotype over device genotypes: nqtl QTL with t(effect_df)
    effects (N(0, 1) for effect_df <= 0) on the centred genotypes, scaled to
    h2, plus N(0, 1 - h2) noise, y = 100 + g + e.  Returns (bed on cuda:0,
    allele frequencies, y, QTL ids)."""
    import numpy as np

    dev = torch.device("cuda:0")
    rs = np.random.default_rng(seed)
    qtl = np.sort(rs.choice(p, size=min(nqtl, p), replace=False))
    eff = rs.standard_t(effect_df, size=len(qtl)) if effect_df > 0 else rs.standard_normal(len(qtl))
    qtl_t = torch.from_numpy(qtl).to(dev)
    eff_t = torch.from_numpy(eff).to(dev)
    gval = torch.zeros(n, dtype=torch.float64, device=dev)
    bed, counts, gen = _device_bed_chunks(torch, n, p, seed, fmin, fmax, ld_block=ld_block, chunk=chunk)
    for j0, m, x, f in gen():
        sel = (qtl_t >= j0) & (qtl_t < j0 + m)
        if bool(sel.any()):
            cols = qtl_t[sel] - j0
            gval += (x[:n, cols].to(torch.float64) - 2.0 * f[j0 + cols]) @ eff_t[sel]
        del x
    freq = (counts / (2.0 * n)).cpu().numpy()
    gv = gval.cpu().numpy()
    gv *= np.sqrt(h2 / gv.var())
    y = 100.0 + gv + rs.standard_normal(n) * np.sqrt(1.0 - h2)
    return bed, freq, y, qtl


def device_traits(torch, n, p, seed, ntraits=16, nqtl=3000, shared=0.5, chunk=2048):
    """ntraits phenotypes over device genotypes (f ~ U(0.01, 0.5)): each
    trait has nqtl QTL with N(0, 1) effects, a fraction `shared` of them
    common to all traits with the same effect (genetic correlation about
    `shared`), h2 drawn from U(0.1, 0.6).  Returns (bed on cuda:0, allele
    frequencies, y as (ntraits, n))."""
    import numpy as np

    dev = torch.device("cuda:0")
    rs = np.random.default_rng(seed)
    ns = int(nqtl * shared)
    common = rs.choice(p, size=ns, replace=False)
    eff = np.zeros((p, ntraits))
    eff[common] = rs.standard_normal((ns, 1)) * np.ones((1, ntraits))   # shared QTL, same effect
    for t in range(ntraits):
        own = rs.choice(np.setdiff1d(np.arange(p), common), size=nqtl - ns, replace=False)
        eff[own, t] = rs.standard_normal(nqtl - ns)
    qtl = np.flatnonzero(np.abs(eff).sum(1) > 0)
    qtl_t = torch.from_numpy(qtl).to(dev)
    eff_t = torch.from_numpy(eff[qtl]).to(dev)
    gval = torch.zeros((n, ntraits), dtype=torch.float64, device=dev)
    bed, counts, gen = _device_bed_chunks(torch, n, p, seed, 0.01, 0.5, chunk=chunk)
    for j0, m, x, f in gen():
        sel = (qtl_t >= j0) & (qtl_t < j0 + m)
        if bool(sel.any()):
            cols = qtl_t[sel] - j0
            gval += (x[:n, cols].to(torch.float64) - 2.0 * f[j0 + cols]) @ eff_t[sel]
        del x
    freq = (counts / (2.0 * n)).cpu().numpy()
    gv = gval.cpu().numpy()
    h2 = rs.uniform(0.1, 0.6, size=ntraits)
    gv *= np.sqrt(h2 / gv.var(0))
    y = 100.0 + gv + rs.standard_normal((n, ntraits)) * np.sqrt(1.0 - h2)
    return bed, freq, np.ascontiguousarray(y.T)


def device_standardized(torch, bed, n, freq, cols):
    """z_ij = (x_ij - 2 f_j) / sqrt(2 f_j (1 - f_j)) of SNP ids `cols` as an
    (len(cols) x n) float64 device tensor, from the device .bed payload
    (missing genotypes read as 1, z = 0 for f_j or 1 - f_j below machine
    epsilon), as the reference's xscale standardises (utilities.f90:519-539)."""
    import numpy as np

    dev = bed.device
    ci = torch.as_tensor(np.asarray(cols, dtype=np.int64), device=dev)
    by = bed[ci].to(torch.int32)                               # (m, nbytes)
    sh = torch.tensor([0, 2, 4, 6], device=dev, dtype=torch.int32)
    code = ((by[:, :, None] >> sh) & 3).reshape(len(cols), -1)[:, :n]
    x = code - ((code >> 1) & 1)                                # 00->0, 01(missing)->1, 10->1, 11->2
    f = torch.as_tensor(np.asarray(freq, dtype=np.float64)[np.asarray(cols)], device=dev)
    eps = np.finfo(np.float64).eps
    ok = (f >= eps) & ((1.0 - f) >= eps)
    s = torch.where(ok, torch.sqrt(2.0 * f * (1.0 - f)), torch.ones_like(f))
    z = (x.to(torch.float64) - (2.0 * f)[:, None]) / s[:, None]
    return torch.where(ok[:, None], z, torch.zeros_like(z))

"""Accuracy checks of a sampler's state against FP64 evaluations made
independently of the product path under test (torch on the device, as
plumbing): the block products X_s'r + U_s'r_a (models.f90:358) and the
residual identity r = y - Z beta - mu (compute_residuals,
utilities.f90:609-614).  Used by the full-size tests and by bench.py's
accuracy report; the standardisation follows xscale (utilities.f90:519-539).
"""
from __future__ import annotations

import numpy as np

from . import soda, synth


def block_cols(s, b):
    """SNP ids of block b (block-position order)."""
    a, e = soda.block_bounds(s.nloci, s.blocksize, s.nblocks, s.blockt)[b]
    return s.permvec[a:e]


def fp64_products(torch, s, bed, freq, b, chains):
    """X_s'r + U_s'r_a of block b for the given chains from the sampler's
    downloaded state, FP64 in torch: (len(chains), n_b)."""
    cols = block_cols(s, b)
    nb = len(cols)
    Z = synth.device_standardized(torch, bed, s.nind, freq, cols)            # nb x n
    R = torch.tensor(np.stack([s.get("YADJ", s.nind, c) for c in chains]), device=Z.device)
    RA = torch.tensor(np.stack([s.get("YAADJ", s.blocksize, c)[:nb] for c in chains]), device=Z.device)
    XA = torch.tensor(s.get("XA_BLOCK", nb * nb, b).reshape(nb, nb, order="F"), device=Z.device)
    out = Z @ R.T + XA.T @ RA.T                                              # nb x C
    return out.T.cpu().numpy()


def products_error(torch, s, bed, freq, blocks, chains) -> float:
    """Worst norm-wise relative error of the sampler's own products
    (sb_block_rhs) against fp64_products over the blocks and chains."""
    worst = 0.0
    for b in blocks:
        want = fp64_products(torch, s, bed, freq, b, chains)
        got = s.block_rhs(b)
        for k, c in enumerate(chains):
            worst = max(worst, float(np.linalg.norm(got[c] - want[k]) / np.linalg.norm(want[k])))
    return worst


def residual_identity(torch, s, bed, freq, chains, rows=None, snp_chunk=8192) -> float:
    """max over chains of ||r - (y - Z beta - mu)|| / ||y - Z beta - mu||,
    over all individuals or the given rows; Z beta in FP64 from the .bed
    bytes, SNP chunk by chunk."""
    n, p = s.nind, s.nloci
    dev = bed.device
    idx = np.arange(n) if rows is None else np.sort(np.asarray(rows))
    it = torch.as_tensor(idx, device=dev)
    byte, sh = it // 4, (2 * (it % 4)).to(torch.int32)
    f = torch.as_tensor(np.asarray(freq, dtype=np.float64), device=dev)
    eps = np.finfo(np.float64).eps
    B = torch.tensor(np.stack([s.get("G", p, c) for c in chains]), device=dev)          # C x p
    zb = torch.zeros((len(chains), len(idx)), dtype=torch.float64, device=dev)
    for j0 in range(0, p, snp_chunk):
        j1 = min(p, j0 + snp_chunk)
        code = (bed[j0:j1][:, byte].to(torch.int32) >> sh) & 3                          # m x rows
        x = (code - ((code >> 1) & 1)).to(torch.float64)
        fc = f[j0:j1]
        ok = (fc >= eps) & ((1.0 - fc) >= eps)
        sd = torch.where(ok, torch.sqrt(2.0 * fc * (1.0 - fc)), torch.ones_like(fc))
        z = torch.where(ok[:, None], (x - (2.0 * fc)[:, None]) / sd[:, None], torch.zeros_like(x))
        zb += B[:, j0:j1] @ z
    worst = 0.0
    for k, c in enumerate(chains):
        y = s.get("WHY", n, c)[idx]
        mu = s.get("SCALARS", 4, c)[0]
        r = s.get("YADJ", n, c)[idx]
        want = y - zb[k].cpu().numpy() - mu
        worst = max(worst, float(np.linalg.norm(r - want) / np.linalg.norm(want)))
    return worst


def sample_blocks(nblocks: int, k: int = 8) -> list[int]:
    """k blocks spread evenly over the partition (first and last included)."""
    return sorted(set(np.linspace(0, nblocks - 1, k).round().astype(int).tolist()))

"""PLINK binary genotype files (.bed/.bim/.fam) and the packed-count layout.

The reference reads SNP-major PLINK files (magic ``6c 1b 01``) and maps each
2-bit code ``c`` to an allele count with ``c - ((c >> 1) & 1)``: 00->0, 01
(missing)->1, 10->1, 11->2 (utilities.f90:446-451 for the dense load,
utilities.f90:489 for the packed load).  Phenotypes come from column 6 of the
.fam file (utilities.f90:342).
"""
from __future__ import annotations

import os
import numpy as np

BED_MAGIC = bytes([0x6C, 0x1B, 0x01])

# allele count -> PLINK 2-bit code (0 -> 00, 1 -> 10, 2 -> 11)
_COUNT_TO_CODE = np.array([0b00, 0b10, 0b11], dtype=np.uint8)
# PLINK 2-bit code -> allele count, missing (01) read as 1
_CODE_TO_COUNT = np.array([0, 1, 1, 2], dtype=np.int8)


def bytes_per_snp(nind: int) -> int:
    return (nind + 3) // 4


def encode_bed(counts: np.ndarray, missing: np.ndarray | None = None) -> np.ndarray:
    """Encode an (nind, nloci) int8 count matrix as SNP-major .bed payload.

    Returns a (nloci, ceil(nind/4)) uint8 array (without the 3 magic bytes).
    ``missing`` (bool, same shape) writes code 01 for those genotypes.
    """
    counts = np.asarray(counts)
    nind, nloci = counts.shape
    codes = _COUNT_TO_CODE[counts.T.astype(np.intp)]
    if missing is not None:
        codes = np.where(np.asarray(missing).T, np.uint8(0b01), codes)
    nb = bytes_per_snp(nind)
    pad = nb * 4 - nind
    if pad:
        codes = np.concatenate([codes, np.zeros((nloci, pad), np.uint8)], axis=1)
    codes = codes.reshape(nloci, nb, 4)
    out = codes[:, :, 0] | (codes[:, :, 1] << 2) | (codes[:, :, 2] << 4) | (codes[:, :, 3] << 6)
    return np.ascontiguousarray(out.astype(np.uint8))


def decode_bed(bed: np.ndarray, nind: int) -> np.ndarray:
    """Inverse of :func:`encode_bed`: (nloci, nbytes) uint8 -> (nind, nloci) int8 counts."""
    bed = np.asarray(bed, dtype=np.uint8)
    nloci, nb = bed.shape
    codes = np.stack([(bed >> s) & 3 for s in (0, 2, 4, 6)], axis=2).reshape(nloci, nb * 4)
    return np.ascontiguousarray(_CODE_TO_COUNT[codes[:, :nind]].T)


def write_plink(prefix: str, bed: np.ndarray, why: np.ndarray, snp_ids=None,
                ind_ids=None) -> None:
    """Write prefix.bed/.bim/.fam.  ``bed`` is the SNP-major payload."""
    bed = np.asarray(bed, dtype=np.uint8)
    nloci = bed.shape[0]
    nind = len(why)
    d = os.path.dirname(prefix)
    if d:
        os.makedirs(d, exist_ok=True)
    with open(prefix + ".bed", "wb") as f:
        f.write(BED_MAGIC)
        f.write(bed.tobytes())
    if snp_ids is None:
        snp_ids = [f"snp{j + 1}" for j in range(nloci)]
    with open(prefix + ".bim", "w") as f:
        for j, s in enumerate(snp_ids):
            f.write(f"1\t{s}\t0\t{j + 1}\t1\t2\n")
    if ind_ids is None:
        ind_ids = [str(i + 1) for i in range(nind)]
    with open(prefix + ".fam", "w") as f:
        for i, s in enumerate(ind_ids):
            f.write(f"{s} {s} 0 0 0 {float(why[i])!r}\n")


def write_groups(path: str, grp: np.ndarray, snp_ids=None) -> None:
    """SNP-group file for BayesRc: ``snp_id group`` per line (utilities.f90:304-327)."""
    n = len(grp)
    if snp_ids is None:
        snp_ids = [f"snp{j + 1}" for j in range(n)]
    with open(path, "w") as f:
        for s, g in zip(snp_ids, grp):
            f.write(f"{s} {int(g)}\n")


def read_bed(prefix: str, nind: int, nloci: int) -> np.ndarray:
    """Read prefix.bed and return the (nloci, ceil(nind/4)) payload."""
    raw = np.fromfile(prefix + ".bed", dtype=np.uint8)
    if raw.size < 3 or bytes(raw[:3]) != BED_MAGIC:
        raise ValueError(f"{prefix}.bed is not a SNP-major PLINK file")
    nb = bytes_per_snp(nind)
    body = raw[3:]
    if body.size != nb * nloci:
        raise ValueError(f"{prefix}.bed holds {body.size} bytes, expected {nb * nloci}")
    return body.reshape(nloci, nb)


def read_fam_pheno(prefix: str) -> tuple[list[str], np.ndarray]:
    ids, y = [], []
    with open(prefix + ".fam") as f:
        for line in f:
            t = line.split()
            if not t:
                continue
            ids.append(t[1])
            y.append(float(t[5]))
    return ids, np.array(y, dtype=np.float64)


def count_lines(path: str) -> int:
    with open(path) as f:
        return sum(1 for line in f if line.strip())


def read_plink(prefix: str):
    """Return (bed payload, why, nind, nloci) for a SNP-major PLINK prefix."""
    ids, why = read_fam_pheno(prefix)
    nind = len(ids)
    nloci = count_lines(prefix + ".bim")
    return read_bed(prefix, nind, nloci), why, nind, nloci


def allele_freq(bed: np.ndarray, nind: int) -> np.ndarray:
    """Allele frequency p_j = sum_i x_ij / (2 n) (utilities.f90:527), exact in FP64."""
    counts = decode_bed(bed, nind)
    s = counts.sum(axis=0, dtype=np.int64).astype(np.float64)
    return s / (2.0 * nind)

"""Genomic values of the prediction run on the GPU (``-predict``).

The reference's prediction run (SODABayes.f90:145-148) reads the training
run's allele frequencies (.frq), posterior mean effects (.eff) and mean (.mod)
(setup_prediction, utilities.f90:636-660, and xscale's prediction branch,
utilities.f90:540-553), standardises the new genotypes with those
frequencies and writes ``pred_i = mu + sum_j z_ij g_j`` (compute_dgv,
utilities.f90:616-622) to the .gv file (write_dgv, utilities.f90:625-633).
``predict`` runs that product through the library's ``sb_predict``; the
readers and the writer keep the reference's file formats.
"""
from __future__ import annotations

import ctypes as C
import math

import numpy as np

from . import _lib
from .plink import read_bed, count_lines


def _second_column(path: str, skip: int = 0, count: int | None = None) -> np.ndarray:
    """List-directed read of ``dum, value`` per line (D exponents accepted)."""
    vals = []
    with open(path) as f:
        for k, line in enumerate(f):
            if k < skip or not line.strip():
                continue
            vals.append(float(line.split()[1].replace("D", "E").replace("d", "e")))
            if count is not None and len(vals) == count:
                break
    return np.array(vals, dtype=np.float64)


def read_frq(path: str, nloci: int | None = None) -> np.ndarray:
    """Allele frequencies, one ``id freq`` line per SNP (utilities.f90:541-544)."""
    return _second_column(path, 0, nloci)


def read_eff(path: str, nloci: int | None = None) -> np.ndarray:
    """Posterior mean effects after one header line (utilities.f90:647-654)."""
    return _second_column(path, 1, nloci)


def read_mu(path: str) -> float:
    """Mean from the first line of the .mod file (utilities.f90:656-658)."""
    return float(_second_column(path, 0, 1)[0])


def predict(bed, nind: int, freq: np.ndarray, effects: np.ndarray, mu: float) -> np.ndarray:
    """Genomic values of ``nind`` individuals.

    ``bed`` is the SNP-major .bed payload without magic bytes, shaped
    (nloci, ceil(nind/4)) uint8: a numpy array (uploaded by the call) or a
    contiguous device tensor (read in place).  Returns ``nind`` doubles.
    """
    freq = np.ascontiguousarray(freq, dtype=np.float64)
    effects = np.ascontiguousarray(effects, dtype=np.float64)
    nloci = freq.size
    if effects.size != nloci:
        raise ValueError("freq and effects differ in length")
    nbytes = (nind + 3) // 4
    if hasattr(bed, "data_ptr"):
        if not bed.is_cuda or not bed.is_contiguous() or bed.numel() != nloci * nbytes:
            raise ValueError("device bed must be a contiguous cuda tensor of nloci*ceil(nind/4) bytes")
        ptr, on_dev, keep = C.c_void_p(bed.data_ptr()), 1, bed
    else:
        keep = np.ascontiguousarray(bed, dtype=np.uint8)
        if keep.size != nloci * nbytes:
            raise ValueError("bed must hold nloci*ceil(nind/4) bytes")
        ptr, on_dev = keep.ctypes.data_as(C.c_void_p), 0
    pred = np.empty(nind, dtype=np.float64)
    dp = C.POINTER(C.c_double)
    _lib.check(_lib.load().sb_predict(int(nind), int(nloci), ptr, on_dev, freq.ctypes.data_as(dp),
                                     effects.ctypes.data_as(dp), float(mu), pred.ctypes.data_as(dp)))
    del keep
    return pred


def format_e15_7(v: float) -> str:
    """Fortran ``E15.7``: ``0.ddddddd`` mantissa, signed two-digit exponent."""
    if v == 0.0 or not math.isfinite(v):
        return f"{'0.0000000E+00' if v == 0.0 else str(v):>15}"
    m, e = f"{abs(v):.6e}".split("e")   # d.dddddd -> 0.ddddddd, exponent + 1
    s = f"{'-' if v < 0 else ''}0.{m.replace('.', '')}E{int(e) + 1:+03d}"
    return f"{s:>15}"


def write_gv(path: str, ids, pred: np.ndarray) -> None:
    """One ``A20,1x,E15.7`` line per individual (write_dgv)."""
    with open(path, "w") as f:
        for i, v in zip(ids, pred):
            f.write(f"{str(i):<20.20} {format_e15_7(float(v))}\n")


def predict_files(bfile: str, model: str, out: str) -> np.ndarray:
    """The prediction run: ``bfile``.{bed,bim,fam} scored with ``model``.{frq,eff,mod}."""
    nind = count_lines(bfile + ".fam")
    nloci = count_lines(bfile + ".bim")
    bed = read_bed(bfile, nind, nloci)
    pred = predict(bed, nind, read_frq(model + ".frq", nloci), read_eff(model + ".eff", nloci),
                   read_mu(model + ".mod"))
    with open(bfile + ".fam") as f:
        ids = [line.split()[1] for line in f if line.strip()]
    write_gv(out, ids, pred)
    return pred

"""ctypes binding of libsodabayes.so (the C ABI in include/sodabayes.h).

The library is the only compute path: if it is missing or cannot be loaded
this module raises instead of falling back to anything else.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libsodabayes.so")

SB_OK = 0
SB_ERR_ARG = -1
SB_ERR_NOT_PD = -5
METHODS = {"BayesA": 1, "BayesB": 2, "BayesC": 3, "BayesCpi": 4, "BayesR": 5, "BayesRc": 6}
RNG_REPLAY, RNG_COUNTER = 0, 1
PREC_FP64, PREC_MIXED = 0, 1

Q = dict(G=1, CLASS=2, VARA_S=3, YADJ=4, YAADJ=5, P=6, SCALARS=7, GSTORE=8, INDISTSTORE=9,
         VARA_SSTORE=10, STORES=11, SNPSTORE=12, VARSTORE=13, PSTORE=14, XA=15, XXD=16, GRAM=17,
         SNPINDIST=18, VARINDIST=19, KTIMES=20, SKIPPED=21, NAIND=22, WHY=23, XA_BLOCK=24)

# kernel timing slots of the KTIMES query (order of SB_KT_* in the header)
KT_SLOTS = ["rhs", "draw", "list", "update", "rquant", "xq", "ut0", "drawb", "dquant", "xupd", "ut1", "rared"]

# every function include/sodabayes.h declares
EXPORTS = ["sb_rng_seed", "sb_permutate", "sb_create", "sb_destroy", "sb_rng_set", "sb_rng_get",
           "sb_build_aug", "sb_test_shift_d", "sb_mcmc_init", "sb_mcmc_init_chains", "sb_mcmc_run", "sb_hyp_width", "sb_get",
           "sb_block_rhs", "sb_get_info", "sb_predict", "sb_predict_noalloc", "sb_state_size", "sb_state_save", "sb_state_load",
           "sb_counter_uniforms", "sb_layout_stream", "sb_last_error"]


class SbDims(C.Structure):
    _fields_ = [("method", C.c_int32), ("nind", C.c_int32), ("nloci", C.c_int32),
                ("nmix", C.c_int32), ("ngroup", C.c_int32), ("nblocks", C.c_int32),
                ("blocksize", C.c_int32), ("blockt", C.c_int32), ("nchains", C.c_int32),
                ("rng_mode", C.c_int32), ("precision", C.c_int32), ("bed_on_device", C.c_int32)]


class SbPriors(C.Structure):
    _fields_ = [("pizero", C.c_double), ("scale_va", C.c_double), ("scale_ve", C.c_double),
                ("df_va", C.c_double), ("df_ve", C.c_double), ("vara0", C.c_double),
                ("mix", C.c_double * 20), ("alpha", C.c_double * 20),
                ("numit", C.c_int32), ("burnin", C.c_int32), ("thin", C.c_int32),
                ("resync", C.c_int32), ("seed", C.c_int64)]


class SbError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"libsodabayes error {code}: {msg}")
        self.code = code


_lib = None


def load():
    """Load libsodabayes.so (raises if it is not built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is not built: run `python -m sodabayes_amd.build`")
    L = C.CDLL(LIB_PATH)
    vp, u64p, i32p, dp = C.c_void_p, C.POINTER(C.c_uint64), C.POINTER(C.c_int32), C.POINTER(C.c_double)
    L.sb_rng_seed.argtypes = [C.c_int64, C.c_int32, u64p]
    L.sb_permutate.argtypes = [C.c_int32, C.c_int32, u64p, i32p]
    L.sb_create.argtypes = [C.POINTER(SbDims), vp, dp, i32p, i32p, C.POINTER(vp)]
    L.sb_destroy.argtypes = [vp]
    L.sb_rng_set.argtypes = [vp, u64p]
    L.sb_rng_get.argtypes = [vp, u64p]
    L.sb_build_aug.argtypes = [vp, dp, i32p, i32p]
    L.sb_test_shift_d.argtypes = [vp, C.c_int32, C.c_double]
    L.sb_mcmc_init.argtypes = [vp, C.POINTER(SbPriors), dp]
    L.sb_mcmc_init_chains.argtypes = [vp, C.POINTER(SbPriors), dp, C.c_int64, C.POINTER(C.c_uint8), C.c_int64]
    L.sb_mcmc_run.argtypes = [vp, C.c_int32, C.c_int32, dp]
    L.sb_hyp_width.argtypes = [vp]
    L.sb_hyp_width.restype = C.c_int32
    L.sb_get.argtypes = [vp, C.c_int32, C.c_int32, dp, C.c_int64]
    L.sb_get_info.argtypes = [vp, dp, C.c_int32]
    L.sb_block_rhs.argtypes = [vp, C.c_int32, dp]
    L.sb_predict.argtypes = [C.c_int32, C.c_int32, vp, C.c_int32, dp, dp, C.c_double, dp]
    L.sb_predict_noalloc.argtypes = [C.c_int32, C.c_int32, vp, vp, vp, C.c_double, vp, vp]
    L.sb_state_size.argtypes = [vp, C.POINTER(C.c_int64)]
    L.sb_state_save.argtypes = [vp, vp, C.c_int64]
    L.sb_state_load.argtypes = [vp, vp, C.c_int64]
    L.sb_layout_stream.argtypes = [vp, dp]
    L.sb_counter_uniforms.argtypes = [C.c_int64, C.c_int32, C.POINTER(C.c_uint32), C.c_int32, C.c_int32, dp]
    L.sb_last_error.restype = C.c_char_p
    for f in EXPORTS:
        if f not in ("sb_hyp_width", "sb_last_error"):
            getattr(L, f).restype = C.c_int
    _lib = L
    return L


def check(rc: int) -> None:
    if rc != SB_OK:
        raise SbError(rc, load().sb_last_error().decode(errors="replace"))

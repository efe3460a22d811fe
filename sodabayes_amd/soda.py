"""Host side of the SODA branch, above the C ABI (include/sodabayes.h).

Mirrors what run_program_sodaBayes does around its two SODA calls
(SODABayes.f90:123-142): option values as the reference parses them, allele
frequencies, the seeded stream, the block partition (soda.f90:407-417) and
permutation (soda.f90:429-435), then setup_soda -> :meth:`SodaSampler.setup`
and soda_bayes -> :meth:`SodaSampler.init` / :meth:`SodaSampler.run`.
Results come back under the reference's variable names.
"""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass, field

import numpy as np

from . import _lib
from .plink import allele_freq

MAXDIST = 20   # globals.f90:8
MAXGRP = 100   # globals.f90:34


def f32(x: float) -> float:
    """cast_float (utilities.f90:168-171) reads options into a default real."""
    return float(np.float32(x))


@dataclass
class SodaOptions:
    """The MCMC options of useroptions.f90:24-54 with the reference defaults.

    Float options are rounded through float32 as cast_float does; ``mix`` and
    ``alpha`` accept the reference's comma-separated strings.
    """
    method: str = "BayesCpi"
    pi: float = 0.99
    scale_va: float = 0.01
    scale_ve: float = 0.01
    df_va: float = 0.01
    df_ve: float = 0.01
    alpha: object = 1.0
    numit: int = 50000
    burnin: int = 5000
    thin: int = 1
    nmix: int = 2
    mix: object = "0.0,1.0"
    seed: int = 0
    nthreads: int = 1
    blocksize: int = 0
    nblocks: int = 0
    rblocks: bool = False
    sodaoff: bool = False         # -SODAOff: the single-site sampler (models.f90:607-962)
    vara0: float = 0.0            # module vara before the job (models.f90:99-101)
    # build-side choices (no reference counterpart)
    rng: str = "replay"           # "replay" (xoroshiro stream 0) or "counter" (Philox)
    precision: str = "fp64"       # "fp64" or "mixed" (int8-limb MFMA products, chains batched)
    chains: int = 1               # independent chains sharing the genotypes (mixed only)
    resync: int = 0               # iterations between residual resyncs; 0: the reference's 1000

    def method_code(self) -> int:
        for k, v in _lib.METHODS.items():
            if k.upper() == str(self.method).upper():
                return v
        raise ValueError("supported methods: BayesA, BayesB, BayesC, BayesCpi, BayesR, and BayesRc")

    def resolved_nmix(self) -> int:
        """parse_nmix, useroptions.f90:375-400"""
        m = self.method_code()
        if self.nmix > MAXDIST:
            raise ValueError(f"nmix is too big. maximum allowed: {MAXDIST}")
        if m == 1:
            return 1
        if m in (2, 3, 4):
            return 2
        return int(self.nmix)

    def mix_values(self) -> list[float]:
        """parse_priors' -mix handling, useroptions.f90:501-518"""
        m, nmix = self.method_code(), self.resolved_nmix()
        if m in (1, 2):
            return [0.0] * nmix
        if m in (3, 4):
            return [0.0, 1.0]
        vals = self.mix.split(",") if isinstance(self.mix, str) else list(self.mix)
        if len(vals) != nmix:
            raise ValueError(f"Number of mixture classes {nmix} but {len(vals)} effect sizes specified (mix)")
        return [f32(float(v)) for v in vals]

    def alpha_values(self) -> list[float]:
        """parse_priors' -alpha handling, useroptions.f90:519-535"""
        nmix = self.resolved_nmix()
        vals = self.alpha.split(",") if isinstance(self.alpha, str) else (
            [self.alpha] if np.isscalar(self.alpha) else list(self.alpha))
        if len(vals) == 1:
            return [f32(float(vals[0]))] * nmix
        if len(vals) != nmix:
            raise ValueError(f"Number of mixture classes {nmix} but {len(vals)} Dirichlet priors (alpha)")
        return [f32(float(v)) for v in vals]


def partition(nloci: int, blocksize: int = 0, nblocks: int = 0, nthreads: int = 1):
    """Block partition of setup_soda, soda.f90:407-421.

    Returns (blocksize, nblocks, blockt); blocksize <= 1 means SODA is off.
    """
    if blocksize > 0:
        nb = math.ceil(nloci / blocksize)
        return blocksize, nb, nb
    if nblocks > 0:
        bs = math.ceil(nloci / nblocks)
        return bs, nblocks, nloci - (bs - 1) * nblocks
    bs = nthreads
    nb = math.ceil(nloci / bs)
    return bs, nb, nb


def block_bounds(nloci: int, blocksize: int, nblocks: int, blockt: int):
    """0-based [start, end) of each block (soda.f90:463-469)."""
    out = []
    for b in range(1, nblocks + 1):
        if b <= blockt:
            ls = (b - 1) * blocksize + 1
            le = min(ls + blocksize - 1, nloci)
        else:
            ls = blocksize * blockt + (b - blockt - 1) * (blocksize - 1) + 1
            le = min(ls + blocksize - 2, nloci)
        out.append((ls - 1, le))
    return out


def group_index(labels) -> tuple[np.ndarray, int]:
    """SNP groups numbered by first appearance (get_size / load_snp_grp,
    utilities.f90:276-327).  Returns 0-based indices and the group count."""
    seen: dict[str, int] = {}
    idx = np.empty(len(labels), dtype=np.int32)
    for j, lab in enumerate(labels):
        key = str(lab)
        if key not in seen:
            if len(seen) >= MAXGRP:
                raise ValueError(f"max no. of SNP groups: {MAXGRP}")
            seen[key] = len(seen)
        idx[j] = seen[key]
    return idx, len(seen)


def rng_seed(seed: int, stream: int = 0) -> np.ndarray:
    st = np.zeros(2, dtype=np.uint64)
    _lib.check(_lib.load().sb_rng_seed(int(seed), stream, st.ctypes.data_as(C.POINTER(C.c_uint64))))
    return st


def permutate(n: int, s: int, state: np.ndarray) -> np.ndarray:
    p = np.zeros(n, dtype=np.int32)
    _lib.check(_lib.load().sb_permutate(n, s, state.ctypes.data_as(C.POINTER(C.c_uint64)),
                                        p.ctypes.data_as(C.POINTER(C.c_int32))))
    return p


@dataclass
class SodaSampler:
    """One SODA job on the GPU.

    ``bed`` is the SNP-major PLINK payload (nloci x ceil(nind/4)); ``grp`` the
    SNP group labels for BayesRc.  ``bed_device_ptr`` (optional) passes a
    device pointer to the same payload instead of host memory.

    Batched traits and folds: ``why`` may be (chains, nind), one phenotype
    per chain, and ``heldout`` (nind) or (chains, nind) flags validation
    individuals whose phenotypes a chain does not use (imputed every
    iteration; sb_mcmc_init_chains).
    """
    opts: SodaOptions
    nind: int
    nloci: int
    why: np.ndarray
    bed: np.ndarray | None = None
    grp: object = None
    freq: np.ndarray | None = None
    bed_device_ptr: int | None = None
    heldout: np.ndarray | None = None
    # tests only: (block, amount) lowers that block's d before the
    # factorisation (sb_test_shift_d), to reach the pivoted-Cholesky fallback
    test_dshift: tuple | None = None
    handle: C.c_void_p = field(default=None, init=False)

    def __post_init__(self):
        o = self.opts
        self.method = o.method_code()
        self.nmix = o.resolved_nmix()
        if self.method == 6:
            if self.grp is None:
                raise ValueError("BayesRc needs SNP groups")
            self.grp_idx, self.ngroup = group_index(self.grp)
        else:
            self.grp_idx, self.ngroup = None, 1
        bs, nb, bt = partition(self.nloci, o.blocksize, o.nblocks, o.nthreads)
        # -SODAOff, or a partition into blocks of one SNP, which switches SODA
        # off (soda.f90:418-421; the reference then crashes in soda_bayes):
        # the single-site sampler, nblocks = blocksize = 0 at the boundary
        self.regular = bool(o.sodaoff) or bs <= 1
        if self.regular:
            bs, nb, bt = 0, 0, 0
        self.blocksize, self.nblocks, self.blockt = bs, nb, bt
        if self.freq is None:
            if self.bed is None:
                raise ValueError("freq is required when the genotypes are on the device")
            self.freq = allele_freq(self.bed, self.nind)
        if o.seed == 0 and o.rng == "replay":
            raise ValueError("replay needs a nonzero seed (seed 0 reads the system clock)")
        self.state = rng_seed(o.seed)
        if o.rblocks and not self.regular:
            self.permvec = permutate(self.nloci, bs, self.state)
        else:
            self.permvec = np.arange(self.nloci, dtype=np.int32)

    # -- setup_soda --------------------------------------------------------
    def setup(self) -> np.ndarray:
        L = _lib.load()
        o = self.opts
        d = _lib.SbDims(method=self.method, nind=self.nind, nloci=self.nloci, nmix=self.nmix,
                        ngroup=self.ngroup, nblocks=self.nblocks, blocksize=self.blocksize,
                        blockt=self.blockt, nchains=int(o.chains),
                        rng_mode=_lib.RNG_REPLAY if o.rng == "replay" else _lib.RNG_COUNTER,
                        precision=_lib.PREC_FP64 if o.precision == "fp64" else _lib.PREC_MIXED,
                        bed_on_device=1 if self.bed_device_ptr else 0)
        freq = np.ascontiguousarray(self.freq, dtype=np.float64)
        perm = np.ascontiguousarray(self.permvec, dtype=np.int32)
        if self.bed_device_ptr:
            bed_ptr = C.c_void_p(self.bed_device_ptr)
        else:
            self._bed = np.ascontiguousarray(self.bed, dtype=np.uint8)
            bed_ptr = self._bed.ctypes.data_as(C.c_void_p)
        grp = None
        if self.grp_idx is not None:
            self._grp = np.ascontiguousarray(self.grp_idx, dtype=np.int32)
            grp = self._grp.ctypes.data_as(C.POINTER(C.c_int32))
        h = C.c_void_p()
        _lib.check(L.sb_create(C.byref(d), bed_ptr, freq.ctypes.data_as(C.POINTER(C.c_double)),
                               perm.ctypes.data_as(C.POINTER(C.c_int32)), grp, C.byref(h)))
        self.handle = h
        _lib.check(L.sb_rng_set(h, self.state.ctypes.data_as(C.POINTER(C.c_uint64))))
        if self.test_dshift is not None:
            _lib.check(L.sb_test_shift_d(h, int(self.test_dshift[0]), float(self.test_dshift[1])))
        xxd = np.zeros(self.nblocks)
        fb, fc = C.c_int32(0), C.c_int32(0)
        rc = L.sb_build_aug(h, xxd.ctypes.data_as(C.POINTER(C.c_double)), C.byref(fb), C.byref(fc))
        _lib.check(rc)
        self.xxd = xxd
        na = self.get("NAIND", 2)
        self.naind, self.nskipped = int(na[0]), int(na[1])
        return xxd

    def skipped(self) -> np.ndarray:
        """1-based ids of the SNPs the pivoted-Cholesky fallback skipped, in
        the order of their updates (i_skipped, soda.f90:505-507)."""
        if not self.nskipped:
            return np.zeros(0, dtype=np.int64)
        return self.get("SKIPPED", self.nskipped).astype(np.int64)

    # -- soda_bayes --------------------------------------------------------
    def init(self) -> None:
        o = self.opts
        pr = _lib.SbPriors()
        pr.pizero, pr.scale_va, pr.scale_ve = f32(o.pi), f32(o.scale_va), f32(o.scale_ve)
        pr.df_va, pr.df_ve, pr.vara0 = f32(o.df_va), f32(o.df_ve), float(o.vara0)
        for k, m in enumerate(o.mix_values()):
            pr.mix[k] = m
        for k, a in enumerate(o.alpha_values()):
            pr.alpha[k] = a
        pr.numit, pr.burnin, pr.thin, pr.seed = o.numit, o.burnin, o.thin, int(o.seed)
        pr.resync = int(o.resync)
        self._why = np.ascontiguousarray(self.why, dtype=np.float64)
        chains = int(o.chains)
        if self._why.ndim == 1 and self.heldout is None:
            _lib.check(_lib.load().sb_mcmc_init(self.handle, C.byref(pr),
                                                self._why.ctypes.data_as(C.POINTER(C.c_double))))
        else:
            if self._why.shape not in ((self.nind,), (chains, self.nind)):
                raise ValueError(f"why must be ({self.nind},) or ({chains}, {self.nind})")
            ldy = self.nind if self._why.ndim == 2 else 0
            hp, ldh = None, 0
            if self.heldout is not None:
                self._held = np.ascontiguousarray(self.heldout, dtype=np.uint8)
                if self._held.shape not in ((self.nind,), (chains, self.nind)):
                    raise ValueError(f"heldout must be ({self.nind},) or ({chains}, {self.nind})")
                hp = self._held.ctypes.data_as(C.POINTER(C.c_uint8))
                ldh = self.nind if self._held.ndim == 2 else 0
            _lib.check(_lib.load().sb_mcmc_init_chains(self.handle, C.byref(pr),
                                                       self._why.ctypes.data_as(C.POINTER(C.c_double)), ldy,
                                                       hp, ldh))
        self.rep = 0

    def run(self, count: int | None = None, trace: bool = True) -> np.ndarray:
        """Run ``count`` iterations (default: up to numit); returns the trace
        records [rep, included, mu, vara, vare, snpindist..., varindist...]."""
        L = _lib.load()
        if count is None:
            count = self.opts.numit - self.rep
        first = self.rep + 1
        thin = max(1, self.opts.thin)
        nrec = sum(1 for r in range(first, first + count) if r % thin == 0)
        w = L.sb_hyp_width(self.handle)
        hyp = np.zeros((max(nrec, 1), w))
        ptr = hyp.ctypes.data_as(C.POINTER(C.c_double)) if trace else None
        _lib.check(L.sb_mcmc_run(self.handle, first, count, ptr))
        self.rep += count
        return hyp[:nrec]

    def save_state(self) -> bytes:
        """Checkpoint of every chain after its last completed iteration
        (sb_state_save)."""
        L = _lib.load()
        size = C.c_int64()
        _lib.check(L.sb_state_size(self.handle, C.byref(size)))
        buf = np.zeros(size.value, dtype=np.uint8)
        _lib.check(L.sb_state_save(self.handle, buf.ctypes.data_as(C.c_void_p), size.value))
        return buf.tobytes()

    def load_state(self, blob: bytes) -> None:
        """Continues the chains of a checkpoint (after setup(), in place of
        init(); sb_state_load)."""
        buf = np.frombuffer(blob, dtype=np.uint8).copy()
        _lib.check(_lib.load().sb_state_load(self.handle, buf.ctypes.data_as(C.c_void_p), buf.size))
        self.rep = int(self.info_raw()[1])

    def get(self, name: str, n: int, chain: int = 0) -> np.ndarray:
        out = np.zeros(n)
        _lib.check(_lib.load().sb_get(self.handle, _lib.Q[name], chain,
                                      out.ctypes.data_as(C.POINTER(C.c_double)), n))
        return out

    def info_raw(self) -> np.ndarray:
        v = np.zeros(15)
        _lib.check(_lib.load().sb_get_info(self.handle, v.ctypes.data_as(C.POINTER(C.c_double)), 15))
        return v

    def kernel_times(self) -> dict:
        """Per-kernel HIP-event timing accumulated over the runs made with
        SODABAYES_KERNEL_TIMING=1: {slot: (total ms, launches, algorithmic bytes)}."""
        v = self.get("KTIMES", 3 * len(_lib.KT_SLOTS))
        return {k: (v[3 * i], v[3 * i + 1], v[3 * i + 2]) for i, k in enumerate(_lib.KT_SLOTS)}

    def block_rhs(self, block: int) -> np.ndarray:
        """X_s'r + A_s'r_a of a block for the current state: (chains, n_b)."""
        a, e = block_bounds(self.nloci, self.blocksize, self.nblocks, self.blockt)[block]
        out = np.zeros((int(self.opts.chains), e - a))
        _lib.check(_lib.load().sb_block_rhs(self.handle, block, out.ctypes.data_as(C.POINTER(C.c_double))))
        return out

    def layout_stream(self) -> dict:
        """Read-only bandwidth of this handle's operand layouts
        (sb_layout_stream): {layout: (bytes per block, us per block, GB/s)}."""
        v = np.zeros(12)
        _lib.check(_lib.load().sb_layout_stream(self.handle, v.ctypes.data_as(C.POINTER(C.c_double))))
        names = (["geno_sw", "geno_t", "Ut0", "Ut1"] if self.opts.precision == "mixed" else ["geno", "U"])
        return {k: (v[3 * i], v[3 * i + 1], v[3 * i + 2]) for i, k in enumerate(names)}

    def info(self) -> dict:
        v = self.info_raw()
        return dict(setup_seconds=v[0], iterations=v[1], sweep_ms=v[2], device_bytes=v[3],
                    chunk_supersteps=int(v[12]), rows_per_wave=int(v[13]), tuned_block_us=v[14])

    def results(self, chain: int = 0) -> dict:
        """Final state and posterior means under the reference's names
        (models.f90:575-594)."""
        p, nb = self.nloci, self.nmix * self.ngroup
        m = self.method
        get = lambda name, n: self.get(name, n, chain)  # noqa: E731
        st = get("STORES", 5)
        cnt = st[0]
        sc = get("SCALARS", 4)
        out = {
            "nblocks": self.nblocks, "blocksize": self.blocksize, "blockt": self.blockt,
            "naind": self.naind, "nskipped": self.nskipped, "freq": self.freq, "permvec": self.permvec + 1,
            "xxd": self.get("XXD", self.nblocks),
            "mu": sc[0], "vare": sc[1], "vara": sc[2],
            "mu_store": st[1] / cnt, "vare_store": st[2] / cnt,
            "g": get("G", p), "gstore": get("GSTORE", p) / cnt,
            "yadj": get("YADJ", self.nind), "yAadj": get("YAADJ", self.blocksize),
            "varstore": get("VARSTORE", nb) / cnt, "varindist": get("VARINDIST", nb),
        }
        if m != 1:
            grp1 = (self.grp_idx + 1) if self.grp_idx is not None else np.ones(p)
            out["snptracker"] = np.concatenate([grp1, get("CLASS", p)])
            out["snpindist"] = get("SNPINDIST", nb)
            out["snpstore"] = get("SNPSTORE", nb) / cnt
            out["indiststore"] = get("INDISTSTORE", p * self.nmix) / cnt
            out["included_store"] = st[4] / cnt
            out["p"] = get("P", nb)
        if m not in (1, 2, 3):
            out["pstore"] = get("PSTORE", nb) / cnt
        if m in (1, 2):
            out["vara_s"] = get("VARA_S", p)
            out["vara_sstore"] = get("VARA_SSTORE", p) / cnt
        else:
            out["vara_store"] = st[3] / cnt
        return out

    def close(self) -> None:
        if self.handle:
            _lib.load().sb_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def autotune_blocksize(opts: SodaOptions, nind: int, nloci: int, why, candidates=(1024, 2048, 4096),
                       iters: int = 5, warmup: int = 2, **sampler_kw) -> tuple[int, dict]:
    """SODA subset size with the fastest sweep for this job's shape.

    Every subset size gives a valid SODA sampler of the same posterior (the
    blocking only decides which SNPs share an augmenting matrix), so the
    choice is one of speed.  For each candidate the job is set up (Gram,
    bound, factorisation), initialised and swept ``warmup`` + ``iters``
    times; returns the subset size with the lowest milliseconds per
    iteration and {size: ms per iteration}.  ``sampler_kw`` goes to
    SodaSampler (bed / bed_device_ptr, freq, grp, heldout).
    """
    import time
    from dataclasses import replace

    timings = {}
    for bs in candidates:
        if bs > nloci:
            continue
        o = replace(opts, blocksize=int(bs), nblocks=0, numit=warmup + iters, burnin=0)
        s = SodaSampler(o, nind, nloci, why, **sampler_kw)
        try:
            s.setup()
            s.init()
            s.run(warmup, trace=False)
            t0 = time.perf_counter()
            s.run(iters, trace=False)   # returns after the stream has drained
            timings[int(bs)] = (time.perf_counter() - t0) * 1e3 / iters
        finally:
            s.close()
    if not timings:
        raise ValueError("no candidate subset size fits nloci")
    best = min(timings, key=timings.get)
    return best, timings

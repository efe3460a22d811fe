/*
 * sodabayes.h -- C ABI of the MI355X (gfx950) SODA Gibbs sampler.
 *
 * Drop-in boundary for the SODA branch of the reference program:
 *   SODABayes.f90:139  call setup_soda   -> sb_create + sb_build_aug
 *   SODABayes.f90:142  call soda_bayes   -> sb_mcmc_init + sb_mcmc_run + sb_get
 * The Fortran host keeps option parsing, file IO and the output writers; it
 * binds these functions with `interface ... bind(C)` (iso_c_binding).
 *
 * Conventions: all arrays are plain host pointers in C order unless stated;
 * SNP ids and positions are 0-based; every function returns SB_OK (0) or an
 * error code, and sb_last_error() gives the message of the last failure on
 * the calling thread.  A handle is not re-entrant: calls on one handle must
 * not overlap.  One handle owns one HIP stream on the current device.
 */
#ifndef SODABAYES_H
#define SODABAYES_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SB_OK 0
#define SB_ERR_ARG (-1)      /* invalid argument or size */
#define SB_ERR_HIP (-2)      /* HIP runtime error (message in sb_last_error) */
#define SB_ERR_NOMEM (-3)    /* device allocation failed */
#define SB_ERR_STATE (-4)    /* call out of order (e.g. run before init) */
#define SB_ERR_NOT_PD (-5)   /* dI - X_s'X_s not positive definite: see sb_build_aug */

/* Bayesian methods, globals.f90:22 */
enum { SB_BAYES_A = 1, SB_BAYES_B = 2, SB_BAYES_C = 3, SB_BAYES_CPI = 4,
       SB_BAYES_R = 5, SB_BAYES_RC = 6 };

/* Random-number modes.
 * REPLAY : every draw comes from the reference's xoroshiro128+ stream 0, in
 *          the order of a -nthreads 1 run (parallelrng.f90, models.f90).
 * COUNTER: Philox4x32-10 keyed by (seed, chain), counter (iteration, SNP,
 *          draw): every SNP draws independently, in parallel. */
enum { SB_RNG_REPLAY = 0, SB_RNG_COUNTER = 1 };

/* Arithmetic of the genotype products X_s' r.
 * FP64 : FP64 products and sums.
 * MIXED: exact int8 MFMA on fixed-point residual limbs with int32 sums,
 *        combined in FP64 (norm-wise error stated in sb_get_info). */
enum { SB_PREC_FP64 = 0, SB_PREC_MIXED = 1 };

typedef struct {
    int32_t method;             /* SB_BAYES_* */
    int32_t nind, nloci;        /* n individuals, p SNPs */
    int32_t nmix, ngroup;       /* mixture classes (1 for BayesA), SNP groups */
    int32_t nblocks, blocksize, blockt; /* resolved partition, soda.f90:407-417;
                                         * all 0: -SODAOff, the single-site
                                         * sampler regular_bayes
                                         * (models.f90:607-962), FP64 only */
    int32_t nchains;            /* independent chains on the same genotypes */
    int32_t rng_mode;           /* SB_RNG_* (REPLAY needs nchains == 1) */
    int32_t precision;          /* SB_PREC_* */
    int32_t bed_on_device;      /* 1: `bed` in sb_create is a device pointer */
} sb_dims;

/* Priors and chain settings (useroptions.f90:488-552, already widened from
 * the reference's float32 option parsing by the caller). */
typedef struct {
    double pizero, scale_va, scale_ve, df_va, df_ve;
    double vara0;               /* module value of vara before the job */
    double mix[20];             /* class variances as a fraction of Va */
    double alpha[20];           /* Dirichlet prior */
    int32_t numit, burnin, thin;
    int32_t resync;             /* iterations between residual resyncs r = y - X g - mu
                                 * (compute_residuals, models.f90:571-573); 0: the
                                 * reference's 1000 */
    int64_t seed;               /* COUNTER mode key (REPLAY uses sb_rng_set) */
} sb_priors;

typedef struct sb_handle sb_handle;

/* ---- host-side RNG of the reference (parallelrng.f90), for the host ---- */

/* stream 0 of set_random_seed(seed) (utilities.f90:504-517): state[2] */
int sb_rng_seed(int64_t seed, int32_t stream, uint64_t state[2]);
/* permutate(p, n, s) (utilities.f90:588-607): permvec 0-based, advances state */
int sb_permutate(int32_t n, int32_t s, uint64_t state[2], int32_t *permvec);

/* ---- setup (replaces setup_soda, soda.f90:403-443) ---- */

/* bed: nloci x ceil(nind/4) SNP-major PLINK payload (no magic bytes), in
 * SNP-id order; freq: allele frequency per SNP id (utilities.f90:527);
 * permvec: SNP id at each block position (soda.f90:429-435); snpgrp: 0-based
 * group per SNP id (BayesRc) or NULL. */
int sb_create(const sb_dims *dims, const uint8_t *bed, const double *freq,
              const int32_t *permvec, const int32_t *snpgrp, sb_handle **out);
int sb_destroy(sb_handle *h);

/* REPLAY mode: load / read back the stream-0 xoroshiro128+ state. */
int sb_rng_set(sb_handle *h, const uint64_t state[2]);
int sb_rng_get(sb_handle *h, uint64_t state[2]);

/* Builds the augmenting matrices (get_aug_matrix, soda.f90:445-518):
 * int8-MFMA Gram X_s'X_s (exact int32) with the FP64 standardisation of
 * xxproduct (soda.f90:520-556), power-method bound d = lambda + 0.001
 * (soda.f90:378-401, start vectors drawn from the replay stream), and the
 * FP64 Cholesky of dI - X_s'X_s (soda.f90:8-47).  xxd: nblocks values.
 * A block whose dI - X_s'X_s is not positive definite takes the reference's
 * fallback (soda.f90:492-510): pivoted Cholesky (soda.f90:49-124) on the
 * host, a dense augmenting block, and the SNPs beyond its rank skipped in
 * the block sweep and updated one by one after it (SB_Q_SKIPPED,
 * SB_Q_NAIND).  Returns SB_ERR_NOT_PD when the pivoted factorisation fails
 * too (the reference stops there), or, with *fail_block / *fail_col
 * (1-based) set, for any failing block under SB_PREC_MIXED: the fallback
 * runs on the FP64 path only.
 * For the single-site sampler (nblocks = 0) it builds no matrices and
 * leaves xxd untouched: every SNP is updated one by one, in permvec order. */
int sb_build_aug(sb_handle *h, double *xxd, int32_t *fail_block, int32_t *fail_col);

/* Test hook (no counterpart in the reference; for tests that must reach the
 * pivoted-Cholesky fallback on small data): the next sb_build_aug on this
 * handle lowers block `block`'s (0-based) bound d by `amount` before the
 * factorisation, so that dI - X_s'X_s is indefinite there.  block < 0
 * clears it.  Nothing else sets it: a job never runs with a shifted d
 * unless its caller asks for one. */
int sb_test_shift_d(sb_handle *h, int32_t block, double amount);

/* ---- MCMC (replaces soda_bayes, models.f90:9-605) ---- */

/* Initial state, models.f90:56-128: why has nind values. */
int sb_mcmc_init(sb_handle *h, const sb_priors *pr, const double *why);
/* Batched traits and cross-validation folds (no counterpart in the reference,
 * whose jobs are one phenotype on all individuals): the initial state of
 * every chain from its own phenotype, chain c's nind values at why + c*ldy
 * (ldy = 0: one column for all chains, as sb_mcmc_init).  heldout (may be
 * NULL): chain c's nind flags at heldout + c*ldh; a nonzero flag marks an
 * individual whose phenotype chain c does not use (a validation row).  Such
 * a row's phenotype is treated as missing: it starts at the chain's training
 * mean and is redrawn every iteration from its full conditional
 * y ~ N(mu + x'g, vare), i.e. its residual becomes a fresh N(0, vare) draw
 * after vare's update, so the chain samples the posterior given the
 * training rows while the augmenting matrices stay shared.  Held-out rows
 * need SB_RNG_COUNTER. */
int sb_mcmc_init_chains(sb_handle *h, const sb_priors *pr, const double *why, int64_t ldy,
                        const uint8_t *heldout, int64_t ldh);
/* Runs iterations rep = first .. first+count-1 (models.f90:129-574): the
 * sweep, hyperparameter draws, posterior accumulation and the residual
 * resync every 1000 iterations.  hyp (optional, host) receives one record
 * per iteration with mod(rep, thin) == 0 for chain 0:
 *   [rep, included, mu, vara, vare, snpindist(ngroup*nmix), varindist(...)]
 * i.e. 5 + 2*ngroup*nmix doubles per record. */
int sb_mcmc_run(sb_handle *h, int32_t first, int32_t count, double *hyp);
/* doubles per hyp record */
int32_t sb_hyp_width(const sb_handle *h);

/* Downloads one quantity of one chain.  Per-SNP arrays are in SNP-id order. */
enum {
    SB_Q_G = 1,          /* effects g (nloci) */
    SB_Q_CLASS,          /* snptracker(:,2) as double, 1-based (nloci) */
    SB_Q_VARA_S,         /* per-SNP variances, BayesA/B (nloci) */
    SB_Q_YADJ,           /* residuals (nind) */
    SB_Q_YAADJ,          /* augmented residuals (blocksize) */
    SB_Q_P,              /* mixture proportions (ngroup*nmix, column-major) */
    SB_Q_SCALARS,        /* mu, vare, vara, included (4) */
    SB_Q_GSTORE,         /* sum of g over kept samples (nloci) */
    SB_Q_INDISTSTORE,    /* class counts (nloci*nmix, column-major nloci x nmix) */
    SB_Q_VARA_SSTORE,    /* sum of vara_s (nloci) */
    SB_Q_STORES,         /* counter, mu, vare, vara, included sums (5) */
    SB_Q_SNPSTORE,       /* (ngroup*nmix) */
    SB_Q_VARSTORE,       /* (ngroup*nmix) */
    SB_Q_PSTORE,         /* (ngroup*nmix) */
    SB_Q_XA,             /* augmenting matrices, blocksize x nloci column-major */
    SB_Q_XXD,            /* d per block (nblocks) */
    SB_Q_GRAM,           /* standardised Gram of block `chain` (n_b x n_b), recomputed */
    SB_Q_SNPINDIST,      /* last iteration's SNPs per (group, class) (ngroup*nmix) */
    SB_Q_VARINDIST,      /* last iteration's sum of g^2 per (group, class) (ngroup*nmix) */
    SB_Q_KTIMES,         /* per-kernel timing of the runs made with SODABAYES_KERNEL_TIMING=1:
                            SB_KT_N x (summed ms, launches, algorithmic bytes), slot order below */
    SB_Q_SKIPPED,        /* SNPs skipped by the pivoted-Cholesky fallback, 1-based ids in update
                            order (i_skipped, soda.f90:505-507) (nskipped) */
    SB_Q_NAIND,          /* naind, nskipped (2) (soda.f90:440, 505, 514) */
    SB_Q_WHY,            /* the chain's phenotype (nind), held-out rows at their current draw */
    SB_Q_XA_BLOCK        /* augmenting matrix of block `chain` (n_b x n_b column-major, block positions) */
};

/* Kernel timing slots of SB_Q_KTIMES (diagnostics; no counterpart in the
 * reference).  FP64 path: block products, draw, changed-SNP list, update.
 * Batched path: residual statistics + limb quantisation, X_s'Q with U_s'R_a,
 * limb combination + draw, delta quantisation, X_s C update with U_s D, r_a
 * reduction.  The U products run inside the X_s'Q and update launches, so
 * SB_KT_UT0 and SB_KT_UT1 stay empty (kept for the slot order). */
enum {
    SB_KT_RHS = 0, SB_KT_DRAW, SB_KT_LIST, SB_KT_UPDATE,
    SB_KT_RQUANT, SB_KT_XQ, SB_KT_UT0, SB_KT_DRAWB, SB_KT_DQUANT, SB_KT_XUPD, SB_KT_UT1, SB_KT_RARED,
    SB_KT_N
};
int sb_get(sb_handle *h, int32_t what, int32_t chain, double *out, int64_t count);

/* Diagnostics (no counterpart in the reference): X_s'r + A_s'r_a of every SNP
 * of block `block` for the current state, through the handle's product path
 * (FP64 or mixed), without drawing.  out: n_b x nchains, chain-major. */
int sb_block_rhs(sb_handle *h, int32_t block, double *out);

/* Diagnostics: fills up to n doubles: [setup seconds, iterations run,
 * last sweep milliseconds, device bytes allocated, then -- with the
 * environment variable SODABAYES_KERNEL_TIMING=1 during sb_mcmc_run -- summed
 * HIP-event milliseconds and launch counts of the block kernels:
 * rhs ms, rhs launches, rhs algorithmic bytes, draw ms, draw launches,
 * update ms, update launches; then the largest power-method iteration
 * count of sb_build_aug; then the batched path's tile shape: super-steps per
 * X_s'Q chunk, row tiles per wave, and the tuned block-0 products + update
 * microseconds (0 when the shape was fixed by SODABAYES_KSS /
 * SODABAYES_ROWS_PER_WAVE)]. */
int sb_get_info(sb_handle *h, double *out, int32_t n);

/* ---- prediction (replaces compute_dgv, utilities.f90:616-622, after
 * xscale's prediction branch, utilities.f90:540-553; -predict,
 * SODABayes.f90:145-148) ---- */

/* Genomic values pred_i = mu + sum_j z_ij g_j of nind individuals over nloci
 * SNPs, z_ij = (x_ij - 2 freq_j) / sqrt(2 freq_j (1 - freq_j)), z = 0 where
 * freq_j < eps or 1 - freq_j < eps.  bed: SNP-major PLINK payload without
 * the magic bytes (nloci x ceil(nind/4) bytes, missing read as 1), on the
 * host or (bed_on_device != 0) in device memory; freq, effects: nloci
 * doubles in .bim order (the .frq / .eff columns setup_prediction reads);
 * pred: nind doubles out.  No handle: the call allocates and frees its own
 * device buffers. */
int sb_predict(int32_t nind, int32_t nloci, const uint8_t *bed, int32_t bed_on_device, const double *freq,
               const double *effects, double mu, double *pred);

/* Prediction that allocates nothing (no counterpart in the reference; for
 * callers that cannot allocate device memory: code under HIP graph capture,
 * serving loops, harnesses that forbid hipMalloc).  The same genomic values
 * as sb_predict, from device pointers only: bed (nloci x ceil(nind/4) .bed
 * payload), freq and effects (nloci doubles each) in, pred (nind doubles)
 * out, on `stream` (a hipStream_t; NULL: the default stream), asynchronous
 * like any kernel launch.  The per-chunk partial sums go to a fixed device
 * array of the library of SB_PREDICT_NOALLOC_MAX_PARTIALS doubles (16 MiB;
 * each device keeps its own copy until the process ends), which bounds
 * nind to that many individuals (SB_ERR_ARG beyond).  No hipMalloc or
 * hipFree is called.  Every call on a device uses that device's one array:
 * calls that overlap on one device race. */
#define SB_PREDICT_NOALLOC_MAX_PARTIALS 2097152
int sb_predict_noalloc(int32_t nind, int32_t nloci, const uint8_t *bed, const double *freq,
                       const double *effects, double mu, double *pred, void *stream);

/* ---- device-state checkpoints (no counterpart in the reference; for long
 * chains that must survive a job's end) ---- */

/* Bytes of a checkpoint of this handle's chains (after sb_mcmc_init). */
int sb_state_size(sb_handle *h, int64_t *bytes);
/* Copies every chain's state after its last completed iteration into buf
 * (size >= sb_state_size): priors, scalars and the replay generator state,
 * r, r_a, phenotypes and held-out flags, effects, classes, per-SNP
 * variances and the posterior sums, with the iteration count. */
int sb_state_save(sb_handle *h, void *buf, int64_t size);
/* Restores a checkpoint into a handle built on the same inputs (sb_create +
 * sb_build_aug with the same dims and genotypes; checked through the dims,
 * naind and the bounds d, which must match bit for bit) in place of
 * sb_mcmc_init.  sb_mcmc_run then continues from iteration count + 1 and
 * gives the chain the uninterrupted run would have given. */
int sb_state_load(sb_handle *h, const void *buf, int64_t size);

/* Diagnostics (no counterpart in the reference): the read bandwidth of this
 * handle's operand layouts alone.  For each layout, one read-only pass per
 * block (1 KB per wave load, one launch per block as the sweep launches its
 * kernels) over every block; out[3k .. 3k+2] = bytes per block,
 * microseconds per block, GB/s, for k = 0..3: the batched path's SNP-major
 * and individual-major genotype tiles and its two int8-limb U tile layouts, or
 * the FP64 path's genotype words and packed U (k = 2, 3 zero).  12 doubles. */
int sb_layout_stream(sb_handle *h, double *out);

/* Diagnostics of the COUNTER generator (no counterpart in the reference):
 * for each of n draw sites k, the first `per` uniforms in [0, 1) of the
 * stream the sampler opens for chain `chain` of a handle seeded `seed`, at
 * counter (iteration ctr[3k], site ctr[3k+1], tag ctr[3k+2]); out holds
 * n x per doubles, site-major.  Runs on the current device. */
int sb_counter_uniforms(int64_t seed, int32_t chain, const uint32_t *ctr, int32_t n, int32_t per,
                        double *out);

/* Message of the last failure on this thread ("" if none). */
const char *sb_last_error(void);

#ifdef __cplusplus
}
#endif
#endif

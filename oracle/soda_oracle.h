/*
 * ORACLE -- test infrastructure only.  A plain-C, FP64, single-threaded
 * restatement of the reference's SODA path (src/soda.f90, src/models.f90
 * soda_bayes, src/parallelrng.f90, the hot subset of src/utilities.f90),
 * used by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg as
 * the checker.  The product never links or calls it.
 *
 * Pinning (each function is checked by tests/test_oracle_*.py):
 *   RNG + distributions  -- bit-exact against the reference-linked golden
 *                           driver (tests/golden/rng_*.bin) and against
 *                           moment / KS checks
 *   xxproduct            -- brute-force integer Gram on tiny inputs, golden XA
 *   power method         -- golden xxd, numpy eigvalsh
 *   choleskyPD / PSD     -- golden XA, scipy cholesky, U^T U = M, dpstrf rank
 *   soda_bayes (T=1)     -- final chain state of reference runs
 *                           (tests/golden/soda_*.bin) for every method
 * Semantics follow the reference with -nthreads 1 (every draw from stream 0).
 */
#ifndef SODA_ORACLE_H
#define SODA_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct { uint64_t s1, s2; } orc_rng;

/* Philox4x32-10 (Salmon et al., SC'11): the build's counter-based generator,
 * not a reference routine; pinned by the Random123 known-answer vectors */
void     orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
uint64_t orc_chain_key(uint64_t seed, int chain);
void     orc_counter_uniforms(uint64_t key, uint32_t iter, uint32_t site, uint32_t tag, int n, double *out);
void     orc_counter_first(uint64_t key, uint32_t iter, uint32_t site0, uint32_t tag, int n, double *out);

/* parallelrng.f90 */
void     orc_rng_seed(int64_t seed, int nstreams, orc_rng *streams);
uint64_t orc_rng_next(orc_rng *s);
double   orc_uniform(orc_rng *s);
void     orc_uniform_array(orc_rng *s, int n, double *out);
double   orc_normal(orc_rng *s, double mean, double sd);
double   orc_gamma(orc_rng *s, double shape, double scale);
double   orc_sinvchi2(orc_rng *s, double dof, double scale);
void     orc_dirichlet(orc_rng *s, int n, const double *a, double *x);

/* utilities.f90 / soda.f90 building blocks */
int64_t  orc_nbatch(int nind);
void     orc_pack_genbits(const uint8_t *bed, int nind, int nloci, int64_t *genbits);
void     orc_freq(const uint8_t *bed, int nind, int nloci, double *freq);
void     orc_standardize(const uint8_t *bed, int nind, int nloci, const double *freq,
                         double *X);
double   orc_xxproduct(const int64_t *a, const int64_t *b, int64_t nbatch, int nind,
                       double fj, double fk);
int64_t  orc_gram_count(const int64_t *a, const int64_t *b, int64_t nbatch);
void     orc_gram_block(const uint8_t *bed, int nind, int m, const double *f, double *G);
void     orc_predict(const uint8_t *bed, int nind, int nloci, const double *freq, const double *eff,
                     double mu, double *pred);
int      orc_power_method(int n, const double *A, orc_rng *rng, double *lambda);
int      orc_cholesky_pd(int n, const double *A, double *U);
int      orc_cholesky_psd(int n, double *A, int *piv, int *rank);
void     orc_permutate(int *p, int n, int s, orc_rng *rng);

/* Whole job: setup_soda + soda_bayes.  Methods 1..6 = A, B, C, Cpi, R, Rc. */
typedef struct {
    int32_t method, nind, nloci, nmix, ngroup;
    int32_t numit, burnin, thin;
    int32_t blocksize, nblocks;   /* options; 0 = not given */
    int32_t rblocks, nthreads, maxthreads;
    int32_t dshift_block;         /* test hook: block whose d is lowered (-1: none) */
    int64_t seed;
    double  pizero, scale_va, scale_ve, df_va, df_ve;
    double  vara0;                /* value of module vara before the job */
    double  mix[20], alpha[20];
    double  dshift;               /* amount subtracted from that block's d */
} orc_opts;

/* Runs the job and writes a tagged dump (same records as the golden
 * driver plus per-iteration hyp_* traces).  grp: 1-based SNP group per SNP
 * (BayesRc) or NULL.  Returns 0 on success, negative on error. */
int orc_run(const orc_opts *o, const uint8_t *bed, const double *why,
            const int32_t *grp, const char *dump_path);

#ifdef __cplusplus
}
#endif
#endif

/*
 * ORACLE -- test infrastructure only (see soda_oracle.h).  Plain-C FP64
 * restatement of the reference's SODA path with -nthreads 1 semantics.
 * Each function cites the reference lines it follows.  Parity status of
 * each function is listed in soda_oracle.h.
 */
#define _POSIX_C_SOURCE 200809L
#include "soda_oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

/* parallelrng.f90:20 -- PI is a default-real (f32) literal widened to f64 */
static const double PI_F32 = (double)3.141592653589793238462f;
/* soda.f90:385 tol = 1.0e-14 (f32 literal), soda.f90:485 0.001 (f32) */
static const double POWER_TOL = (double)1.0e-14f;
static const double D_EPS_ADD = (double)0.001f;
static const double DEPS = 2.220446049250313080847e-16; /* epsilon(1.0d0) */

/* ------------------------------------------------------------------ RNG -- */

/* Philox4x32-10: J. K. Salmon, M. A. Moraes, R. O. Dror, D. E. Shaw,
 * "Parallel random numbers: as easy as 1, 2, 3", SC'11, section 3.2.  Not a
 * routine of the reference (whose generator is xoroshiro128+): it is the
 * build's counter-based generator, restated here from the paper so that
 * the device's uniforms can be checked bit for bit.  One round maps
 * (c0, c1, c2, c3) with key (k0, k1) to
 *   (hi(M1 c2) ^ c1 ^ k0, lo(M1 c2), hi(M0 c0) ^ c3 ^ k1, lo(M0 c0)),
 * M0 = 0xD2511F53, M1 = 0xCD9E8D57; the key is bumped by the Weyl constants
 * (0x9E3779B9, 0xBB67AE85) between rounds; 10 rounds.  Pinned by the
 * known-answer vectors the authors publish with their Random123 library
 * (tests/test_oracle_math.py). */
void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4])
{
    uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3], k0 = key[0], k1 = key[1];
    for (int r = 0; r < 10; r++) {
        if (r > 0) {
            k0 += 0x9E3779B9u;
            k1 += 0xBB67AE85u;
        }
        const uint64_t p0 = (uint64_t)0xD2511F53u * c0, p1 = (uint64_t)0xCD9E8D57u * c2;
        const uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0, n1 = (uint32_t)p1;
        const uint32_t n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1, n3 = (uint32_t)p0;
        c0 = n0;
        c1 = n1;
        c2 = n2;
        c3 = n3;
    }
    out[0] = c0;
    out[1] = c1;
    out[2] = c2;
    out[3] = c3;
}

/* Key of chain c in the counter-based mode: seed ^ (golden-ratio odd
 * constant * (c + 1)), split into (low, high) 32-bit key words. */
uint64_t orc_chain_key(uint64_t seed, int chain)
{
    return seed ^ (0x9E3779B97F4A7C15ULL * (uint64_t)(chain + 1));
}

static inline double int64_to_real64(uint64_t x);

/* The counter-based stream of one draw site: block s = Philox(counter
 * (iter, site, tag, s), key) gives two 64-bit words (w0 << 32 | w1) and
 * (w2 << 32 | w3), each mapped to [0, 1) as the reference maps its 64-bit
 * outputs (rng_int64_to_real64, parallelrng.f90:242-249). */
void orc_counter_uniforms(uint64_t key, uint32_t iter, uint32_t site, uint32_t tag, int n, double *out)
{
    const uint32_t k[2] = {(uint32_t)key, (uint32_t)(key >> 32)};
    for (int i = 0, s = 0; i < n; s++) {
        const uint32_t c[4] = {iter, site, tag, (uint32_t)s};
        uint32_t b[4];
        orc_philox4x32_10(c, k, b);
        out[i++] = int64_to_real64(((uint64_t)b[0] << 32) | b[1]);
        if (i < n) out[i++] = int64_to_real64(((uint64_t)b[2] << 32) | b[3]);
    }
}

/* First uniform of each of n consecutive sites (site0, site0 + 1, ...) of
 * one (key, iteration, tag): the per-SNP draws of one sweep. */
void orc_counter_first(uint64_t key, uint32_t iter, uint32_t site0, uint32_t tag, int n, double *out)
{
    for (int j = 0; j < n; j++) orc_counter_uniforms(key, iter, site0 + (uint32_t)j, tag, 1, out + j);
}

static inline uint64_t rotl64(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }

/* rng_xoroshiro128plus, parallelrng.f90:100-123 */
uint64_t orc_rng_next(orc_rng *s)
{
    uint64_t s1 = s->s1, s2 = s->s2;
    uint64_t r = s1 + s2;
    s2 ^= s1;
    s1 = rotl64(s1, 55) ^ s2 ^ (s2 << 14);
    s2 = rotl64(s2, 36);
    s->s1 = s1;
    s->s2 = s2;
    return r;
}

/* rng_jump, parallelrng.f90:126-149 */
static void rng_jump(orc_rng *s)
{
    static const uint64_t jmp[2] = {0xbeac0467eba5facbULL, 0xd86b048b86aa9922ULL};
    uint64_t a = 0, b = 0;
    for (int i = 0; i < 2; i++)
        for (int bit = 0; bit < 64; bit++) {
            if (jmp[i] & (1ULL << bit)) {
                a ^= s->s1;
                b ^= s->s2;
            }
            orc_rng_next(s);
        }
    s->s1 = a;
    s->s2 = b;
}

/* set_random_seed utilities.f90:504-517 + rng_zigset parallelrng.f90:54-78:
 * seed (|seed|, |seed|+1), stream k = jump^(k+1)(seed). */
void orc_rng_seed(int64_t seed, int nstreams, orc_rng *streams)
{
    orc_rng s;
    uint64_t a = (uint64_t)(seed < 0 ? -seed : seed);
    s.s1 = a;
    s.s2 = a + 1;
    for (int k = 0; k < nstreams; k++) {
        rng_jump(&s);
        streams[k] = s;
    }
}

/* rng_int64_to_real64, parallelrng.f90:242-249 */
static inline double int64_to_real64(uint64_t x)
{
    uint64_t b = 0x3FF0000000000000ULL | (x >> 12);
    double d;
    memcpy(&d, &b, sizeof d);
    return d - 1.0;
}

/* rng_uni, parallelrng.f90:163-177 */
double orc_uniform(orc_rng *s) { return int64_to_real64(orc_rng_next(s)); }

/* rng_uni_array, parallelrng.f90:179-194 */
void orc_uniform_array(orc_rng *s, int n, double *out)
{
    for (int i = 0; i < n; i++) out[i] = orc_uniform(s);
}

/* rand_normal, parallelrng.f90:264-277 (sin branch only, f32 PI) */
double orc_normal(orc_rng *s, double mean, double sd)
{
    if (sd <= 0.0) return mean;
    double u1 = orc_uniform(s);
    double u2 = orc_uniform(s);
    /* `x**0.5d0` is lowered to sqrt by the reference's LLVM-based build */
    double r = sqrt(-2.0 * log(u1));
    double theta = 2.0 * PI_F32 * u2;
    return mean + sd * r * sin(theta);
}

/* rand_gamma, parallelrng.f90:294-333 (Marsaglia-Tsang) */
double orc_gamma(orc_rng *s, double shape, double scale)
{
    if (shape >= 1.0) {
        double d = shape - 1.0 / 3.0;
        double c = 1.0 / sqrt(9.0 * d);
        for (;;) {
            double x = orc_normal(s, 0.0, 1.0);
            double v = 1.0 + c * x;
            while (v <= 0.0) {
                x = orc_normal(s, 0.0, 1.0);
                v = 1.0 + c * x;
            }
            v = v * v * v;
            double u = orc_uniform(s);
            double xsq = x * x;
            if ((u < 1.0 - 0.0331 * xsq * xsq) ||
                (log(u) < 0.5 * xsq + d * (1.0 - v + log(v))))
                return scale * d * v;
        }
    }
    double g = orc_gamma(s, shape + 1.0, 1.0);
    double w = orc_uniform(s);
    return scale * g * pow(w, 1.0 / shape);
}

/* rand_scaled_inverse_chi_square + rand_inverse_gamma, parallelrng.f90:348-362 */
double orc_sinvchi2(orc_rng *s, double dof, double scale)
{
    double shape = 0.5 * dof;
    double sc = 0.5 * dof * scale;
    return 1.0 / orc_gamma(s, shape, 1.0 / sc);
}

/* rand_dirichlet, parallelrng.f90:445-455 */
void orc_dirichlet(orc_rng *s, int n, const double *a, double *x)
{
    double sx = 0.0;
    for (int i = 0; i < n; i++) x[i] = orc_gamma(s, a[i], 1.0);
    for (int i = 0; i < n; i++) sx += x[i];
    for (int i = 0; i < n; i++) x[i] = x[i] / sx;
}

/* ------------------------------------------------------------ genotypes -- */

/* soda.f90:436 nbatch = ceiling(nind / 32) */
int64_t orc_nbatch(int nind) { return ((int64_t)nind + 31) / 32; }

/* load_geno_bits, utilities.f90:462-502: little-endian 64-bit words of each
 * SNP's .bed bytes, w - ((w >> 1) & 0x5555...) turns codes into counts. */
void orc_pack_genbits(const uint8_t *bed, int nind, int nloci, int64_t *genbits)
{
    const int64_t nb = (nind + 3) / 4, nbatch = orc_nbatch(nind);
    const uint64_t M5 = 0x5555555555555555ULL;
    for (int j = 0; j < nloci; j++) {
        const uint8_t *col = bed + (int64_t)j * nb;
        for (int64_t w = 0; w < nbatch; w++) {
            uint64_t v = 0;
            for (int b = 0; b < 8; b++) {
                int64_t idx = w * 8 + b;
                if (idx < nb) v |= (uint64_t)col[idx] << (8 * b);
            }
            v = v - ((v >> 1) & M5);
            genbits[(int64_t)j * nbatch + w] = (int64_t)v;
        }
    }
}

static inline int bed_count(const uint8_t *col, int i)
{
    int c = (col[i >> 2] >> (2 * (i & 3))) & 3;
    return c - ((c >> 1) & 1); /* utilities.f90:449-451, missing -> 1 */
}

/* xscale (mcmc branch), utilities.f90:525-534: p = sum(x) / (2 n) */
void orc_freq(const uint8_t *bed, int nind, int nloci, double *freq)
{
    const int64_t nb = (nind + 3) / 4;
    for (int j = 0; j < nloci; j++) {
        const uint8_t *col = bed + (int64_t)j * nb;
        double s = 0.0;
        for (int i = 0; i < nind; i++) s += (double)bed_count(col, i);
        freq[j] = s / (2.0 * nind);
    }
}

/* xscale, utilities.f90:528-532: (x - 2p)/sqrt(2p(1-p)), or 0 if monomorphic */
void orc_standardize(const uint8_t *bed, int nind, int nloci, const double *freq, double *X)
{
    const int64_t nb = (nind + 3) / 4;
    for (int j = 0; j < nloci; j++) {
        const uint8_t *col = bed + (int64_t)j * nb;
        double p = freq[j];
        double *xc = X + (int64_t)j * nind;
        if (p < DEPS || (1.0 - p) < DEPS) {
            for (int i = 0; i < nind; i++) xc[i] = 0.0;
        } else {
            double sd = sqrt(2.0 * p * (1.0 - p));
            for (int i = 0; i < nind; i++) xc[i] = ((double)bed_count(col, i) - 2.0 * p) / sd;
        }
    }
}

/* Genomic values of the prediction run (-predict): xscale's prediction branch
 * standardises the genotypes with the frequencies read from the .frq file
 * (utilities.f90:540-553), compute_dgv adds mu to the row dot products with
 * the effects (utilities.f90:616-622): pred_i = mu + sum_j X(i,j) g_j, the
 * sum over j in order. */
void orc_predict(const uint8_t *bed, int nind, int nloci, const double *freq, const double *eff, double mu,
                 double *pred)
{
    const int64_t nb = (nind + 3) / 4;
    double *sd = malloc(sizeof(double) * (size_t)nloci);
    for (int j = 0; j < nloci; j++) {
        double p = freq[j];
        sd[j] = (p < DEPS || (1.0 - p) < DEPS) ? 0.0 : sqrt(2.0 * p * (1.0 - p));
    }
    for (int i = 0; i < nind; i++) {
        double s = 0.0;
        for (int j = 0; j < nloci; j++) {
            double x = sd[j] > 0.0 ? ((double)bed_count(bed + (int64_t)j * nb, i) - 2.0 * freq[j]) / sd[j] : 0.0;
            s += x * eff[j];
        }
        pred[i] = mu + s;
    }
    free(sd);
}

/* Integer part of xxproduct, soda.f90:541-553: 3 nd + 2 ng + nc = sum x_a x_b */
__attribute__((target("popcnt"))) int64_t orc_gram_count(const int64_t *a, const int64_t *b, int64_t nbatch)
{
    const uint64_t M5 = 0x5555555555555555ULL;
    int64_t nc = 0, nd = 0, ng = 0;
    for (int64_t i = 0; i < nbatch; i++) {
        uint64_t x = (uint64_t)a[i], y = (uint64_t)b[i];
        uint64_t c = x & y;
        uint64_t d = (c >> 1) & M5;
        uint64_t e = ((x >> 1) & y) & M5;
        uint64_t f = (x & (y >> 1)) & M5;
        uint64_t g = e | f;
        nc += __builtin_popcountll(c);
        nd += __builtin_popcountll(d);
        ng += __builtin_popcountll(g);
    }
    return nd * 3 + ng * 2 + nc;
}

/* xxproduct, soda.f90:520-556 */
double orc_xxproduct(const int64_t *a, const int64_t *b, int64_t nbatch, int nind,
                     double fj, double fk)
{
    double pj = fj, pk = fk, qj = 1.0 - pj, qk = 1.0 - pk;
    if (pj <= DEPS || qj <= DEPS || pk <= DEPS || qk <= DEPS) return 0.0;
    pj = pj * 2.0;
    pk = pk * 2.0;
    double res = (double)orc_gram_count(a, b, nbatch);
    res = (res - (double)nind * pj * pk) / sqrt(pj * pk * qj * qk);
    return res;
}

/* Standardised Gram of m SNP columns (get_aug_matrix's loop over pairs,
 * soda.f90:472-483, each entry xxproduct soda.f90:520-556): bed holds the m
 * columns' .bed rows (m x ceil(nind/4)), f their frequencies; G is m x m. */
void orc_gram_block(const uint8_t *bed, int nind, int m, const double *f, double *G)
{
    const int64_t nbatch = orc_nbatch(nind);
    int64_t *gb = malloc(sizeof(int64_t) * (size_t)nbatch * (size_t)m);
    orc_pack_genbits(bed, nind, m, gb);
    for (int a = 0; a < m; a++)
        for (int b = 0; b <= a; b++) {
            const double v = orc_xxproduct(gb + (int64_t)a * nbatch, gb + (int64_t)b * nbatch, nbatch, nind, f[a],
                                           f[b]);
            G[(int64_t)a * m + b] = v;
            G[(int64_t)b * m + a] = v;
        }
    free(gb);
}

/* power_method, soda.f90:378-401.  A is n x n column-major.  The start
 * vector is n stream uniforms (random(n,y,p), parallelrng.f90:179-194). */
int orc_power_method(int n, const double *A, orc_rng *rng, double *lambda)
{
    double *y = malloc(sizeof(double) * (size_t)n);
    double *v = malloc(sizeof(double) * (size_t)n);
    long maxit = 1000L * n;
    double lam = 0.0;
    long it;
    orc_uniform_array(rng, n, y);
    for (it = 1; it <= maxit; it++) {
        double ss = 0.0;
        for (int i = 0; i < n; i++) ss += y[i] * y[i];
        double nrm = sqrt(ss);
        for (int i = 0; i < n; i++) v[i] = y[i] / nrm;
        for (int i = 0; i < n; i++) y[i] = 0.0;
        for (int j = 0; j < n; j++) {
            const double *col = A + (int64_t)j * n;
            double vj = v[j];
            for (int i = 0; i < n; i++) y[i] += col[i] * vj;
        }
        lam = 0.0;
        for (int i = 0; i < n; i++) lam += y[i] * v[i];
        double delta = 0.0;
        for (int i = 0; i < n; i++) {
            double t = y[i] - lam * v[i];
            delta += t * t;
        }
        delta = delta / fabs(lam);
        if (delta <= POWER_TOL) break;
    }
    *lambda = fabs(lam);
    free(y);
    free(v);
    return (int)(it > maxit ? maxit : it);
}

/* choleskyPD, soda.f90:8-47 (upper U, dot-product form).  Returns 0, -1, or
 * the 1-based failing pivot.  U is fully written (zeros below the diagonal);
 * rows >= the failing pivot keep the zeros set at soda.f90:32. */
int orc_cholesky_pd(int n, const double *A, double *U)
{
    if (n <= 0) return -1;
    for (int i = 0; i < n; i++)
        if (A[(int64_t)i * n + i] <= 0.0) return -1;
    for (int64_t k = 0; k < (int64_t)n * n; k++) U[k] = 0.0;
#define UU(r, c) U[(int64_t)(c) * n + (r)]
#define AA(r, c) A[(int64_t)(c) * n + (r)]
    for (int i = 0; i < n; i++) {
        double dot = 0.0;
        for (int k = 0; k < i; k++) dot += UU(k, i) * UU(k, i);
        double tmp = AA(i, i) - dot;
        if (tmp <= 0.0 || isnan(tmp)) return i + 1;
        UU(i, i) = sqrt(tmp);
        for (int j = i + 1; j < n; j++) {
            double d2 = 0.0;
            for (int k = 0; k < i; k++) d2 += UU(k, j) * UU(k, i);
            UU(i, j) = (AA(j, i) - d2) / UU(i, i);
        }
    }
#undef UU
#undef AA
    return 0;
}

/* choleskyPSD with its LAPACK-style helpers, soda.f90:49-124, 128-376.
 * A (n x n column-major) is overwritten as the reference does; piv 1-based. */
int orc_cholesky_psd(int n, double *A, int *piv, int *rank)
{
#define AP(r, c) A[(int64_t)(c) * n + (r)]
    if (n <= 0) return -1;
    double *work = calloc((size_t)2 * n, sizeof(double));
    for (int p = 0; p < n; p++) piv[p] = p + 1;
    /* IDMAX over the diagonal: first index of the maximum */
    int pvt = 0;
    double r = AP(0, 0);
    for (int i = 1; i < n; i++)
        if (AP(i, i) > r) {
            r = AP(i, i);
            pvt = i;
        }
    double Ajj = AP(pvt, pvt);
    if (Ajj <= DEPS) {
        *rank = 0;
        free(work);
        return -1;
    }
    double DSTOP = n * DEPS * Ajj;
    for (int j = 0; j < n; j++) {
        for (int p = j; p < n; p++) {
            if (j > 0) work[p] = work[p] + AP(j - 1, p) * AP(j - 1, p);
            work[n + p] = AP(p, p) - work[p];
        }
        if (j > 0) {
            int it = 0;
            double rr = work[n + j];
            for (int i = 1; i < n - j; i++)
                if (work[n + j + i] > rr) {
                    rr = work[n + j + i];
                    it = i;
                }
            pvt = it + j;
            Ajj = work[n + pvt];
            if (Ajj <= DSTOP) {
                AP(j, j) = Ajj;
                *rank = j;
                free(work);
                return 0;
            }
        }
        if (j != pvt) {
            double t;
            AP(pvt, pvt) = AP(j, j);
            for (int i = 0; i < j; i++) { /* dswap(j-1, A(1,j),1, A(1,pvt),1) */
                t = AP(i, j);
                AP(i, j) = AP(i, pvt);
                AP(i, pvt) = t;
            }
            for (int i = 0; i < pvt - j - 1; i++) { /* dswap(pvt-j-1, A(j,j+1),n, A(j+1,pvt),1) */
                t = AP(j, j + 1 + i);
                AP(j, j + 1 + i) = AP(j + 1 + i, pvt);
                AP(j + 1 + i, pvt) = t;
            }
            for (int i = 0; i < n - pvt - 1; i++) { /* dswap(n-pvt, A(j,pvt+1),n, A(pvt,pvt+1),n) */
                t = AP(j, pvt + 1 + i);
                AP(j, pvt + 1 + i) = AP(pvt, pvt + 1 + i);
                AP(pvt, pvt + 1 + i) = t;
            }
            t = work[j];
            work[j] = work[pvt];
            work[pvt] = t;
            int ti = piv[pvt];
            piv[pvt] = piv[j];
            piv[j] = ti;
        }
        Ajj = sqrt(Ajj);
        AP(j, j) = Ajj;
        if (j < n - 1) {
            /* DGEMV('T', j-1, n-j, -1, A(1,j+1), n, A(1,j), 1, 1, A(j,j+1), n) */
            if (j > 0)
                for (int c = j + 1; c < n; c++) {
                    double temp = 0.0;
                    for (int i = 0; i < j; i++) temp += AP(i, c) * AP(i, j);
                    AP(j, c) = AP(j, c) + (-1.0) * temp;
                }
            /* DSCAL(n-j, 1.0/Ajj, A(j,j+1), n) */
            double da = 1.0 / Ajj;
            for (int c = j + 1; c < n; c++) AP(j, c) = da * AP(j, c);
        }
    }
    *rank = n;
    free(work);
    return 0;
#undef AP
}

/* permutate, utilities.f90:588-607 (p 1-based SNP ids) */
void orc_permutate(int *p, int n, int s, orc_rng *rng)
{
    double *u = malloc(sizeof(double) * (size_t)s);
    for (int i = 0; i < n; i++) p[i] = i + 1;
    for (int i = 1; i <= n; i += s) {
        int m = (n - i + 1) < s ? (n - i + 1) : s;
        orc_uniform_array(rng, s, u);
        for (int j = 1; j <= m; j++) {
            int ipj = i + j - 1;
            int k = (int)(u[j - 1] * (n - ipj + 1)) + ipj;
            int t = p[ipj - 1];
            p[ipj - 1] = p[k - 1];
            p[k - 1] = t;
        }
    }
    free(u);
}

/* ----------------------------------------------------------------- dump -- */

static void put_d(FILE *f, const char *name, const double *a, int64_t n)
{
    char nm[16];
    memset(nm, ' ', 16);
    memcpy(nm, name, strlen(name) < 16 ? strlen(name) : 16);
    int64_t hdr[2] = {1, n};
    fwrite(nm, 1, 16, f);
    fwrite(hdr, 8, 2, f);
    fwrite(a, 8, (size_t)n, f);
}

static void put_s(FILE *f, const char *name, double x) { put_d(f, name, &x, 1); }

static void put_i(FILE *f, const char *name, const int *a, int64_t n)
{
    char nm[16];
    memset(nm, ' ', 16);
    memcpy(nm, name, strlen(name) < 16 ? strlen(name) : 16);
    int64_t hdr[2] = {2, n};
    fwrite(nm, 1, 16, f);
    fwrite(hdr, 8, 2, f);
    for (int64_t i = 0; i < n; i++) {
        int64_t v = a[i];
        fwrite(&v, 8, 1, f);
    }
}

static void put_n(FILE *f, const char *name, int x) { put_i(f, name, &x, 1); }

/* ------------------------------------------------------------ whole job -- */

enum { BAYES_A = 1, BAYES_B, BAYES_C, BAYES_CPI, BAYES_R, BAYES_RC };

typedef struct {
    int nind, nloci, nmix, ngroup, method;
    int blocksize, nblocks, blockt, naind, naind_alloc, nskipped, naless;
    double *X;       /* nind x nloci standardized, column-major */
    double *freq;
    double *XA;      /* naind_alloc x nloci */
    double *xxd;
    int *permvec;    /* 1-based */
    int *skip_j;
    int *i_skipped;  /* 1-based SNP ids */
} orc_soda;

/* block bounds, soda.f90:463-469 (repeated at models.f90:136-142) */
static void block_bounds(const orc_soda *S, int b /*1-based*/, int *ls, int *le)
{
    if (b <= S->blockt) {
        *ls = (b - 1) * S->blocksize + 1;
        int e = *ls + S->blocksize - 1;
        *le = e < S->nloci ? e : S->nloci;
    } else {
        *ls = S->blocksize * S->blockt + (b - S->blockt - 1) * (S->blocksize - 1) + 1;
        int e = *ls + S->blocksize - 2;
        *le = e < S->nloci ? e : S->nloci;
    }
}

/* setup_soda + get_aug_matrix, soda.f90:403-518 */
static int setup_soda(const orc_opts *o, const uint8_t *bed, orc_soda *S, orc_rng *rng0)
{
    int nloci = S->nloci;
    if (o->blocksize > 0) {
        S->blocksize = o->blocksize;
        S->nblocks = (int)ceil((double)nloci / S->blocksize);
        S->blockt = S->nblocks;
    } else if (o->nblocks > 0) {
        S->nblocks = o->nblocks;
        S->blocksize = (int)ceil((double)nloci / S->nblocks);
        S->blockt = nloci - (S->blocksize - 1) * S->nblocks;
    } else {
        S->blocksize = o->nthreads;
        S->nblocks = (int)ceil((double)nloci / S->blocksize);
        S->blockt = S->nblocks;
    }
    if (S->blocksize <= 1) return -2; /* SODA switched off (soda.f90:418-421) */
    S->naind = S->blocksize;
    S->naind_alloc = S->blocksize;
    S->xxd = calloc((size_t)S->nblocks, sizeof(double));
    S->XA = calloc((size_t)S->naind_alloc * nloci, sizeof(double));
    S->skip_j = calloc((size_t)nloci, sizeof(int));
    S->i_skipped = calloc((size_t)nloci, sizeof(int));
    S->permvec = malloc(sizeof(int) * (size_t)nloci);
    if (o->rblocks)
        orc_permutate(S->permvec, nloci, S->blocksize, rng0);
    else
        for (int k = 0; k < nloci; k++) S->permvec[k] = k + 1;
    int64_t nbatch = orc_nbatch(S->nind);
    int64_t *genbits = malloc(sizeof(int64_t) * (size_t)(nbatch * nloci));
    orc_pack_genbits(bed, S->nind, nloci, genbits);

    int bs = S->blocksize;
    double *xx = malloc(sizeof(double) * (size_t)bs * bs);
    double *U = malloc(sizeof(double) * (size_t)bs * bs);
    int *piv = malloc(sizeof(int) * (size_t)bs);
    S->nskipped = 0;
    S->naless = nloci;
    for (int b = 1; b <= S->nblocks; b++) {
        int ls, le;
        block_bounds(S, b, &ls, &le);
        int n = le - ls + 1;
        /* xxtemp(1:n,1:n) as an n x n matrix */
        for (int j = 1; j <= n; j++) {
            int jj = S->permvec[ls + j - 2];
            for (int k = 1; k <= j; k++) {
                int kk = S->permvec[ls + k - 2];
                double v = orc_xxproduct(genbits + (int64_t)(jj - 1) * nbatch,
                                         genbits + (int64_t)(kk - 1) * nbatch, nbatch,
                                         S->nind, S->freq[jj - 1], S->freq[kk - 1]);
                xx[(int64_t)(j - 1) * n + (k - 1)] = v;
                xx[(int64_t)(k - 1) * n + (j - 1)] = v;
            }
        }
        double d;
        orc_power_method(n, xx, rng0, &d);
        d = d + D_EPS_ADD;
        /* test hook (no counterpart in the reference): a lowered d makes
         * dI - G indefinite, so the pivoted-Cholesky fallback and the
         * skipped-SNP updates run; the golden driver applies the same shift */
        if (b - 1 == o->dshift_block) d = d - o->dshift;
        S->xxd[b - 1] = d;
        for (int64_t k = 0; k < (int64_t)n * n; k++) xx[k] = -xx[k];
        for (int j = 0; j < n; j++) xx[(int64_t)j * n + j] = d + xx[(int64_t)j * n + j];
        int status = orc_cholesky_pd(n, xx, U);
        /* XA(1:n, ls:le) = U (copy-out of the array section) */
        for (int c = 0; c < n; c++)
            for (int r = 0; r < n; r++)
                S->XA[(int64_t)(ls - 1 + c) * S->naind_alloc + r] = U[(int64_t)c * n + r];
        if (status != 0) {
            int rank;
            int st2 = orc_cholesky_psd(n, xx, piv, &rank);
            if (st2 != 0) {
                free(xx); free(U); free(piv); free(genbits);
                return -3; /* soda.f90:496-500 */
            }
            for (int j = 1; j <= n; j++) {
                int jj = piv[j - 1];
                if (j <= rank) {
                    for (int r = 1; r <= j; r++)
                        S->XA[(int64_t)(ls + jj - 2) * S->naind_alloc + (r - 1)] =
                            xx[(int64_t)(j - 1) * n + (r - 1)];
                } else {
                    S->skip_j[ls + jj - 2] = 1;
                    S->nskipped++;
                    S->i_skipped[S->nskipped - 1] = S->permvec[ls + jj - 2];
                }
            }
        } else {
            S->naless = 0;
        }
        int cand = bs - n + S->nskipped;
        if (cand < S->naless) S->naless = cand;
    }
    S->naind = S->naind - S->naless;
    free(xx);
    free(U);
    free(piv);
    free(genbits);
    return 0;
}

/* conditional class probabilities, models.f90:369-390 */
static void class_probs(int nmix, const double *s, double *stemp)
{
    for (int kk = 0; kk < nmix; kk++) {
        double skk = s[kk], sk = 0.0;
        int overflow = 0;
        for (int j = 0; j < nmix; j++) {
            if (j == kk) continue;
            double clike = s[j] - skk;
            if (clike < -700) continue;
            if (clike > 700) {
                overflow = 1;
                break;
            }
            sk = sk + exp(clike);
        }
        stemp[kk] = overflow ? 0.0 : 1.0 / (1.0 + sk);
    }
}

static double dotv(const double *a, const double *b, int n)
{
    double s = 0.0;
    for (int i = 0; i < n; i++) s += a[i] * b[i];
    return s;
}

/* compute_residuals, utilities.f90:609-614 (row-wise dot over SNP ids) */
static void compute_residuals(const orc_soda *S, const double *why, const double *g, double mu,
                              double *yadj)
{
    for (int i = 0; i < S->nind; i++) {
        double s = 0.0;
        for (int j = 0; j < S->nloci; j++) s += S->X[(int64_t)j * S->nind + i] * g[j];
        yadj[i] = why[i] - s - mu;
    }
}

int orc_run(const orc_opts *o, const uint8_t *bed, const double *why, const int32_t *grp,
            const char *dump_path)
{
    orc_soda S;
    memset(&S, 0, sizeof S);
    const int n = o->nind, p = o->nloci, method = o->method;
    S.nind = n;
    S.nloci = p;
    S.method = method;
    int nmix = method == BAYES_A ? 1 : (method <= BAYES_CPI ? 2 : o->nmix);
    int ngroup = method == BAYES_RC ? o->ngroup : 1;
    S.nmix = nmix;
    S.ngroup = ngroup;
    if (nmix > 20 || ngroup < 1) return -1;
    double mix[20], alpha[20];
    for (int k = 0; k < 20; k++) {
        mix[k] = o->mix[k];
        alpha[k] = o->alpha[k];
    }
    if (method == BAYES_C || method == BAYES_CPI) { /* useroptions.f90:503-507 */
        mix[0] = 0.0;
        mix[1] = 1.0;
    }

    S.freq = malloc(sizeof(double) * (size_t)p);
    S.X = malloc(sizeof(double) * (size_t)n * p);
    orc_freq(bed, n, p, S.freq);
    orc_standardize(bed, n, p, S.freq, S.X);

    int nstreams = o->maxthreads > 0 ? o->maxthreads : 1;
    orc_rng *streams = malloc(sizeof(orc_rng) * (size_t)nstreams);
    orc_rng_seed(o->seed, nstreams, streams);
    orc_rng *r0 = &streams[0];

    int st = setup_soda(o, bed, &S, r0);
    if (st != 0) return st;

    FILE *f = fopen(dump_path, "wb");
    if (!f) return -4;
    put_n(f, "nind", n);
    put_n(f, "nloci", p);
    put_n(f, "nblocks", S.nblocks);
    put_n(f, "blocksize", S.blocksize);
    put_n(f, "blockt", S.blockt);
    put_n(f, "naind", S.naind);
    put_n(f, "nskipped", S.nskipped);
    put_d(f, "freq", S.freq, p);
    put_i(f, "permvec", S.permvec, p);
    put_d(f, "xxd", S.xxd, S.nblocks);
    put_d(f, "XA", S.XA, (int64_t)S.naind_alloc * p);
    put_i(f, "skip_j", S.skip_j, p);
    if (S.nskipped > 0) put_i(f, "i_skipped", S.i_skipped, S.nskipped);

    /* ------------------------------------------- soda_bayes init 56-128 -- */
    const double nnind = (double)n;
    double *g = calloc((size_t)p, sizeof(double));
    double *gstore = calloc((size_t)p, sizeof(double));
    double *yadj = malloc(sizeof(double) * (size_t)n);
    double *priv = malloc(sizeof(double) * (size_t)n);
    double *yAadj = calloc((size_t)S.naind_alloc, sizeof(double));
    double *privA = malloc(sizeof(double) * (size_t)S.naind_alloc);
    double *xpx = calloc((size_t)p, sizeof(double));
    int *trk1 = malloc(sizeof(int) * (size_t)p); /* snptracker(:,1) */
    int *trk2 = malloc(sizeof(int) * (size_t)p); /* snptracker(:,2) */
    double *vara_s = calloc((size_t)p, sizeof(double));
    double *vara_sstore = calloc((size_t)p, sizeof(double));
    double *indiststore = calloc((size_t)p * nmix, sizeof(double));
    double P[100 * 20], log_p[100 * 20], snpstore[100 * 20], pstore[100 * 20];
    double varindist[100 * 20], varstore[100 * 20];
    int snpindist[100 * 20];
    double gp[20], vare_gp[20], s[20], stemp[20];
    double *gtemp = calloc((size_t)S.blocksize, sizeof(double));
    int *gflag = calloc((size_t)S.blocksize, sizeof(int));
#define IDX(i, j) ((j) * ngroup + (i)) /* (group, class) column-major */
    memset(snpstore, 0, sizeof snpstore);
    memset(pstore, 0, sizeof pstore);
    memset(varindist, 0, sizeof varindist);
    memset(varstore, 0, sizeof varstore);
    memset(snpindist, 0, sizeof snpindist);
    memset(P, 0, sizeof P);
    double included_store = 0.0, included = 0.0;
    double vara = o->vara0, vare, mu, mu_store = 0.0, vare_store = 0.0, vara_store = 0.0;
    double scale_va = o->scale_va;
    const double scale_ve = o->scale_ve, df_va = o->df_va, df_ve = o->df_ve;

    if (method != BAYES_A) {
        for (int j = 0; j < p; j++) {
            trk1[j] = (method == BAYES_RC) ? grp[j] : 1;
            trk2[j] = 2;
        }
        if (method == BAYES_B || method == BAYES_C) {
            for (int i = 0; i < ngroup; i++) {
                P[IDX(i, 0)] = o->pizero;
                P[IDX(i, 1)] = 1.0 - o->pizero;
            }
        } else {
            for (int i = 0; i < ngroup; i++) {
                P[IDX(i, 0)] = 0.5;
                double sum = 0.0;
                for (int k = 1; k < nmix; k++) P[IDX(i, k)] = 1.0 / mix[k];
                for (int k = 1; k < nmix; k++) sum += P[IDX(i, k)];
                for (int k = 1; k < nmix; k++) P[IDX(i, k)] = 0.5 * P[IDX(i, k)] / sum;
            }
        }
    }
    for (int k = 0; k < S.nskipped; k++) {
        int j = S.i_skipped[k] - 1;
        xpx[j] = dotv(S.X + (int64_t)j * n, S.X + (int64_t)j * n, n);
    }
    {
        double sw = 0.0;
        for (int i = 0; i < n; i++) sw += why[i];
        mu = sw / nnind;
    }
    compute_residuals(&S, why, g, mu, yadj);
    if (df_va <= 0.0) {
        double sw = 0.0, vary = 0.0;
        for (int i = 0; i < n; i++) sw += why[i];
        double yhat = sw / nnind;
        for (int i = 0; i < n; i++) vary += (why[i] - yhat) * (why[i] - yhat);
        vary = vary / (nnind - 1.0);
        if (method == BAYES_A) {
            vara = scale_va * vary / p;
        } else if (method == BAYES_B) {
            vara = scale_va * vary / p / P[IDX(0, 1)];
        } else {
            for (int i = 0; i < ngroup; i++) {
                double d = 0.0;
                for (int k = 0; k < nmix; k++) d += P[IDX(i, k)] * mix[k];
                vara = vara + d;
            }
            vara = scale_va * vary / p / vara;
        }
        scale_va = 0.0;
    } else {
        vara = scale_va * df_va / (df_va + 2.0);
    }
    if (method == BAYES_A || method == BAYES_B) {
        for (int j = 0; j < p; j++) vara_s[j] = vara;
    } else {
        for (int k = 0; k < nmix; k++) gp[k] = mix[k] * vara;
    }
    double scale = (dotv(yadj, yadj, n) + scale_ve * df_ve) / (nnind + df_ve);
    vare = orc_sinvchi2(r0, nnind + df_ve, scale);
    for (int i = 0; i < S.naind; i++) yAadj[i] = orc_normal(r0, 0.0, sqrt(vare));

    int nrec = o->numit / (o->thin > 0 ? o->thin : 1) + 1, irec = 0;
    double *h_mu = calloc((size_t)nrec, sizeof(double));
    double *h_vara = calloc((size_t)nrec, sizeof(double));
    double *h_vare = calloc((size_t)nrec, sizeof(double));
    int *h_incl = calloc((size_t)nrec, sizeof(int));
    int *h_snpind = calloc((size_t)nrec * ngroup * nmix, sizeof(int));
    double *h_varind = calloc((size_t)nrec * ngroup * nmix, sizeof(double));
    int counter = 0;
    const int NA = S.naind_alloc;

    for (int rep = 1; rep <= o->numit; rep++) {
        /* models.f90:130-132 */
        for (int i = 0; i < n; i++) yadj[i] = yadj[i] + mu;
        {
            double sy = 0.0;
            for (int i = 0; i < n; i++) sy += yadj[i];
            mu = orc_normal(r0, sy / nnind, sqrt(vare / nnind));
        }
        for (int i = 0; i < n; i++) yadj[i] = yadj[i] - mu;

        if (method == BAYES_A) { /* models.f90:133-194 */
            for (int b = 1; b <= S.nblocks; b++) {
                int ls, le;
                block_bounds(&S, b, &ls, &le);
                double zz = S.xxd[b - 1];
                for (int k = ls; k <= le; k++) {
                    if (S.skip_j[k - 1]) continue;
                    int jk = k - ls;
                    int snp = S.permvec[k - 1] - 1;
                    const double *z = S.X + (int64_t)snp * n;
                    const double *za = S.XA + (int64_t)(k - 1) * NA;
                    double gk = g[snp];
                    double rhs = dotv(yadj, z, n) + dotv(yAadj, za, NA);
                    rhs = rhs + zz * gk;
                    gtemp[jk] = gk;
                    double v1 = zz + vare / vara_s[snp];
                    gk = orc_normal(r0, rhs / v1, sqrt(vare / v1));
                    g[snp] = gk;
                    double sc = (gk * gk + scale_va * df_va) / (df_va + 1.0);
                    vara_s[snp] = orc_sinvchi2(r0, df_va + 1.0, sc);
                }
                memset(priv, 0, sizeof(double) * (size_t)n);
                memset(privA, 0, sizeof(double) * (size_t)NA);
                for (int k = ls; k <= le; k++) {
                    if (S.skip_j[k - 1]) continue;
                    int jk = k - ls;
                    int snp = S.permvec[k - 1] - 1;
                    const double *z = S.X + (int64_t)snp * n;
                    const double *za = S.XA + (int64_t)(k - 1) * NA;
                    double gdiff = gtemp[jk] - g[snp];
                    for (int i = 0; i < n; i++) priv[i] = priv[i] + z[i] * gdiff;
                    for (int i = 0; i < NA; i++) privA[i] = privA[i] + za[i] * gdiff;
                }
                for (int i = 0; i < n; i++) yadj[i] = yadj[i] + priv[i];
                for (int i = 0; i < NA; i++) yAadj[i] = yAadj[i] + privA[i];
            }
            for (int k = 0; k < S.nskipped; k++) {
                int snp = S.i_skipped[k] - 1;
                const double *z = S.X + (int64_t)snp * n;
                double zz = xpx[snp];
                double gk = g[snp];
                double rhs = dotv(yadj, z, n);
                rhs = rhs + zz * gk;
                double v1 = zz + vare / vara_s[snp];
                double gt = orc_normal(r0, rhs / v1, sqrt(vare / v1));
                gk = gk - gt;
                for (int i = 0; i < n; i++) yadj[i] = yadj[i] + z[i] * gk;
                g[snp] = gt;
                double sc = (gt * gt + scale_va * df_va) / (df_va + 1.0);
                vara_s[snp] = orc_sinvchi2(r0, df_va + 1.0, sc);
            }
            varindist[0] = dotv(g, g, p);
        } else if (method == BAYES_B) { /* models.f90:195-325 */
            log_p[0] = log(P[IDX(0, 0)]);
            log_p[1] = log(P[IDX(0, 1)]);
            for (int b = 1; b <= S.nblocks; b++) {
                int ls, le;
                block_bounds(&S, b, &ls, &le);
                double zz = S.xxd[b - 1];
                double zz_vare = zz / vare;
                for (int k = ls; k <= le; k++) {
                    if (S.skip_j[k - 1]) continue;
                    int jk = k - ls;
                    int snp = S.permvec[k - 1] - 1;
                    const double *z = S.X + (int64_t)snp * n;
                    const double *za = S.XA + (int64_t)(k - 1) * NA;
                    double gk = g[snp];
                    double rhs = dotv(yadj, z, n) + dotv(yAadj, za, NA);
                    if (trk2[snp] > 1) rhs = rhs + zz * gk;
                    s[0] = log_p[0];
                    double logdetV = log(vara_s[snp] * zz_vare + 1.0);
                    double uhat = rhs / (zz + vare / vara_s[snp]);
                    s[1] = -0.5 * (logdetV - (rhs * uhat / vare)) + log_p[1];
                    double clike = s[0] - s[1], sk;
                    if (clike > 700)
                        sk = 0.0;
                    else if (clike < -700)
                        sk = 1.0;
                    else
                        sk = 1.0 / (1.0 + exp(clike));
                    double r = orc_uniform(r0);
                    int flag = (r < sk) ? 2 : 1;
                    gtemp[jk] = gk;
                    if (flag == 1) {
                        gk = orc_normal(r0, 0.0, sqrt(vara_s[snp]));
                        g[snp] = 0.0;
                        if (trk2[snp] == 1) gflag[jk] = 1;
                    } else {
                        double v1 = zz + vare / vara_s[snp];
                        gk = orc_normal(r0, rhs / v1, sqrt(vare / v1));
                        g[snp] = gk;
                    }
                    trk2[snp] = flag;
                    double sc = (gk * gk + scale_va * df_va) / (df_va + 1.0);
                    vara_s[snp] = orc_sinvchi2(r0, df_va + 1.0, sc);
                }
                memset(priv, 0, sizeof(double) * (size_t)n);
                memset(privA, 0, sizeof(double) * (size_t)NA);
                for (int k = ls; k <= le; k++) {
                    if (S.skip_j[k - 1]) continue;
                    int jk = k - ls;
                    if (gflag[jk] == 1) {
                        gflag[jk] = 0;
                        continue;
                    }
                    int snp = S.permvec[k - 1] - 1;
                    const double *z = S.X + (int64_t)snp * n;
                    const double *za = S.XA + (int64_t)(k - 1) * NA;
                    double gdiff = gtemp[jk] - g[snp];
                    for (int i = 0; i < n; i++) priv[i] = priv[i] + z[i] * gdiff;
                    for (int i = 0; i < NA; i++) privA[i] = privA[i] + za[i] * gdiff;
                }
                for (int i = 0; i < n; i++) yadj[i] = yadj[i] + priv[i];
                for (int i = 0; i < NA; i++) yAadj[i] = yAadj[i] + privA[i];
            }
            /* skipped SNPs (models.f90:276-320); the reference reads an
             * unset log_p here -- the oracle uses the block's log_p. */
            for (int k = 0; k < S.nskipped; k++) {
                int snp = S.i_skipped[k] - 1;
                const double *z = S.X + (int64_t)snp * n;
                double zz = xpx[snp], zz_vare = zz / vare, gk = g[snp];
                double rhs = dotv(yadj, z, n);
                if (trk2[snp] > 1) rhs = rhs + zz * gk;
                s[0] = log_p[0];
                double logdetV = log(vara_s[snp] * zz_vare + 1.0);
                double uhat = rhs / (zz + vare / vara_s[snp]);
                s[1] = -0.5 * (logdetV - (rhs * uhat / vare)) + log_p[1];
                double clike = s[0] - s[1], sk;
                if (clike > 700)
                    sk = 0.0;
                else if (clike < -700)
                    sk = 1.0;
                else
                    sk = 1.0 / (1.0 + exp(clike));
                double r = orc_uniform(r0);
                int flag = (r < sk) ? 2 : 1;
                double gt;
                if (flag == 1) {
                    gt = orc_normal(r0, 0.0, sqrt(vara_s[snp]));
                    if (trk2[snp] > 1)
                        for (int i = 0; i < n; i++) yadj[i] = yadj[i] + z[i] * gk;
                    g[snp] = 0.0;
                } else {
                    double v1 = zz + vare / vara_s[snp];
                    gt = orc_normal(r0, rhs / v1, sqrt(vare / v1));
                    gk = gk - gt;
                    for (int i = 0; i < n; i++) yadj[i] = yadj[i] + z[i] * gk;
                    g[snp] = gt;
                }
                trk2[snp] = flag;
                double sc = (gt * gt + scale_va * df_va) / (df_va + 1.0);
                vara_s[snp] = orc_sinvchi2(r0, df_va + 1.0, sc);
            }
            int inc = 0;
            double v2 = 0.0;
            for (int j = 0; j < p; j++)
                if (trk2[j] == 2) {
                    inc++;
                    v2 += g[j] * g[j];
                }
            included = inc;
            snpindist[0] = p - inc;
            snpindist[1] = inc;
            varindist[0] = 0.0;
            varindist[1] = v2;
        } else { /* BayesC, Cpi, R, Rc: models.f90:326-513 */
            for (int i = 0; i < ngroup; i++)
                for (int j = 0; j < nmix; j++) log_p[IDX(i, j)] = log(P[IDX(i, j)]);
            for (int k = 1; k < nmix; k++) vare_gp[k] = vare / gp[k];
            for (int b = 1; b <= S.nblocks; b++) {
                int ls, le;
                block_bounds(&S, b, &ls, &le);
                double zz = S.xxd[b - 1];
                double zz_vare = zz / vare;
                for (int k = ls; k <= le; k++) {
                    if (S.skip_j[k - 1]) continue;
                    int jk = k - ls;
                    int snp = S.permvec[k - 1] - 1;
                    const double *z = S.X + (int64_t)snp * n;
                    const double *za = S.XA + (int64_t)(k - 1) * NA;
                    double gk = g[snp];
                    double rhs = dotv(yadj, z, n) + dotv(yAadj, za, NA);
                    if (trk2[snp] > 1) rhs = rhs + zz * gk;
                    int sg = trk1[snp] - 1;
                    s[0] = log_p[IDX(sg, 0)];
                    for (int kk = 1; kk < nmix; kk++) {
                        double logdetV = log(gp[kk] * zz_vare + 1.0);
                        double uhat = rhs / (zz + vare_gp[kk]);
                        s[kk] = -0.5 * (logdetV - (rhs * uhat / vare)) + log_p[IDX(sg, kk)];
                    }
                    class_probs(nmix, s, stemp);
                    double ssculm = 0.0, r = orc_uniform(r0);
                    int flag = 1;
                    for (int kk = 0; kk < nmix; kk++) {
                        ssculm = ssculm + stemp[kk];
                        if (r < ssculm) {
                            flag = kk + 1;
                            break;
                        }
                    }
                    gtemp[jk] = gk;
                    if (flag == 1) {
                        if (trk2[snp] == 1) gflag[jk] = 1;
                        gk = 0.0;
                    } else {
                        double v1 = zz + vare / gp[flag - 1];
                        gk = orc_normal(r0, rhs / v1, sqrt(vare / v1));
                    }
                    trk2[snp] = flag;
                    g[snp] = gk;
                }
                memset(priv, 0, sizeof(double) * (size_t)n);
                memset(privA, 0, sizeof(double) * (size_t)NA);
                for (int k = ls; k <= le; k++) {
                    if (S.skip_j[k - 1]) continue;
                    int jk = k - ls;
                    if (gflag[jk] == 1) {
                        gflag[jk] = 0;
                        continue;
                    }
                    int snp = S.permvec[k - 1] - 1;
                    const double *z = S.X + (int64_t)snp * n;
                    const double *za = S.XA + (int64_t)(k - 1) * NA;
                    double gdiff = gtemp[jk] - g[snp];
                    for (int i = 0; i < n; i++) priv[i] = priv[i] + z[i] * gdiff;
                    for (int i = 0; i < NA; i++) privA[i] = privA[i] + za[i] * gdiff;
                }
                for (int i = 0; i < n; i++) yadj[i] = yadj[i] + priv[i];
                for (int i = 0; i < NA; i++) yAadj[i] = yAadj[i] + privA[i];
            }
            /* skipped SNPs (models.f90:433-495); the reference reads unset
             * log_p / vare_gp here -- the oracle uses the block's values. */
            for (int k = 0; k < S.nskipped; k++) {
                int snp = S.i_skipped[k] - 1;
                const double *z = S.X + (int64_t)snp * n;
                double zz = xpx[snp], zz_vare = zz / vare, gk = g[snp];
                double rhs = dotv(yadj, z, n);
                if (trk2[snp] > 1) rhs = rhs + zz * gk;
                int sg = trk1[snp] - 1;
                s[0] = log_p[IDX(sg, 0)];
                for (int kk = 1; kk < nmix; kk++) {
                    double logdetV = log(gp[kk] * zz_vare + 1.0);
                    double uhat = rhs / (zz + vare_gp[kk]);
                    s[kk] = -0.5 * (logdetV - (rhs * uhat / vare)) + log_p[IDX(sg, kk)];
                }
                class_probs(nmix, s, stemp);
                double ssculm = 0.0, r = orc_uniform(r0), gt;
                int flag = 1;
                for (int kk = 0; kk < nmix; kk++) {
                    ssculm = ssculm + stemp[kk];
                    if (r < ssculm) {
                        flag = kk + 1;
                        break;
                    }
                }
                if (flag == 1) {
                    gt = 0.0;
                    if (trk2[snp] > 1)
                        for (int i = 0; i < n; i++) yadj[i] = yadj[i] + z[i] * gk;
                } else {
                    double v1 = zz + vare / gp[flag - 1];
                    gt = orc_normal(r0, rhs / v1, sqrt(vare / v1));
                    gk = gk - gt;
                    for (int i = 0; i < n; i++) yadj[i] = yadj[i] + z[i] * gk;
                }
                g[snp] = gt;
                trk2[snp] = flag;
            }
            int excl = 0;
            for (int i = 0; i < ngroup; i++) {
                for (int j = 0; j < nmix; j++) {
                    int cnt = 0;
                    double vs = 0.0;
                    for (int q = 0; q < p; q++)
                        if (trk1[q] == i + 1 && trk2[q] == j + 1) {
                            cnt++;
                            vs += g[q] * g[q];
                        }
                    snpindist[IDX(i, j)] = cnt;
                    varindist[IDX(i, j)] = vs;
                }
                excl += snpindist[IDX(i, 0)];
            }
            included = p - excl;
            double sc = 0.0;
            for (int i = 0; i < ngroup; i++)
                for (int j = 1; j < nmix; j++) sc = sc + varindist[IDX(i, j)] / mix[j];
            sc = (sc + scale_va * df_va) / (df_va + included);
            vara = orc_sinvchi2(r0, included + df_va, sc);
            for (int k = 0; k < nmix; k++) gp[k] = mix[k] * vara;
        }
        /* models.f90:515-525 */
        scale = (dotv(yadj, yadj, n) + scale_ve * df_ve) / (nnind + df_ve);
        vare = orc_sinvchi2(r0, nnind + df_ve, scale);
        if (method != BAYES_A && method != BAYES_B && method != BAYES_C) {
            for (int i = 0; i < ngroup; i++) {
                double dirx[20], pr[20];
                for (int j = 0; j < nmix; j++) dirx[j] = (double)snpindist[IDX(i, j)] + alpha[j];
                orc_dirichlet(r0, nmix, dirx, pr);
                for (int j = 0; j < nmix; j++) P[IDX(i, j)] = pr[j];
            }
        }
        for (int i = 0; i < S.naind; i++) yAadj[i] = orc_normal(r0, 0.0, sqrt(vare));
        /* accumulate, models.f90:526-548 */
        if (rep > o->burnin && rep % o->thin == 0) {
            counter++;
            for (int j = 0; j < p; j++) gstore[j] += g[j];
            mu_store += mu;
            for (int k = 0; k < ngroup * nmix; k++) varstore[k] += varindist[k];
            vare_store += vare;
            if (method != BAYES_A) {
                included_store += included;
                for (int k = 0; k < ngroup * nmix; k++) snpstore[k] += snpindist[k];
                for (int j = 0; j < p; j++) indiststore[(int64_t)(trk2[j] - 1) * p + j] += 1.0;
                if (method != BAYES_B && method != BAYES_C)
                    for (int k = 0; k < ngroup * nmix; k++) pstore[k] += P[k];
            }
            if (method == BAYES_A || method == BAYES_B)
                for (int j = 0; j < p; j++) vara_sstore[j] += vara_s[j];
            else
                vara_store += vara;
        }
        if (rep % o->thin == 0 && irec < nrec) {
            h_mu[irec] = mu;
            h_vara[irec] = vara;
            h_vare[irec] = vare;
            h_incl[irec] = (int)included;
            for (int k = 0; k < ngroup * nmix; k++) {
                h_snpind[(int64_t)irec * ngroup * nmix + k] = snpindist[k];
                h_varind[(int64_t)irec * ngroup * nmix + k] = varindist[k];
            }
            irec++;
        }
        if (rep % 1000 == 0) compute_residuals(&S, why, g, mu, yadj);
    }
    /* posterior means, models.f90:576-594 */
    double cnt = counter;
    for (int j = 0; j < p; j++) gstore[j] /= cnt;
    mu_store /= cnt;
    for (int k = 0; k < ngroup * nmix; k++) varstore[k] /= cnt;
    vare_store /= cnt;
    put_s(f, "mu_store", mu_store);
    put_s(f, "vare", vare);
    put_s(f, "vara", vara);
    put_s(f, "vare_store", vare_store);
    put_d(f, "g", g, p);
    put_d(f, "gstore", gstore, p);
    put_d(f, "yadj", yadj, n);
    put_d(f, "yAadj", yAadj, S.naind_alloc);
    put_d(f, "varstore", varstore, (int64_t)ngroup * nmix);
    put_d(f, "varindist", varindist, (int64_t)ngroup * nmix);
    if (method != BAYES_A) {
        int *trk = malloc(sizeof(int) * 2 * (size_t)p);
        memcpy(trk, trk1, sizeof(int) * (size_t)p);
        memcpy(trk + p, trk2, sizeof(int) * (size_t)p);
        put_i(f, "snptracker", trk, 2 * (int64_t)p);
        free(trk);
        put_i(f, "snpindist", snpindist, (int64_t)ngroup * nmix);
        included_store /= cnt;
        for (int k = 0; k < ngroup * nmix; k++) snpstore[k] /= cnt;
        for (int64_t k = 0; k < (int64_t)p * nmix; k++) indiststore[k] /= cnt;
        put_d(f, "snpstore", snpstore, (int64_t)ngroup * nmix);
        put_d(f, "indiststore", indiststore, (int64_t)p * nmix);
        put_s(f, "included_store", included_store);
        put_d(f, "p", P, (int64_t)ngroup * nmix);
    }
    if (method != BAYES_A && method != BAYES_B && method != BAYES_C) {
        for (int k = 0; k < ngroup * nmix; k++) pstore[k] /= cnt;
        put_d(f, "pstore", pstore, (int64_t)ngroup * nmix);
    }
    if (method == BAYES_A || method == BAYES_B) {
        for (int j = 0; j < p; j++) vara_sstore[j] /= cnt;
        put_d(f, "vara_s", vara_s, p);
        put_d(f, "vara_sstore", vara_sstore, p);
    } else {
        vara_store /= cnt;
        put_s(f, "vara_store", vara_store);
    }
    put_d(f, "hyp_mu", h_mu, irec);
    put_d(f, "hyp_vara", h_vara, irec);
    put_d(f, "hyp_vare", h_vare, irec);
    put_i(f, "hyp_included", h_incl, irec);
    put_i(f, "hyp_snpindist", h_snpind, (int64_t)irec * ngroup * nmix);
    put_d(f, "hyp_varindist", h_varind, (int64_t)irec * ngroup * nmix);
    fclose(f);
#undef IDX
    free(g); free(gstore); free(yadj); free(priv); free(yAadj); free(privA); free(xpx);
    free(trk1); free(trk2); free(vara_s); free(vara_sstore); free(indiststore);
    free(gtemp); free(gflag); free(h_mu); free(h_vara); free(h_vare); free(h_incl);
    free(h_snpind); free(h_varind);
    free(S.freq); free(S.X); free(S.XA); free(S.xxd); free(S.permvec); free(S.skip_j);
    free(S.i_skipped); free(streams);
    return 0;
}

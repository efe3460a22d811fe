// Mixed-precision, chain-batched block products of the SODA sweep.
//
// C chains share one pass over a block's genotypes.  Each chain's residual
// r_c is written as a fixed-point integer q = rint(r_c * 2^e_c) with
// |q| < 2^(8L-2), split into L balanced radix-256 digits (int8), and
//   sum_i x_ij r_ic  ~=  2^-e_c * sum_k 256^k * T_jck,   T_jck = sum_i x_ij d_ick
// where every T is an exact int32 sum from int8 MFMA (v_mfma_i32_16x16x32_i8).
// The only error is the rounding of r to 8L-2 bits of its largest entry.
// The augmented part U_s' R_a stays FP64 (LDS-tiled over all chains).
#include <hip/hip_runtime.h>

#include <cmath>

#include "sb_device.h"
#include "sb_draw.h"
#include "sb_internal.h"

namespace sb {

namespace {
constexpr int NT = 256;
typedef int v4i __attribute__((ext_vector_type(4)));

}  // namespace

// per chain, per workgroup: max |r| (as the bit pattern of a non-negative
// double, which orders like the value) and sum r over [0, n)
__global__ __launch_bounds__(NT) void k_rstat(Par P, const double *r, double *rsum, uint64_t *rmax)
{
    __shared__ double lds[NT / WAVE];
    __shared__ unsigned long long mx[NT / WAVE];
    const int c = blockIdx.z;
    r += (int64_t)c * P.ldr;
    double s = 0.0;
    unsigned long long m = 0;
    for (int i = blockIdx.x * NT + threadIdx.x; i < P.n; i += gridDim.x * NT) {
        double v = r[i];
        s += v;
        unsigned long long b = (unsigned long long)__double_as_longlong(fabs(v));
        m = b > m ? b : m;
    }
    s = block_sum<NT>(s, lds);
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        unsigned long long o = __shfl_xor(m, off, WAVE);
        m = o > m ? o : m;
    }
    if ((threadIdx.x & 63) == 0) mx[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < NT / WAVE; w++) m = mx[w] > m ? mx[w] : m;
        rsum[(int64_t)c * gridDim.x + blockIdx.x] = s;
        rmax[(int64_t)c * gridDim.x + blockIdx.x] = m;
    }
}

// Fixed-point exponent of chain c from the k_rstat partials.
__device__ __forceinline__ int chain_exponent(const uint64_t *rmax, int nwg, int c, int L)
{
    unsigned long long m = 0;
    for (int w = 0; w < nwg; w++) m = rmax[(int64_t)c * nwg + w] > m ? rmax[(int64_t)c * nwg + w] : m;
    double mv = __longlong_as_double((long long)m);
    if (mv == 0.0) return 0;
    int E;
    frexp(mv, &E);              // mv = f * 2^E, 0.5 <= f < 1
    return 8 * L - 2 - E;       // |r| * 2^e < 2^(8L-2)
}

// One thread per (16-individual word, chain): balanced radix-256 digits of
// q = rint(r * 2^e), written limb-major as Q[c*L + k][16 w + slot] with the
// slot order {0,4,8,12, 1,5,9,13, 2,6,10,14, 3,7,11,15}: the order in which
// (word >> 2m) & 0x03030303 unpacks a word of 2-bit genotypes.
__global__ __launch_bounds__(NT) void k_quant(Par P, int L, const double *r, const uint64_t *rmax,
                                              int nwg, int8_t *Q, int64_t ldq, int *expo)
{
    const int c = blockIdx.z;
    const int w = blockIdx.x * NT + threadIdx.x;
    const int e = chain_exponent(rmax, nwg, c, L);
    if (w == 0) expo[c] = e;
    if (w >= P.ldw) return;
    const double *rc = r + (int64_t)c * P.ldr + (int64_t)w * 16;
    int32_t q[16];
#pragma unroll
    for (int t = 0; t < 16; t++) q[t] = (int32_t)rint(ldexp(rc[t], e));
    for (int k = 0; k < L; k++) {
        uint32_t out[4];
#pragma unroll
        for (int m = 0; m < 4; m++) {
            uint32_t v = 0;
#pragma unroll
            for (int b = 0; b < 4; b++) {
                int t = 4 * b + m;                     // individual within the word
                int8_t d = (int8_t)(q[t] & 0xFF);      // balanced digit in [-128, 127]
                v |= ((uint32_t)(uint8_t)d) << (8 * b);
            }
            out[m] = v;
        }
#pragma unroll
        for (int t = 0; t < 16; t++) q[t] = (q[t] - (int32_t)(int8_t)(q[t] & 0xFF)) >> 8;
        uint4 *dst = reinterpret_cast<uint4 *>(Q + (int64_t)(c * L + k) * ldq + (int64_t)w * 16);
        *dst = make_uint4(out[0], out[1], out[2], out[3]);
    }
}

// T[jk][col] += sum_i x_(p0+jk, i) * Q[col][i] for the block's SNPs.
// Workgroup = 64 SNPs (16 per wave) x all NT16 columns over one chunk of
// `kwords` genotype words; int32 partial sums are added atomically (exact).
template <int NTILE>
__global__ __launch_bounds__(NT) void k_xq(Par P, const uint32_t *__restrict__ geno, int p0, int nb,
                                           const int8_t *__restrict__ Q, int64_t ldq, int kwords,
                                           int *__restrict__ T, int ldt)
{
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int g = lane >> 4, l16 = lane & 15;
    const int snp = blockIdx.x * 64 + wave * 16 + l16;
    const bool valid = snp < nb;
    const int w0 = blockIdx.y * kwords;
    const int w1 = min(P.ldw, w0 + kwords);
    const uint32_t *grow = geno + (int64_t)(p0 + (valid ? snp : 0)) * P.ldw;
    v4i acc[NTILE];
#pragma unroll
    for (int t = 0; t < NTILE; t++) acc[t] = v4i{0, 0, 0, 0};
    constexpr uint32_t M = 0x03030303u;
    // super-step: 16 words (256 individuals); lane group g owns words 4g..4g+3
    // (all loads of a super-step are issued before its MFMAs)
    for (int ws = w0; ws < w1; ws += 16) {
        uint4 gw = valid ? *reinterpret_cast<const uint4 *>(grow + ws + 4 * g) : make_uint4(0, 0, 0, 0);
        uint4 qb[4][NTILE];
#pragma unroll
        for (int s = 0; s < 4; s++) {
            const int64_t qoff = (int64_t)(ws + 4 * g + s) * 16;
#pragma unroll
            for (int t = 0; t < NTILE; t++)
                qb[s][t] = *reinterpret_cast<const uint4 *>(Q + (int64_t)(16 * t + l16) * ldq + qoff);
        }
        uint32_t words[4] = {gw.x, gw.y, gw.z, gw.w};
#pragma unroll
        for (int s = 0; s < 4; s++) {
            const uint32_t wd = words[s];
            const long a0 = (long)(((uint64_t)((wd >> 2) & M) << 32) | (wd & M));
            const long a1 = (long)(((uint64_t)((wd >> 6) & M) << 32) | ((wd >> 4) & M));
#pragma unroll
            for (int t = 0; t < NTILE; t++) {
                const long b0 = (long)(((uint64_t)qb[s][t].y << 32) | qb[s][t].x);
                const long b1 = (long)(((uint64_t)qb[s][t].w << 32) | qb[s][t].z);
                acc[t] = __builtin_amdgcn_mfma_i32_16x16x32_i8(a0, b0, acc[t], 0, 0, 0);
                acc[t] = __builtin_amdgcn_mfma_i32_16x16x32_i8(a1, b1, acc[t], 0, 0, 0);
            }
        }
    }
    // acc[t][e]: row = SNP blockIdx.x*64 + wave*16 + 4g + e, column 16 t + l16
#pragma unroll
    for (int t = 0; t < NTILE; t++)
#pragma unroll
        for (int e = 0; e < 4; e++) {
            int row = blockIdx.x * 64 + wave * 16 + 4 * g + e;
            if (row < nb && acc[t][e] != 0) atomicAdd(T + (int64_t)row * ldt + 16 * t + l16, acc[t][e]);
        }
}

// Products with a block's packed upper triangle U on FP64 MFMA
// (v_mfma_f64_16x16x4f64), 64 x 64 U tiles staged in LDS:
//   mode 0 (rhs):    part[g][jk][c] = sum_{i in row group g, i <= jk} U(i, jk) R(i, c)
//   mode 1 (update): part[g][i][c]  = sum_{j in col group g, j >= i} U(i, j) R(j, c)
// R is n_b x C (chain stride rstride) -- r_a for mode 0, delta for mode 1.
// Workgroup: 64 outputs (4 waves x 16) x all chains over one group of
// `span` rows (mode 0) or columns (mode 1).
typedef double v4d __attribute__((ext_vector_type(4)));
template <int MODE>
__global__ __launch_bounds__(NT) void k_utile(Par P, const double *__restrict__ U, int64_t tri0, int nb,
                                              const double *__restrict__ R, int64_t rstride, int span,
                                              double *__restrict__ part)
{
    const int C = P.nchains;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int l16 = lane & 15, lg = lane >> 4;
    const int o0 = blockIdx.x * 64 + wave * 16;   // this wave's 16 outputs
    const int g = blockIdx.y;
    int k0 = g * span, k1 = min(nb, k0 + span);
    if (MODE == 0) {
        if (k0 > blockIdx.x * 64 + 63) return;    // rows entirely below the diagonal
        k1 = min(k1, o0 + 16);
    } else {
        if (k1 <= blockIdx.x * 64) return;        // columns entirely left of the diagonal
        k0 = max(k0, o0);
    }
    if (o0 >= nb) return;
    const int ntc = (C + 15) / 16;
    v4d acc[4];
#pragma unroll
    for (int t = 0; t < 4; t++) acc[t] = v4d{0.0, 0.0, 0.0, 0.0};
    // k-step of 16: lane group lg owns k indices kt + 4 lg + s, s = 0..3
    for (int kt = k0 - (k0 & 3); kt < k1; kt += 16) {
        double a[4];
        const int kb = kt + 4 * lg;
        if (MODE == 0) {
            // A[m][k] = U(row kb+s, col o0+m): 4 consecutive rows of column o0+l16
            const int j = o0 + l16;
            const double *col = U + tri0 + (int64_t)j * (j + 1) / 2;
#pragma unroll
            for (int s = 0; s < 4; s++) {
                const int i = kb + s;
                a[s] = (j < nb && i >= k0 && i < k1 && i <= j) ? col[i] : 0.0;
            }
        } else {
            // A[m][k] = U(row o0+m, col kb+s): row o0+l16 of 4 consecutive columns
            const int i = o0 + l16;
#pragma unroll
            for (int s = 0; s < 4; s++) {
                const int j = kb + s;
                a[s] = (i < nb && j >= k0 && j < k1 && i <= j) ? U[tri0 + (int64_t)j * (j + 1) / 2 + i] : 0.0;
            }
        }
        for (int t = 0; t < ntc; t++) {
            const int ch = t * 16 + l16;
            double bv[4];
            const double *rc = R + (int64_t)(ch < C ? ch : 0) * rstride;
#pragma unroll
            for (int s = 0; s < 4; s++) {
                const int k = kb + s;
                bv[s] = (ch < C && k >= k0 && k < k1) ? rc[k] : 0.0;
            }
#pragma unroll
            for (int s = 0; s < 4; s++) acc[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[s], bv[s], acc[t], 0, 0, 0);
        }
    }
    // D[m][n] of the f64 16x16x4 MFMA: m = lg + 4*v (rows interleaved across
    // the 4 lane groups), n = l16 (chain within tile t)
    for (int t = 0; t < ntc; t++)
#pragma unroll
        for (int v = 0; v < 4; v++) {
            int o = o0 + lg + 4 * v;
            int ch = t * 16 + l16;
            if (o < nb && ch < C) part[((int64_t)g * P.blocksize + o) * C + ch] = acc[t][v];
        }
}

// r_a(i, c) += sum of the mode-1 partials of the column groups that reach row
// i's tile (groups left of the diagonal wrote nothing).
__global__ void k_ra_reduce(Par P, const double *part, int span, int ngroups, int nb, double *ra)
{
    const int i = blockIdx.x * NT + threadIdx.x;
    const int c = blockIdx.y;
    if (i >= nb) return;
    const int C = P.nchains;
    const int i0 = (i / 64) * 64;
    double s = 0.0;
    for (int g = i0 / span; g < ngroups; g++) s += part[((int64_t)g * P.blocksize + i) * C + c];
    ra[(int64_t)c * P.blocksize + i] += s;
}

// Genotypes in individual-major form per block for the update GEMM:
// word (block b, SNP word sw, individual i) holds the 2-bit counts of block
// positions 16 sw .. 16 sw + 15 of individual i (position t at bits 2t).
__global__ void k_transpose(const uint32_t *__restrict__ geno, int ldw, int p0, int nb, int64_t ldi,
                            uint32_t *__restrict__ gt)
{
    const int sw = blockIdx.y;
    const int iw = blockIdx.x * blockDim.x + threadIdx.x;  // word of 16 individuals
    if (iw >= ldw) return;
    uint32_t in[16];
#pragma unroll
    for (int t = 0; t < 16; t++) {
        int pos = 16 * sw + t;
        in[t] = pos < nb ? geno[(int64_t)(p0 + pos) * ldw + iw] : 0u;
    }
    uint32_t out[16];
#pragma unroll
    for (int li = 0; li < 16; li++) {
        uint32_t w = 0;
#pragma unroll
        for (int t = 0; t < 16; t++) w |= ((in[t] >> (2 * li)) & 3u) << (2 * t);
        out[li] = w;
    }
    uint4 *dst = reinterpret_cast<uint4 *>(gt + (int64_t)sw * ldi + (int64_t)iw * 16);
#pragma unroll
    for (int q = 0; q < 4; q++) dst[q] = make_uint4(out[4 * q], out[4 * q + 1], out[4 * q + 2], out[4 * q + 3]);
}

// Per chain: c_j = delta_j / s_j over the block, K_c = sum_j 2 f_j c_j, and
// c quantised to L balanced int8 limbs in the MFMA slot order (Dq rows
// c*L + k, 16 slots per 16-SNP word).
__global__ __launch_bounds__(NT) void k_dquant(Par P, int L, const double *delta, const double *f2,
                                               const double *sdv, int p0, int nb, int nsw, int8_t *Dq,
                                               int64_t ldd, int *expo, double *kconst)
{
    __shared__ double lds[NT / WAVE];
    __shared__ unsigned long long mx[NT / WAVE];
    __shared__ int e_s;
    const int c = blockIdx.x;
    const double *dc = delta + (int64_t)c * P.p + p0;
    double kc = 0.0;
    unsigned long long m = 0;
    for (int j = threadIdx.x; j < nb; j += NT) {
        double s = sdv[p0 + j];
        double cj = s > 0.0 ? dc[j] / s : 0.0;
        kc = fma(f2[p0 + j], cj, kc);
        unsigned long long bits = (unsigned long long)__double_as_longlong(fabs(cj));
        m = bits > m ? bits : m;
    }
    kc = block_sum<NT>(kc, lds);
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        unsigned long long o = __shfl_xor(m, off, WAVE);
        m = o > m ? o : m;
    }
    if ((threadIdx.x & 63) == 0) mx[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < NT / WAVE; w++) m = mx[w] > m ? mx[w] : m;
        double mv = __longlong_as_double((long long)m);
        int E = 0;
        if (mv > 0.0) frexp(mv, &E);
        e_s = mv > 0.0 ? 8 * L - 2 - E : 0;
        expo[c] = e_s;
        kconst[c] = kc;
    }
    __syncthreads();
    const int e = e_s;
    for (int j = threadIdx.x; j < nsw * 16; j += NT) {
        int32_t q = 0;
        if (j < nb) {
            double s = sdv[p0 + j];
            double cj = s > 0.0 ? dc[j] / s : 0.0;
            q = (int32_t)rint(ldexp(cj, e));
        }
        const int t = j & 15, slot = 4 * (t & 3) + (t >> 2);
        const int64_t col = (int64_t)(j & ~15) + slot;
        for (int k = 0; k < L; k++) {
            int8_t d = (int8_t)(q & 0xFF);
            Dq[(int64_t)(c * L + k) * ldd + col] = d;
            q = (q - (int32_t)d) >> 8;
        }
    }
}

// r_c += X_s c - K_c for every chain, as the int8 MFMA product of the
// individual-major genotypes with the limbs of c (exact int32 sums), then
// the FP64 limb combination.  Workgroup = 64 individuals (4 waves x 16).
template <int NTILE>
__global__ __launch_bounds__(NT) void k_xupd(Par P, const uint32_t *__restrict__ gt, int64_t ldi, int nsw,
                                             const int8_t *__restrict__ Dq, int64_t ldd, int L,
                                             const int *__restrict__ expo, const double *__restrict__ kconst,
                                             double *__restrict__ r)
{
    __shared__ int Ts[4][16][16 * NTILE + 1];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int g = lane >> 4, l16 = lane & 15;
    const int i0 = blockIdx.x * 64 + wave * 16;
    v4i acc[NTILE];
#pragma unroll
    for (int t = 0; t < NTILE; t++) acc[t] = v4i{0, 0, 0, 0};
    constexpr uint32_t M = 0x03030303u;
    constexpr int PF = 4;  // super-steps fetched ahead of their MFMAs
    for (int s0 = 0; s0 < nsw; s0 += 4 * PF) {
        uint32_t wd[PF];
        uint4 qb[PF][NTILE];
#pragma unroll
        for (int f = 0; f < PF; f++) {
            const int sw = s0 + 4 * f + g;
            wd[f] = sw < nsw ? gt[(int64_t)sw * ldi + i0 + l16] : 0u;
#pragma unroll
            for (int t = 0; t < NTILE; t++)
                qb[f][t] = sw < nsw ? *reinterpret_cast<const uint4 *>(Dq + (int64_t)(16 * t + l16) * ldd + (int64_t)sw * 16)
                                    : make_uint4(0, 0, 0, 0);
        }
#pragma unroll
        for (int f = 0; f < PF; f++) {
            const long a0 = (long)(((uint64_t)((wd[f] >> 2) & M) << 32) | (wd[f] & M));
            const long a1 = (long)(((uint64_t)((wd[f] >> 6) & M) << 32) | ((wd[f] >> 4) & M));
#pragma unroll
            for (int t = 0; t < NTILE; t++) {
                const long b0 = (long)(((uint64_t)qb[f][t].y << 32) | qb[f][t].x);
                const long b1 = (long)(((uint64_t)qb[f][t].w << 32) | qb[f][t].z);
                acc[t] = __builtin_amdgcn_mfma_i32_16x16x32_i8(a0, b0, acc[t], 0, 0, 0);
                acc[t] = __builtin_amdgcn_mfma_i32_16x16x32_i8(a1, b1, acc[t], 0, 0, 0);
            }
        }
    }
    // acc[t][e]: individual i0 + 4g + e, column 16 t + l16
#pragma unroll
    for (int t = 0; t < NTILE; t++)
#pragma unroll
        for (int e = 0; e < 4; e++) Ts[wave][4 * g + e][16 * t + l16] = acc[t][e];
    __syncthreads();
    const int C = P.nchains;
    for (int pr = lane; pr < 16 * C; pr += 64) {
        const int li = pr & 15, c = pr >> 4;
        const int i = i0 + li;
        if (i >= P.n) continue;
        long long S = 0;
        for (int k = L - 1; k >= 0; k--) S = S * 256 + (long long)Ts[wave][li][c * L + k];
        double* rc = r + (int64_t)c * P.ldr;
        rc[i] = rc[i] + (ldexp((double)S, -expo[c]) - kconst[c]);
    }
}

// One thread per (SNP, chain): exact limb combination, rhs, draw
// (Philox, counter (iteration, SNP id, tag)), and clearing of T.
__global__ __launch_bounds__(NT) void k_draw_b(Par P, const Scal *sc, int L, int *T, int ldt,
                                               const int *expo, const double *rsum, int nwg_r,
                                               const double *rapart, int rpg, const double *f2,
                                               const double *sdv, const int32_t *grp,
                                               const int32_t *snp_of_pos, int p0, int nb, double zz,
                                               int rep, double *beta, int32_t *cls, double *vara_s,
                                               double *delta, uint8_t *chg, double *rhs_out)
{
    const int jk = blockIdx.x * NT + threadIdx.x;
    const int c = blockIdx.y;
    if (jk >= nb) return;
    const int C = P.nchains;
    int *trow = T + (int64_t)jk * ldt + c * L;
    long long S = 0;
    for (int k = L - 1; k >= 0; k--) {
        S = S * 256 + (long long)trow[k];
        trow[k] = 0;
    }
    double sx = ldexp((double)S, -expo[c]);
    double sr = 0.0;
    for (int w = 0; w < nwg_r; w++) sr += rsum[(int64_t)c * nwg_r + w];
    const int pos = p0 + jk;
    const double s = sdv[pos];
    double rhs = s > 0.0 ? (sx - f2[pos] * sr) / s : 0.0;
    // row groups entirely below this column's diagonal hold no partial
    for (int g = 0; g <= jk / rpg; g++) rhs += rapart[((int64_t)g * P.blocksize + jk) * C + c];
    if (rhs_out) {  // diagnostics: the products only, no draw
        rhs_out[(int64_t)c * nb + jk] = rhs;
        return;
    }
    const Scal *scc = sc + c;
    PhiloxSrc src(chain_key(P, c), (uint32_t)rep, (uint32_t)snp_of_pos[pos], 2u);
    const int64_t o = (int64_t)c * P.p;
    draw_snp(P, scc, src, rhs, zz, zz / scc->vare, pos, grp[pos], beta + o, cls + o, vara_s + o,
                    delta + o, chg + o);
}

// ----------------------------------------------------------------- host --

int batch_block(sb_handle *h, int b, double zz, int rep)
{
    const Par &P = h->par;
    hipStream_t st = h->stream;
    const int p0 = h->ls[b], nb = h->ls[b + 1] - p0;
    const int C = P.nchains, L = h->limbs;
    const int nwg_r = h->nwg_r;
    k_rstat<<<dim3(nwg_r, 1, C), NT, 0, st>>>(P, h->r, h->b_rsum, h->b_rmax);
    k_quant<<<dim3((P.ldw + NT - 1) / NT, 1, C), NT, 0, st>>>(P, L, h->r, h->b_rmax, nwg_r, h->b_Q,
                                                             h->b_ldq, h->b_expo);
    const int kwords = h->kwords;
    dim3 gx((nb + 63) / 64, (P.ldw + kwords - 1) / kwords);
    switch (h->b_ntile) {
    case 1: k_xq<1><<<gx, NT, 0, st>>>(P, h->geno, p0, nb, h->b_Q, h->b_ldq, kwords, h->b_T, h->b_ldt); break;
    case 2: k_xq<2><<<gx, NT, 0, st>>>(P, h->geno, p0, nb, h->b_Q, h->b_ldq, kwords, h->b_T, h->b_ldt); break;
    case 3: k_xq<3><<<gx, NT, 0, st>>>(P, h->geno, p0, nb, h->b_Q, h->b_ldq, kwords, h->b_T, h->b_ldt); break;
    case 4: k_xq<4><<<gx, NT, 0, st>>>(P, h->geno, p0, nb, h->b_Q, h->b_ldq, kwords, h->b_T, h->b_ldt); break;
    case 6: k_xq<6><<<gx, NT, 0, st>>>(P, h->geno, p0, nb, h->b_Q, h->b_ldq, kwords, h->b_T, h->b_ldt); break;
    case 8: k_xq<8><<<gx, NT, 0, st>>>(P, h->geno, p0, nb, h->b_Q, h->b_ldq, kwords, h->b_T, h->b_ldt); break;
    case 12: k_xq<12><<<gx, NT, 0, st>>>(P, h->geno, p0, nb, h->b_Q, h->b_ldq, kwords, h->b_T, h->b_ldt); break;
    default: return SB_ERR_ARG;
    }
    const int nrg = (nb + h->rows_per_group - 1) / h->rows_per_group;
    k_utile<0><<<dim3((nb + 63) / 64, nrg), NT, 0, st>>>(P, h->U, h->tri_off[b], nb, h->ra, P.blocksize,
                                                        h->rows_per_group, h->b_rapart);
    return hip_check(hipGetLastError(), "batch products");
}

int batch_update(sb_handle *h, int b, const double *delta)
{
    const Par &P = h->par;
    hipStream_t st = h->stream;
    const int p0 = h->ls[b], nb = h->ls[b + 1] - p0;
    const int C = P.nchains, L = h->limbs;
    const int nsw = (nb + 15) / 16;
    k_dquant<<<C, NT, 0, st>>>(P, L, delta, h->f2, h->sdv, p0, nb, nsw, h->b_Dq, h->b_ldd, h->b_expo2,
                               h->b_kc);
    const int64_t ldi = P.ldr;
    const uint32_t *gt = h->geno_t + h->gt_off[b];
    const int nit = (int)((ldi + 63) / 64);
    switch (h->b_ntile) {
    case 1: k_xupd<1><<<nit, NT, 0, st>>>(P, gt, ldi, nsw, h->b_Dq, h->b_ldd, L, h->b_expo2, h->b_kc, h->r); break;
    case 2: k_xupd<2><<<nit, NT, 0, st>>>(P, gt, ldi, nsw, h->b_Dq, h->b_ldd, L, h->b_expo2, h->b_kc, h->r); break;
    case 3: k_xupd<3><<<nit, NT, 0, st>>>(P, gt, ldi, nsw, h->b_Dq, h->b_ldd, L, h->b_expo2, h->b_kc, h->r); break;
    case 4: k_xupd<4><<<nit, NT, 0, st>>>(P, gt, ldi, nsw, h->b_Dq, h->b_ldd, L, h->b_expo2, h->b_kc, h->r); break;
    case 6: k_xupd<6><<<nit, NT, 0, st>>>(P, gt, ldi, nsw, h->b_Dq, h->b_ldd, L, h->b_expo2, h->b_kc, h->r); break;
    case 8: k_xupd<8><<<nit, NT, 0, st>>>(P, gt, ldi, nsw, h->b_Dq, h->b_ldd, L, h->b_expo2, h->b_kc, h->r); break;
    case 12: k_xupd<12><<<nit, NT, 0, st>>>(P, gt, ldi, nsw, h->b_Dq, h->b_ldd, L, h->b_expo2, h->b_kc, h->r); break;
    default: return SB_ERR_ARG;
    }
    const int ncg = (nb + h->rows_per_group - 1) / h->rows_per_group;
    k_utile<1><<<dim3((nb + 63) / 64, ncg), NT, 0, st>>>(P, h->U, h->tri_off[b], nb, delta + p0, P.p,
                                                        h->rows_per_group, h->b_uapart);
    k_ra_reduce<<<dim3((nb + NT - 1) / NT, C), NT, 0, st>>>(P, h->b_uapart, h->rows_per_group, ncg, nb, h->ra);
    return hip_check(hipGetLastError(), "batch update");
}

int batch_transpose(sb_handle *h)
{
    const Par &P = h->par;
    for (int b = 0; b < P.nblocks; b++) {
        const int p0 = h->ls[b], nb = h->ls[b + 1] - p0;
        const int nsw = (nb + 15) / 16;
        k_transpose<<<dim3((P.ldw + NT - 1) / NT, nsw), NT, 0, h->stream>>>(h->geno, P.ldw, p0, nb, P.ldr,
                                                                          h->geno_t + h->gt_off[b]);
    }
    int rc = hip_check(hipGetLastError(), "transpose");
    if (!rc) rc = hip_check(hipStreamSynchronize(h->stream), "transpose");
    return rc;
}

int batch_draw(sb_handle *h, int b, double zz, int rep, double *delta, uint8_t *chg, double *rhs_out)
{
    const Par &P = h->par;
    const int p0 = h->ls[b], nb = h->ls[b + 1] - p0;
    k_draw_b<<<dim3((nb + NT - 1) / NT, P.nchains), NT, 0, h->stream>>>(
        P, h->scal, h->limbs, h->b_T, h->b_ldt, h->b_expo, h->b_rsum, h->nwg_r, h->b_rapart,
        h->rows_per_group, h->f2,
        h->sdv, h->grp, h->d_pos_by_id + P.p, p0, nb, zz, rep, h->beta, h->cls, h->vara_s, delta, chg,
        rhs_out);
    return hip_check(hipGetLastError(), "batch draw");
}

}  // namespace sb

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
    for (int ws = w0; ws < w1; ws += 16) {
        uint4 gw = valid ? *reinterpret_cast<const uint4 *>(grow + ws + 4 * g) : make_uint4(0, 0, 0, 0);
        uint32_t words[4] = {gw.x, gw.y, gw.z, gw.w};
#pragma unroll
        for (int s = 0; s < 4; s++) {
            const uint32_t wd = words[s];
            const long a0 = (long)(((uint64_t)((wd >> 2) & M) << 32) | (wd & M));
            const long a1 = (long)(((uint64_t)((wd >> 6) & M) << 32) | ((wd >> 4) & M));
            const int64_t qoff = (int64_t)(ws + 4 * g + s) * 16;
#pragma unroll
            for (int t = 0; t < NTILE; t++) {
                const uint4 qb = *reinterpret_cast<const uint4 *>(Q + (int64_t)(16 * t + l16) * ldq + qoff);
                const long b0 = (long)(((uint64_t)qb.y << 32) | qb.x);
                const long b1 = (long)(((uint64_t)qb.w << 32) | qb.z);
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

// RA_part[rg][jk][c] = sum over rows i of row group rg (i <= jk) of
// U(i, jk) * ra_c(i): 64 columns x all chains per workgroup, U and R_a
// tiles staged in LDS.
__global__ __launch_bounds__(NT) void k_ua(Par P, const double *__restrict__ U, int64_t tri0, int nb,
                                           const double *__restrict__ ra, int rows_per_group,
                                           double *__restrict__ part)
{
    __shared__ double Ut[64][65];
    __shared__ double Rt[64][65];
    const int C = P.nchains;
    const int col0 = blockIdx.x * 64;
    const int rg = blockIdx.y;
    const int r0 = rg * rows_per_group;
    if (r0 >= nb || r0 > col0 + 63) return;  // group entirely below the diagonal
    const int r1 = min(min(nb, r0 + rows_per_group), col0 + 64);
    const int tc = threadIdx.x & 63, cg = threadIdx.x >> 6;  // column, chain group
    double acc[16];
#pragma unroll
    for (int k = 0; k < 16; k++) acc[k] = 0.0;
    for (int rt = r0; rt < r1; rt += 64) {
        for (int e = threadIdx.x; e < 64 * 64; e += NT) {
            int rr = e & 63, cc = e >> 6;
            int i = rt + rr, j = col0 + cc;
            Ut[rr][cc] = (j < nb && i <= j && i < r1) ? U[tri0 + (int64_t)j * (j + 1) / 2 + i] : 0.0;
        }
        for (int e = threadIdx.x; e < 64 * 64; e += NT) {
            int rr = e & 63, ch = e >> 6;
            int i = rt + rr;
            Rt[rr][ch] = (ch < C && i < r1) ? ra[(int64_t)ch * P.blocksize + i] : 0.0;
        }
        __syncthreads();
        for (int rr = 0; rr < 64; rr++) {
            double u = Ut[rr][tc];
#pragma unroll
            for (int k = 0; k < 16; k++) acc[k] = fma(u, Rt[rr][cg * 16 + k], acc[k]);
        }
        __syncthreads();
    }
    const int j = col0 + tc;
    if (j >= nb) return;
    for (int k = 0; k < 16; k++) {
        int ch = cg * 16 + k;
        if (ch < C) part[((int64_t)rg * P.blocksize + j) * C + ch] = acc[k];
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
    k_ua<<<dim3((nb + 63) / 64, nrg), NT, 0, st>>>(P, h->U, h->tri_off[b], nb, h->ra, h->rows_per_group,
                                                   h->b_rapart);
    return hip_check(hipGetLastError(), "batch products");
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

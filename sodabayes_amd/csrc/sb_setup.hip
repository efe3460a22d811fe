// Setup of the SODA augmenting matrices on gfx950 (replaces setup_soda and
// get_aug_matrix, soda.f90:403-518):
//   k_pack        .bed bytes -> 2-bit allele counts, block-position order
//   k_gram        X_s'X_s on int8 MFMA (exact int32) + xxproduct's FP64
//                 standardisation (soda.f90:520-556), all blocks in one launch
//   k_pm_*        batched power method (soda.f90:378-401)
//   k_chol_*      batched blocked Cholesky of dI - X_s'X_s (soda.f90:8-47)
//   k_pack_tri    upper factors -> packed column-major triangles
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>

#include "sb_device.h"
#include "sb_internal.h"

namespace sb {
namespace {

constexpr double DEPS = 2.220446049250313080847e-16;    // epsilon(1.0d0)
constexpr double POWER_TOL = (double)1.0e-14f;          // soda.f90:385 (f32 literal)
constexpr double D_EPS_ADD = (double)0.001f;            // soda.f90:485 (f32 literal)
constexpr int TILE = 64;                                // Cholesky tile

typedef int v4i __attribute__((ext_vector_type(4)));

// ------------------------------------------------------------------ pack --

// One 32-bit word = 16 genotypes (4 .bed bytes, little-endian); the code ->
// count transform w - ((w >> 1) & 0x55..) is utilities.f90:489.  Bytes past
// the SNP's ceil(n/4) are zero, as in load_geno_bits' zero-filled tail.
__global__ void k_pack(const uint8_t *__restrict__ bed, int64_t nbytes, const int32_t *permvec,
                       int p, int ldw, uint32_t *__restrict__ geno)
{
    int pos = blockIdx.x;
    int snp = permvec[pos];
    const uint8_t *col = bed + (int64_t)snp * nbytes;
    for (int w = blockIdx.y * blockDim.x + threadIdx.x; w < ldw; w += gridDim.y * blockDim.x) {
        uint32_t v = 0;
        int64_t b0 = (int64_t)w * 4;
        if (b0 + 3 < nbytes) {
            v = (uint32_t)col[b0] | ((uint32_t)col[b0 + 1] << 8) | ((uint32_t)col[b0 + 2] << 16) |
                ((uint32_t)col[b0 + 3] << 24);
        } else {
            for (int b = 0; b < 4; b++)
                if (b0 + b < nbytes) v |= (uint32_t)col[b0 + b] << (8 * b);
        }
        geno[(int64_t)pos * ldw + w] = v - ((v >> 1) & 0x55555555u);
    }
}

// ------------------------------------------------------------------ gram --

// 8 two-bit counts -> 8 int8 lanes of a 64-bit MFMA operand.
__device__ __forceinline__ uint32_t spread8(uint32_t b)
{
    uint32_t t = (b | (b << 12)) & 0x000F000Fu;
    return (t | (t << 6)) & 0x03030303u;
}
__device__ __forceinline__ long spread16(uint32_t x)
{
    return (long)(((uint64_t)spread8((x >> 8) & 0xFFu) << 32) | spread8(x & 0xFFu));
}

// xxproduct's FP64 epilogue (soda.f90:527-554) for the pair (lo, hi) of block
// positions, lo <= hi: the reference calls xxproduct(jj = hi, kk = lo).
__device__ __forceinline__ double std_gram(int S, double fhi, double flo, int nind)
{
    double pj = fhi, pk = flo, qj = 1.0 - pj, qk = 1.0 - pk;
    if (pj <= DEPS || qj <= DEPS || pk <= DEPS || qk <= DEPS) return 0.0;
    pj = pj * 2.0;
    pk = pk * 2.0;
    double res = (double)S;
    return (res - (double)nind * pj * pk) / sqrt(pj * pk * qj * qk);
}

// One workgroup = 64x64 tile of one block's Gram (4 waves of 32x32, each
// 2x2 MFMA 16x16x32 i8 accumulators).  Lane group g = lane/16 streams 32
// consecutive individuals per 128-individual step; A and B use the same
// individual -> k mapping, so the int32 sums are exact and order-free.
__global__ __launch_bounds__(256) void k_gram(const uint32_t *__restrict__ geno, int ldw,
                                              const int32_t *__restrict__ ls,
                                              const int64_t *__restrict__ goff,
                                              const double *__restrict__ freqp, int nind,
                                              double *__restrict__ G, int only_block)
{
    const int b = only_block >= 0 ? only_block : (int)blockIdx.y;
    const int p0 = ls[b], nb = ls[b + 1] - ls[b];
    const int T = (nb + 63) / 64;
    int t = blockIdx.x;
    if (t >= T * (T + 1) / 2) return;
    int ti = 0;
    while (t >= T - ti) {
        t -= T - ti;
        ti++;
    }
    const int tj = ti + t;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int r0 = ti * 64 + (wave >> 1) * 32, c0 = tj * 64 + (wave & 1) * 32;
    const int g = lane >> 4, l16 = lane & 15;
    const uint32_t *base = geno + (int64_t)p0 * ldw;
    const uint32_t *pa[2], *pb[2];
    bool va[2], vb[2];
#pragma unroll
    for (int f = 0; f < 2; f++) {
        int ra = r0 + 16 * f + l16, cb = c0 + 16 * f + l16;
        va[f] = ra < nb;
        vb[f] = cb < nb;
        pa[f] = base + (int64_t)(va[f] ? ra : 0) * ldw + g * 2;
        pb[f] = base + (int64_t)(vb[f] ? cb : 0) * ldw + g * 2;
    }
    v4i acc[2][2];
#pragma unroll
    for (int a = 0; a < 2; a++)
#pragma unroll
        for (int c = 0; c < 2; c++) acc[a][c] = v4i{0, 0, 0, 0};
    const int nsteps = ldw / 8;
    for (int ks = 0; ks < nsteps; ks++) {
        uint64_t wa[2], wb[2];
#pragma unroll
        for (int f = 0; f < 2; f++) {
            wa[f] = va[f] ? *(const uint64_t *)(pa[f] + ks * 8) : 0ull;
            wb[f] = vb[f] ? *(const uint64_t *)(pb[f] + ks * 8) : 0ull;
        }
#pragma unroll
        for (int c = 0; c < 4; c++) {
            long a8[2], b8[2];
#pragma unroll
            for (int f = 0; f < 2; f++) {
                a8[f] = spread16((uint32_t)(wa[f] >> (16 * c)) & 0xFFFFu);
                b8[f] = spread16((uint32_t)(wb[f] >> (16 * c)) & 0xFFFFu);
            }
#pragma unroll
            for (int fa = 0; fa < 2; fa++)
#pragma unroll
                for (int fb = 0; fb < 2; fb++)
                    acc[fa][fb] =
                        __builtin_amdgcn_mfma_i32_16x16x32_i8(a8[fa], b8[fb], acc[fa][fb], 0, 0, 0);
        }
    }
    double *Gb = G + goff[b];
#pragma unroll
    for (int fa = 0; fa < 2; fa++)
#pragma unroll
        for (int fb = 0; fb < 2; fb++)
#pragma unroll
            for (int e = 0; e < 4; e++) {
                int row = r0 + 16 * fa + 4 * g + e;
                int col = c0 + 16 * fb + l16;
                if (row >= nb || col >= nb) continue;
                int lo = row < col ? row : col, hi = row < col ? col : row;
                double v = std_gram(acc[fa][fb][e], freqp[p0 + hi], freqp[p0 + lo], nind);
                Gb[(int64_t)col * nb + row] = v;
                Gb[(int64_t)row * nb + col] = v;
            }
}

// ---------------------------------------------------------- power method --

// Start vectors: p uniforms of the replay stream in block order (each
// block's power_method draws n_b uniforms from stream 0, soda.f90:390), or
// Philox uniforms in COUNTER mode.
__global__ void k_start_replay(Scal *scal, int p, double *y)
{
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    Xoro s{scal->rng[0], scal->rng[1]};
    for (int i = 0; i < p; i++) y[i] = xoro_uniform(s);
    scal->rng[0] = s.s1;
    scal->rng[1] = s.s2;
}

__global__ void k_start_counter(uint64_t key, int p, double *y)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= p) return;
    PhiloxSrc src(key, 0xFFFFFFFFu, (uint32_t)i, 7u);
    y[i] = src.next();
}

// v = y / sqrt(sum(y**2)) per block (soda.f90:392)
__global__ __launch_bounds__(256) void k_pm_normalize(const int32_t *ls, const double *y, double *v)
{
    __shared__ double lds[8];
    int b = blockIdx.x, p0 = ls[b], nb = ls[b + 1] - p0;
    double s = 0.0;
    for (int i = threadIdx.x; i < nb; i += 256) s += y[p0 + i] * y[p0 + i];
    s = block_sum<256>(s, lds);
    __shared__ double nrm;
    if (threadIdx.x == 0) nrm = sqrt(s);
    __syncthreads();
    for (int i = threadIdx.x; i < nb; i += 256) v[p0 + i] = y[p0 + i] / nrm;
}

// y = A v for active blocks; one wave per output row (A symmetric, so row
// i of A is its contiguous column i), four independent partial sums per
// lane so that four 512-byte loads per wave are in flight.
__global__ __launch_bounds__(256) void k_pm_gemv(const double *__restrict__ G,
                                                 const int64_t *__restrict__ goff,
                                                 const int32_t *__restrict__ ls,
                                                 const int32_t *__restrict__ active,
                                                 const double *__restrict__ v, double *__restrict__ y,
                                                 const int32_t *__restrict__ blist)
{
    int b = blist ? blist[blockIdx.y] : (int)blockIdx.y;
    if (!active[b]) return;
    int p0 = ls[b], nb = ls[b + 1] - p0;
    int i = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
    if (i >= nb) return;
    const double *col = G + goff[b] + (int64_t)i * nb;
    const double *vb = v + p0;
    double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0;
    int j = lane;
    for (; j + 192 < nb; j += 256) {
        a0 = fma(col[j], vb[j], a0);
        a1 = fma(col[j + 64], vb[j + 64], a1);
        a2 = fma(col[j + 128], vb[j + 128], a2);
        a3 = fma(col[j + 192], vb[j + 192], a3);
    }
    for (; j < nb; j += 64) a0 = fma(col[j], vb[j], a0);
    double acc = wave_sum((a0 + a1) + (a2 + a3));
    if (lane == 0) y[p0 + i] = acc;
}

// lambda = y.v, delta = sum((y - lambda v)**2)/|lambda|, stop at
// delta <= tol (f32 1e-14) or maxit = 1000 n; else v = y/|y| (soda.f90:391-398)
__global__ __launch_bounds__(256) void k_pm_post(const int32_t *ls, int32_t *active, int32_t *iters,
                                                 double *lam_out, const double *y, double *v,
                                                 const int32_t *blist)
{
    __shared__ double lds[8];
    __shared__ double lam_s, nrm_s;
    __shared__ int stop_s;
    int b = blist ? blist[blockIdx.x] : (int)blockIdx.x;
    if (!active[b]) return;
    int p0 = ls[b], nb = ls[b + 1] - p0;
    double s = 0.0;
    for (int i = threadIdx.x; i < nb; i += 256) s += y[p0 + i] * v[p0 + i];
    s = block_sum<256>(s, lds);
    if (threadIdx.x == 0) lam_s = s;
    __syncthreads();
    double lam = lam_s, dl = 0.0, ss = 0.0;
    for (int i = threadIdx.x; i < nb; i += 256) {
        double t = y[p0 + i] - lam * v[p0 + i];
        dl += t * t;
        ss += y[p0 + i] * y[p0 + i];
    }
    dl = block_sum<256>(dl, lds);
    ss = block_sum<256>(ss, lds);
    if (threadIdx.x == 0) {
        int it = iters[b] + 1;
        iters[b] = it;
        double delta = dl / fabs(lam);
        stop_s = (delta <= POWER_TOL) || (it >= 1000 * nb);
        if (stop_s) {
            active[b] = 0;
            lam_out[b] = fabs(lam) + D_EPS_ADD;  // xxd(b) = |lambda| + 0.001
        }
        nrm_s = sqrt(ss);
    }
    __syncthreads();
    if (stop_s) return;
    for (int i = threadIdx.x; i < nb; i += 256) v[p0 + i] = y[p0 + i] / nrm_s;
}

// One Lanczos step per block (COUNTER mode bound, see setup_build_aug):
// w = G q_k (in y) -> alpha_k = w.q_k, full re-orthogonalisation of w against
// q_0..q_k (twice), beta_k = |w|, q_(k+1) = w / beta_k.  Q holds the Lanczos
// vectors, vector j of every block at Q + j * p.
__global__ __launch_bounds__(256) void k_lz_step(const int32_t *ls, int p, int k, double *y, double *Q,
                                                 double *alpha, double *beta, int K1)
{
    __shared__ double lds[8];
    __shared__ double coef;
    const int b = blockIdx.x, p0 = ls[b], nb = ls[b + 1] - p0;
    double *w = y + p0;
    for (int pass = 0; pass < 2; pass++) {
        for (int j = 0; j <= k; j++) {
            const double *qj = Q + (int64_t)j * p + p0;
            double s = 0.0;
            for (int i = threadIdx.x; i < nb; i += 256) s += w[i] * qj[i];
            s = block_sum<256>(s, lds);
            if (threadIdx.x == 0) {
                coef = s;
                if (pass == 0 && j == k) alpha[b * K1 + k] = s;
            }
            __syncthreads();
            const double cj = coef;
            for (int i = threadIdx.x; i < nb; i += 256) w[i] -= cj * qj[i];
            __syncthreads();
        }
    }
    double s = 0.0;
    for (int i = threadIdx.x; i < nb; i += 256) s += w[i] * w[i];
    s = block_sum<256>(s, lds);
    if (threadIdx.x == 0) {
        coef = sqrt(s);
        beta[b * K1 + k] = coef;
    }
    __syncthreads();
    const double bk = coef;
    double *qn = Q + (int64_t)(k + 1) * p + p0;
    for (int i = threadIdx.x; i < nb; i += 256) qn[i] = bk > 0.0 ? w[i] / bk : 0.0;
}

// ------------------------------------------------------------- cholesky --

// M = dI - G in place (soda.f90:487-490: xxtemp = -xxtemp; diag = d + diag)
__global__ void k_form_m(double *G, const int64_t *goff, const int32_t *ls, const double *xxd,
                         const int32_t *skip)
{
    int b = blockIdx.y, nb = ls[b + 1] - ls[b];
    if (skip && skip[b] != 0) return;
    double d = xxd[b];
    double *A = G + goff[b];
    int64_t n2 = (int64_t)nb * nb;
    for (int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; k < n2;
         k += (int64_t)gridDim.x * blockDim.x) {
        double m = -A[k];
        if (k % (nb + 1) == 0) m = d + m;
        A[k] = m;
    }
}

// Factor diagonal tile k of every block in LDS (dot-product form of
// choleskyPD, soda.f90:33-44).  fail[b] = first failing 1-based column.
__global__ __launch_bounds__(256) void k_chol_diag(double *G, const int64_t *goff, const int32_t *ls,
                                                   int k, int32_t *fail)
{
    __shared__ double A[TILE][TILE + 1];
    __shared__ double piv;
    __shared__ int bad;
    int b = blockIdx.x, nb = ls[b + 1] - ls[b];
    if (fail[b] != 0) return;
    int o = k * TILE;
    if (o >= nb) return;
    int m = nb - o < TILE ? nb - o : TILE;
    double *Ab = G + goff[b];
    for (int e = threadIdx.x; e < TILE * TILE; e += 256) {
        int r = e % TILE, c = e / TILE;
        A[r][c] = (r < m && c < m && r <= c) ? Ab[(int64_t)(o + c) * nb + (o + r)] : 0.0;
    }
    if (threadIdx.x == 0) bad = 0;
    __syncthreads();
    for (int i = 0; i < m; i++) {
        // A[i][i] currently holds M_ii - sum of previous tiles' updates
        if (threadIdx.x < 64) {
            double s = 0.0;
            for (int q = threadIdx.x; q < i; q += 64) s = fma(A[q][i], A[q][i], s);
            s = wave_sum(s);
            if (threadIdx.x == 0) {
                double tmp = A[i][i] - s;
                if (tmp <= 0.0 || isnan(tmp)) {
                    bad = o + i + 1;
                } else {
                    piv = sqrt(tmp);
                    A[i][i] = piv;
                }
            }
        }
        __syncthreads();
        if (bad) break;
        for (int j = i + 1 + threadIdx.x; j < m; j += 256) {
            double s = 0.0;
            for (int q = 0; q < i; q++) s = fma(A[q][j], A[q][i], s);
            A[i][j] = (A[i][j] - s) / piv;
        }
        __syncthreads();
    }
    if (bad) {
        if (threadIdx.x == 0) fail[b] = bad;
        return;
    }
    for (int e = threadIdx.x; e < TILE * TILE; e += 256) {
        int r = e % TILE, c = e / TILE;
        if (r < m && c < m && r <= c) Ab[(int64_t)(o + c) * nb + (o + r)] = A[r][c];
    }
}

// Row panel: U_kj = U_kk^{-T} A_kj for tiles j > k (forward substitution).
__global__ __launch_bounds__(256) void k_chol_panel(double *G, const int64_t *goff, const int32_t *ls,
                                                    int k, const int32_t *fail)
{
    __shared__ double Ukk[TILE][TILE + 1];
    __shared__ double X[TILE][TILE + 1];
    int b = blockIdx.y, nb = ls[b + 1] - ls[b];
    if (fail[b] != 0) return;
    int o = k * TILE, j0 = (k + 1 + blockIdx.x) * TILE;
    if (j0 >= nb) return;
    int m = nb - o < TILE ? nb - o : TILE;
    int w = nb - j0 < TILE ? nb - j0 : TILE;
    double *Ab = G + goff[b];
    for (int e = threadIdx.x; e < TILE * TILE; e += 256) {
        int r = e % TILE, c = e / TILE;
        Ukk[r][c] = (r < m && c < m && r <= c) ? Ab[(int64_t)(o + c) * nb + (o + r)] : 0.0;
        X[r][c] = (r < m && c < w) ? Ab[(int64_t)(j0 + c) * nb + (o + r)] : 0.0;
    }
    __syncthreads();
    // 4 threads per column: column c, row-slices interleaved; rows solved in order
    int c = threadIdx.x & 63, part = threadIdx.x >> 6;
    for (int i = 0; i < m; i++) {
        if (part == 0 && c < w) {
            double s = 0.0;
            for (int q = 0; q < i; q++) s = fma(Ukk[q][i], X[q][c], s);
            X[i][c] = (X[i][c] - s) / Ukk[i][i];
        }
        __syncthreads();
    }
    for (int e = threadIdx.x; e < TILE * TILE; e += 256) {
        int r = e % TILE, cc = e / TILE;
        if (r < m && cc < w) Ab[(int64_t)(j0 + cc) * nb + (o + r)] = X[r][cc];
    }
}

// Trailing update A_ij -= U_ki' U_kj for k < i <= j (upper tiles only).
__global__ __launch_bounds__(256) void k_chol_update(double *G, const int64_t *goff, const int32_t *ls,
                                                     int k, const int32_t *fail)
{
    __shared__ double Ui[TILE][TILE + 1];
    __shared__ double Uj[TILE][TILE + 1];
    int b = blockIdx.y, nb = ls[b + 1] - ls[b];
    if (fail[b] != 0) return;
    int T = (nb + TILE - 1) / TILE;
    int R = T - k - 1;  // trailing tiles
    int t = blockIdx.x;
    if (R <= 0 || t >= R * (R + 1) / 2) return;
    int ti = 0;
    while (t >= R - ti) {
        t -= R - ti;
        ti++;
    }
    int ii = k + 1 + ti, jj = ii + t;
    int o = k * TILE, i0 = ii * TILE, j0 = jj * TILE;
    int m = nb - o < TILE ? nb - o : TILE;
    int wi = nb - i0 < TILE ? nb - i0 : TILE;
    int wj = nb - j0 < TILE ? nb - j0 : TILE;
    double *Ab = G + goff[b];
    for (int e = threadIdx.x; e < TILE * TILE; e += 256) {
        int r = e % TILE, c = e / TILE;
        Ui[r][c] = (r < m && c < wi) ? Ab[(int64_t)(i0 + c) * nb + (o + r)] : 0.0;
        Uj[r][c] = (r < m && c < wj) ? Ab[(int64_t)(j0 + c) * nb + (o + r)] : 0.0;
    }
    __syncthreads();
    // thread owns a 4x4 micro-tile: rows ri..ri+3 (of tile i), cols cj..cj+3
    int tr = threadIdx.x & 15, tc = threadIdx.x >> 4;
    double acc[4][4];
#pragma unroll
    for (int a = 0; a < 4; a++)
#pragma unroll
        for (int c = 0; c < 4; c++) acc[a][c] = 0.0;
    for (int q = 0; q < m; q++) {
        double ui[4], uj[4];
#pragma unroll
        for (int a = 0; a < 4; a++) {
            ui[a] = Ui[q][tr + 16 * a];
            uj[a] = Uj[q][tc + 16 * a];
        }
#pragma unroll
        for (int a = 0; a < 4; a++)
#pragma unroll
            for (int c = 0; c < 4; c++) acc[a][c] = fma(ui[a], uj[c], acc[a][c]);
    }
#pragma unroll
    for (int a = 0; a < 4; a++)
#pragma unroll
        for (int c = 0; c < 4; c++) {
            int r = tr + 16 * a, cc = tc + 16 * c;
            if (r < wi && cc < wj && (ii < jj || r <= cc)) {
                double *dst = Ab + (int64_t)(j0 + cc) * nb + (i0 + r);
                *dst = *dst - acc[a][c];
            }
        }
}

// dfail = -1 marked blocks that were skipped (already factorised): reset to 0
__global__ void k_clear_skip(int32_t *fail, int nb)
{
    for (int b = threadIdx.x; b < nb; b += blockDim.x)
        if (fail[b] < 0) fail[b] = 0;
}

// Dense upper factor (ld n_b) -> packed column-major triangle.
__global__ void k_pack_tri(const double *G, const int64_t *goff, const int32_t *ls,
                           const int64_t *tri_off, double *U)
{
    int b = blockIdx.y, nb = ls[b + 1] - ls[b];
    int j = blockIdx.x;
    if (j >= nb) return;
    const double *src = G + goff[b] + (int64_t)j * nb;
    double *dst = U + tri_off[b] + (int64_t)j * (j + 1) / 2;
    for (int i = threadIdx.x; i <= j; i += blockDim.x) dst[i] = src[i];
}

}  // namespace

// ------------------------------------------------------------- host side --

int setup_pack(sb_handle *h, const uint8_t *bed, bool on_device)
{
    const Par &P = h->par;
    int64_t nbytes = ((int64_t)P.n + 3) / 4;
    const uint8_t *dbed = bed;
    TmpDev tmp;   // staging copy of a host payload, freed on every return path
    int rc = SB_OK;
    if (!on_device) {
        uint8_t *t = nullptr;
        if ((rc = tmp.alloc(&t, (size_t)(nbytes * P.p), "bed payload"))) return rc;
        if (hip_check(hipMemcpyAsync(t, bed, (size_t)(nbytes * P.p), hipMemcpyHostToDevice, h->stream),
                      "upload bed")) {
            hipStreamSynchronize(h->stream);
            return SB_ERR_HIP;
        }
        dbed = t;
    }
    int32_t *dperm = nullptr;
    rc = dev_alloc(h, (void **)&dperm, sizeof(int32_t) * P.p);
    if (rc) {
        hipStreamSynchronize(h->stream);
        return rc;
    }
    hipMemcpyAsync(dperm, h->permvec.data(), sizeof(int32_t) * P.p, hipMemcpyHostToDevice, h->stream);
    dim3 grid(P.p, std::min(64, (P.ldw + 255) / 256));
    k_pack<<<grid, 256, 0, h->stream>>>(dbed, nbytes, dperm, P.p, P.ldw, h->geno);
    rc = hip_check(hipGetLastError(), "k_pack");
    const int rs = hip_check(hipStreamSynchronize(h->stream), "pack");
    return rc ? rc : rs;
}

namespace {
struct GramBuf {
    double *G = nullptr;
    int64_t *goff = nullptr;
    std::vector<int64_t> hoff;
};

int gram_alloc(sb_handle *h, GramBuf &gb, int only_block)
{
    const Par &P = h->par;
    int nb = P.nblocks;
    gb.hoff.assign(nb + 1, 0);
    for (int b = 0; b < nb; b++) {
        int64_t n = h->ls[b + 1] - h->ls[b];
        int64_t sz = (only_block < 0 || only_block == b) ? n * n : 0;
        gb.hoff[b + 1] = gb.hoff[b] + sz;
    }
    if (only_block >= 0) {  // single block at offset 0
        for (int b = 0; b <= nb; b++) gb.hoff[b] = 0;
    }
    int64_t total = only_block >= 0 ? (int64_t)(h->ls[only_block + 1] - h->ls[only_block]) *
                                          (h->ls[only_block + 1] - h->ls[only_block])
                                    : gb.hoff[nb];
    if (hip_check(hipMalloc(&gb.G, sizeof(double) * (size_t)total), "hipMalloc(gram)")) {
        gb.G = nullptr;
        return SB_ERR_NOMEM;
    }
    if (hip_check(hipMalloc(&gb.goff, sizeof(int64_t) * (nb + 1)), "hipMalloc(goff)")) {
        hipFree(gb.G);
        gb.G = nullptr;
        gb.goff = nullptr;
        return SB_ERR_NOMEM;
    }
    hipMemcpyAsync(gb.goff, gb.hoff.data(), sizeof(int64_t) * (nb + 1), hipMemcpyHostToDevice, h->stream);
    return SB_OK;
}

void gram_free(GramBuf &gb)
{
    if (gb.G) hipFree(gb.G);
    if (gb.goff) hipFree(gb.goff);
    gb.G = nullptr;
    gb.goff = nullptr;
}

int launch_gram(sb_handle *h, GramBuf &gb, const double *freqp, int only_block)
{
    const Par &P = h->par;
    int T = (P.blocksize + 63) / 64;
    dim3 grid(T * (T + 1) / 2, only_block >= 0 ? 1 : P.nblocks);
    k_gram<<<grid, 256, 0, h->stream>>>(h->geno, P.ldw, h->d_ls, gb.goff, freqp, P.n, gb.G, only_block);
    return hip_check(hipGetLastError(), "k_gram");
}
}  // namespace

// Largest eigenvalue of the symmetric tridiagonal matrix (a[0..m), b[0..m-1))
// by Sturm-count bisection.
constexpr int LANCZOS_MAX = 192;  // Lanczos steps of the COUNTER-mode bound

static double tridiag_max(const double *a, const double *b, int m)
{
    double lo = a[0], hi = a[0];
    for (int i = 0; i < m; i++) {
        double r = (i > 0 ? std::fabs(b[i - 1]) : 0.0) + (i + 1 < m ? std::fabs(b[i]) : 0.0);
        lo = std::min(lo, a[i] - r);
        hi = std::max(hi, a[i] + r);
    }
    auto count_below = [&](double x) {
        int c = 0;
        double d = 1.0;
        for (int i = 0; i < m; i++) {
            double bb = i > 0 ? b[i - 1] * b[i - 1] : 0.0;
            d = a[i] - x - (i > 0 ? bb / d : 0.0);
            if (d == 0.0) d = -1e-300;
            if (d < 0.0) c++;
        }
        return c;
    };
    for (int it = 0; it < 200 && hi - lo > 1e-15 * std::max(1.0, std::fabs(hi)); it++) {
        double mid = 0.5 * (lo + hi);
        if (count_below(mid) >= m) hi = mid;  // every eigenvalue below mid
        else lo = mid;
    }
    return hi;
}

// Last component of the unit eigenvector of the symmetric tridiagonal T_m
// (diagonal a, off-diagonal b) for its eigenvalue theta, by two steps of
// inverse iteration.  (T - sigma I) is factored by Gaussian elimination with
// partial pivoting in the tridiagonal form of LAPACK's dgttrf (fill-in on a
// second superdiagonal).
static double tridiag_last_component(const double *a, const double *b, int m, double theta)
{
    if (m == 1) return 1.0;
    const double scale = std::max(std::fabs(theta), 1e-300);
    const double sigma = theta + 1e-10 * scale;
    std::vector<double> dl(m - 1), d(m), du(m - 1), du2(m > 2 ? m - 2 : 1, 0.0), x(m, 1.0);
    std::vector<int> piv(m - 1);
    for (int i = 0; i < m; i++) d[i] = a[i] - sigma;
    for (int i = 0; i + 1 < m; i++) dl[i] = du[i] = b[i];
    for (int i = 0; i + 1 < m; i++) {
        if (std::fabs(d[i]) >= std::fabs(dl[i])) {
            piv[i] = 0;
            const double l = d[i] != 0.0 ? dl[i] / d[i] : 0.0;
            dl[i] = l;
            d[i + 1] -= l * du[i];
        } else {
            piv[i] = 1;
            const double l = d[i] / dl[i];
            d[i] = dl[i];
            dl[i] = l;
            const double t = d[i + 1];
            d[i + 1] = du[i] - l * t;
            du[i] = t;
            if (i + 2 < m) {
                du2[i] = du[i + 1];
                du[i + 1] = -l * du[i + 1];
            }
        }
    }
    for (int i = 0; i < m; i++)
        if (d[i] == 0.0) d[i] = 1e-300 * scale;
    for (int step = 0; step < 2; step++) {
        for (int i = 0; i + 1 < m; i++) {
            if (piv[i]) std::swap(x[i], x[i + 1]);
            x[i + 1] -= dl[i] * x[i];
        }
        x[m - 1] /= d[m - 1];
        x[m - 2] = (x[m - 2] - du[m - 2] * x[m - 1]) / d[m - 2];
        for (int i = m - 3; i >= 0; i--) x[i] = (x[i] - du[i] * x[i + 1] - du2[i] * x[i + 2]) / d[i];
        double nrm = 0.0;
        for (int i = 0; i < m; i++) nrm = std::max(nrm, std::fabs(x[i]));
        for (int i = 0; i < m; i++) x[i] /= nrm;
    }
    double s2 = 0.0;
    for (int i = 0; i < m; i++) s2 += x[i] * x[i];
    return x[m - 1] / std::sqrt(s2);
}

// freq per position is kept in the handle's zlut slot 3 (see sb_api.hip).
static const double *freq_pos(sb_handle *h) { return h->f2 + h->par.p; }

int setup_gram_block(sb_handle *h, int block, double *G_host)
{
    if (block < 0 || block >= h->par.nblocks) return fail(SB_ERR_ARG, "block out of range");
    GramBuf gb;
    int rc = gram_alloc(h, gb, block);
    if (rc) return rc;
    rc = launch_gram(h, gb, freq_pos(h), block);
    if (!rc) {
        int64_t n = h->ls[block + 1] - h->ls[block];
        rc = hip_check(hipMemcpyAsync(G_host, gb.G, sizeof(double) * n * n, hipMemcpyDeviceToHost, h->stream),
                       "download gram");
        if (!rc) rc = hip_check(hipStreamSynchronize(h->stream), "gram sync");
    }
    gram_free(gb);
    return rc;
}

// --------------------------------------------------------- PSD fallback --

// choleskyPD (soda.f90:8-47) on the host, row by row as the reference runs
// it: on a failing pivot, rows 1..i-1 of U are complete and the rest zero;
// with a non-positive diagonal entry U is left untouched (returns -1).
// Column-major, 0-based: A(r, c) = A[c * n + r].
static int host_chol_pd(int n, const double *A, double *U)
{
    for (int i = 0; i < n; i++)
        if (A[(int64_t)i * n + i] <= 0.0) return -1;
    std::fill(U, U + (int64_t)n * n, 0.0);
    for (int i = 0; i < n; i++) {
        double dot = 0.0;
        for (int k = 0; k < i; k++) dot += U[(int64_t)i * n + k] * U[(int64_t)i * n + k];
        const double tmp = A[(int64_t)i * n + i] - dot;
        if (tmp <= 0.0 || std::isnan(tmp)) return i + 1;
        U[(int64_t)i * n + i] = std::sqrt(tmp);
        for (int j = i + 1; j < n; j++) {
            double d2 = 0.0;
            for (int k = 0; k < i; k++) d2 += U[(int64_t)j * n + k] * U[(int64_t)i * n + k];
            U[(int64_t)j * n + i] = (A[(int64_t)i * n + j] - d2) / U[(int64_t)i * n + i];
        }
    }
    return 0;
}

// choleskyPSD (soda.f90:49-124): complete-pivoting Cholesky in place with
// the reference helpers' operation order (IDMAX takes the first maximum;
// the transposed DGEMV sums rows in order; DSCAL multiplies by 1/Ajj).
// piv is 1-based; returns 0, or -1 when the largest diagonal is <= eps.
static int host_chol_psd(int n, double *A, int *piv, int *rank)
{
    auto a = [&](int r, int c) -> double & { return A[(int64_t)c * n + r]; };
    std::vector<double> work(2 * (size_t)n, 0.0);
    for (int q = 0; q < n; q++) piv[q] = q + 1;
    int pvt = 0;
    double best = a(0, 0);
    for (int i = 1; i < n; i++)
        if (a(i, i) > best) {
            pvt = i;
            best = a(i, i);
        }
    double Ajj = a(pvt, pvt);
    if (Ajj <= DEPS) {
        *rank = 0;
        return -1;
    }
    const double DSTOP = n * DEPS * Ajj;
    for (int j = 0; j < n; j++) {
        for (int q = j; q < n; q++) {
            if (j > 0) work[q] = work[q] + a(j - 1, q) * a(j - 1, q);
            work[n + q] = a(q, q) - work[q];
        }
        if (j > 0) {
            int it = 0;
            double r = work[n + j];
            for (int i = 1; i < n - j; i++)
                if (work[n + j + i] > r) {
                    it = i;
                    r = work[n + j + i];
                }
            pvt = it + j;
            Ajj = work[n + pvt];
            if (Ajj <= DSTOP) {
                a(j, j) = Ajj;
                *rank = j;
                return 0;
            }
        }
        if (j != pvt) {
            a(pvt, pvt) = a(j, j);
            for (int i = 0; i < j; i++) std::swap(a(i, j), a(i, pvt));
            for (int i = 0; i < pvt - j - 1; i++) std::swap(a(j, j + 1 + i), a(j + 1 + i, pvt));
            for (int i = 0; i < n - pvt - 1; i++) std::swap(a(j, pvt + 1 + i), a(pvt, pvt + 1 + i));
            std::swap(work[j], work[pvt]);
            std::swap(piv[pvt], piv[j]);
        }
        Ajj = std::sqrt(Ajj);
        a(j, j) = Ajj;
        if (j < n - 1) {
            if (j > 0)
                for (int c = 0; c < n - j - 1; c++) {
                    double temp = 0.0;
                    for (int i = 0; i < j; i++) temp = temp + a(i, j + 1 + c) * a(i, j);
                    a(j, j + 1 + c) = a(j, j + 1 + c) + (-1.0) * temp;
                }
            const double da = 1.0 / Ajj;
            for (int c = 0; c < n - j - 1; c++) a(j, j + 1 + c) = da * a(j, j + 1 + c);
        }
    }
    *rank = n;
    return 0;
}

// Blocks whose dI - G is not positive definite take the reference's fallback
// (get_aug_matrix, soda.f90:490-510) on the host: the Gram is recomputed on
// the GPU and downloaded, choleskyPD runs (its partial factor is what the
// reference leaves in XA), then choleskyPSD; rank columns of the pivoted
// factor go to their block positions, the rest are skipped.  XA of such a
// block is dense (positions are permuted by the pivoting), stored in Ud.
static int psd_fallback(sb_handle *h, GramBuf &gb, const std::vector<int> &failed, const std::vector<double> &xxd)
{
    const Par &P = h->par;
    int64_t dense = 0;
    for (int b : failed) {
        const int64_t nbb = h->ls[b + 1] - h->ls[b];
        h->dense_off[b] = dense;
        dense += nbb * nbb;
    }
    int rc = dev_alloc(h, (void **)&h->Ud, sizeof(double) * (size_t)dense);
    if (rc) return rc;
    // the handle's stream does not synchronise with the null stream: every
    // copy below goes through it
    std::vector<uint8_t> hskip(P.p, 0);
    for (int b : failed) {
        const int p0 = h->ls[b], n = h->ls[b + 1] - p0;
        rc = launch_gram(h, gb, freq_pos(h), b);
        std::vector<double> M((size_t)n * n), U((size_t)n * n, 0.0);
        if (!rc)
            rc = hip_check(hipMemcpyAsync(M.data(), gb.G + gb.hoff[b], sizeof(double) * M.size(),
                                          hipMemcpyDeviceToHost, h->stream),
                           "download gram");
        if (!rc) rc = hip_check(hipStreamSynchronize(h->stream), "download gram");
        if (rc) return rc;
        for (auto &v : M) v = -v;
        for (int j = 0; j < n; j++) M[(size_t)j * n + j] = xxd[b] + M[(size_t)j * n + j];
        host_chol_pd(n, M.data(), U.data());
        std::vector<int> piv(n);
        int rank = 0;
        if (host_chol_psd(n, M.data(), piv.data(), &rank) != 0)
            return fail(SB_ERR_NOT_PD, "D - X'X is not positive semidefinite in block " + std::to_string(b + 1) +
                                           " (soda.f90:496-500)");
        for (int j = 0; j < n; j++) {
            const int pos = piv[j] - 1;
            if (j < rank) {
                for (int r = 0; r <= j; r++) U[(size_t)pos * n + r] = M[(size_t)j * n + r];
            } else {
                hskip[p0 + pos] = 1;
                h->skip_list.push_back(p0 + pos);
            }
        }
        rc = hip_check(hipMemcpyAsync(h->Ud + h->dense_off[b], U.data(), sizeof(double) * U.size(),
                                      hipMemcpyHostToDevice, h->stream),
                       "upload fallback factor");
        if (!rc) rc = hip_check(hipStreamSynchronize(h->stream), "upload fallback factor");
        if (rc) return rc;
    }
    h->nskipped = (int)h->skip_list.size();
    rc = hip_check(hipMemcpyAsync(h->skip, hskip.data(), P.p, hipMemcpyHostToDevice, h->stream), "upload skip flags");
    if (!rc && h->nskipped) {
        rc = dev_alloc(h, (void **)&h->d_skip_list, sizeof(int32_t) * h->nskipped);
        if (!rc) rc = dev_alloc(h, (void **)&h->xpx, sizeof(double) * h->nskipped);
        if (!rc)
            rc = hip_check(hipMemcpyAsync(h->d_skip_list, h->skip_list.data(), sizeof(int32_t) * h->nskipped,
                                          hipMemcpyHostToDevice, h->stream),
                           "upload skipped list");
    }
    if (!rc) rc = hip_check(hipStreamSynchronize(h->stream), "fallback");
    return rc;
}

int setup_build_aug(sb_handle *h, double *xxd_out, int32_t *fail_block, int32_t *fail_col)
{
    auto t0 = std::chrono::steady_clock::now();
    const Par &P = h->par;
    if (P.regular) {
        // -SODAOff (models.f90:607-962): no augmenting matrices; every SNP is
        // updated one by one in id order (permvec = identity, models.f90:672-675)
        // by the single-site sweep, which gets its windows' layouts here
        h->skip_list.resize(P.p);
        for (int k = 0; k < P.p; k++) h->skip_list[k] = h->pos_of[k];
        h->nskipped = P.p;
        int rc = dev_alloc(h, (void **)&h->d_skip_list, sizeof(int32_t) * P.p);
        if (!rc)
            rc = hip_check(hipMemcpy(h->d_skip_list, h->skip_list.data(), sizeof(int32_t) * P.p,
                                     hipMemcpyHostToDevice),
                           "upload SNP order");
        if (!rc) rc = single_setup(h);
        if (rc) return rc;
        h->aug_built = true;
        h->setup_seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        return SB_OK;
    }
    const int nb = P.nblocks;
    GramBuf gb;
    int rc = gram_alloc(h, gb, -1);
    if (rc) return rc;
    rc = launch_gram(h, gb, freq_pos(h), -1);
    if (rc) {
        gram_free(gb);
        return rc;
    }
    const char *vb = std::getenv("SODABAYES_VERBOSE");
    const bool verbose = vb && vb[0] == '1';
    auto progress = [&](const char *what, long it, int active) {
        if (!verbose) return;
        hipStreamSynchronize(h->stream);
        double el = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        std::fprintf(stderr, "[sodabayes] setup %.1f s: %s (iteration %ld, %d blocks active)\n", el, what, it,
                     active);
    };
    progress("gram done", 0, nb);
    // fallback state of an earlier call on this handle
    h->skip_list.clear();
    h->nskipped = 0;
    h->dense_off.assign(nb, -1);
    hipMemsetAsync(h->skip, 0, (size_t)P.p, h->stream);
    // sb_test_shift_d (tests only): lower one block's d before the factorisation
    const int shift_block = h->test_dshift_block;
    const double shift_amt = h->test_dshift;
    const bool dshift = shift_block >= 0 && shift_block < nb;
    auto apply_dshift = [&]() {
        double v = 0.0;
        hipMemcpyAsync(&v, h->xxd + shift_block, sizeof v, hipMemcpyDeviceToHost, h->stream);
        hipStreamSynchronize(h->stream);
        v = v - shift_amt;
        hipMemcpyAsync(h->xxd + shift_block, &v, sizeof v, hipMemcpyHostToDevice, h->stream);
        hipStreamSynchronize(h->stream);
    };
    // power method; every buffer of this call is freed on every return path
    TmpDev tmp;
    double *y = nullptr, *v = nullptr;
    int32_t *active = nullptr, *iters = nullptr, *dfail = nullptr, *blist = nullptr;
    if ((rc = tmp.alloc(&y, sizeof(double) * P.p, "power-method vector")) ||
        (rc = tmp.alloc(&v, sizeof(double) * P.p, "power-method vector")) ||
        (rc = tmp.alloc(&active, sizeof(int32_t) * nb, "block flags")) ||
        (rc = tmp.alloc(&iters, sizeof(int32_t) * nb, "block flags")) ||
        (rc = tmp.alloc(&dfail, sizeof(int32_t) * nb, "block flags")) ||
        (rc = tmp.alloc(&blist, sizeof(int32_t) * nb, "block list"))) {
        gram_free(gb);
        return rc;
    }
    std::vector<int32_t> ones(nb, 1);
    hipMemcpyAsync(active, ones.data(), sizeof(int32_t) * nb, hipMemcpyHostToDevice, h->stream);
    hipMemsetAsync(iters, 0, sizeof(int32_t) * nb, h->stream);
    hipMemsetAsync(dfail, 0, sizeof(int32_t) * nb, h->stream);
    if (P.rng_mode == SB_RNG_REPLAY)
        k_start_replay<<<1, 1, 0, h->stream>>>(h->scal, P.p, y);
    else
        k_start_counter<<<(P.p + 255) / 256, 256, 0, h->stream>>>(P.seed, P.p, y);
    k_pm_normalize<<<nb, 256, 0, h->stream>>>(h->d_ls, y, v);
    dim3 ggrid((P.blocksize + 3) / 4, nb);
    std::vector<int32_t> hact(nb);
    // REPLAY reproduces the reference's power iteration (its d to 1e-12).
    // COUNTER needs only a valid bound d > lambda_max: K Lanczos steps with
    // full re-orthogonalisation give the top Ritz value theta (exact for
    // K = n_b) and its residual |beta_K s_K|, an interval around theta that
    // holds an eigenvalue; d = theta + residual + 0.001.  The Cholesky below
    // still checks that dI - G is positive definite.
    const bool lanczos = P.rng_mode == SB_RNG_COUNTER;
    const int64_t maxit = lanczos ? 0 : 1000LL * P.blocksize;
    if (lanczos) {
        const int K = std::min(P.blocksize, LANCZOS_MAX), K1 = K + 1;
        TmpDev lz;   // Lanczos basis: p x (K + 1) doubles (0.9 GB at 600,000 SNPs), freed after the bound
        double *Qv = nullptr, *al = nullptr, *be = nullptr;
        if ((rc = lz.alloc(&Qv, sizeof(double) * (size_t)P.p * (K + 1), "Lanczos basis")) ||
            (rc = lz.alloc(&al, sizeof(double) * nb * K1, "Lanczos coefficients")) ||
            (rc = lz.alloc(&be, sizeof(double) * nb * K1, "Lanczos coefficients"))) {
            gram_free(gb);
            return rc;
        }
        hipMemcpyAsync(Qv, v, sizeof(double) * P.p, hipMemcpyDeviceToDevice, h->stream);
        hipMemsetAsync(active, 0xFF, sizeof(int32_t) * nb, h->stream);
        for (int k = 0; k < K; k++) {
            k_pm_gemv<<<ggrid, 256, 0, h->stream>>>(gb.G, gb.goff, h->d_ls, active, Qv + (int64_t)k * P.p, y,
                                                    nullptr);
            k_lz_step<<<nb, 256, 0, h->stream>>>(h->d_ls, P.p, k, y, Qv, al, be, K1);
        }
        std::vector<double> ha((size_t)nb * K1), hb((size_t)nb * K1), hx(nb);
        hipMemcpyAsync(ha.data(), al, sizeof(double) * nb * K1, hipMemcpyDeviceToHost, h->stream);
        hipMemcpyAsync(hb.data(), be, sizeof(double) * nb * K1, hipMemcpyDeviceToHost, h->stream);
        rc = hip_check(hipStreamSynchronize(h->stream), "lanczos");
        for (int b = 0; b < nb && !rc; b++) {
            const int nbb = h->ls[b + 1] - h->ls[b];
            const double *ab = &ha[(size_t)b * K1], *bb = &hb[(size_t)b * K1];
            int m = std::min(K, nbb);
            bool exact = m == nbb;
            // stop at an invariant subspace (beta ~ 0)
            for (int k = 0; k + 1 < m; k++)
                if (bb[k] <= 1e-13 * std::fabs(ab[0])) {
                    m = k + 1;
                    exact = true;
                    break;
                }
            const double theta = tridiag_max(ab, bb, m);
            const double resid = exact ? 0.0 : std::fabs(bb[m - 1] * tridiag_last_component(ab, bb, m, theta));
            hx[b] = std::fabs(theta) + resid + D_EPS_ADD;
        }
        if (!rc) rc = hip_check(hipMemcpyAsync(h->xxd, hx.data(), sizeof(double) * nb, hipMemcpyHostToDevice,
                                               h->stream), "upload bounds");
        hipStreamSynchronize(h->stream);
        if (rc) {
            gram_free(gb);
            return rc;
        }
        progress("lanczos bound done", K, 0);
    }
    // launches cover the still-active blocks only (blist), refreshed at
    // every host check
    auto power_loop = [&](int64_t limit) {
        hipMemcpyAsync(hact.data(), active, sizeof(int32_t) * nb, hipMemcpyDeviceToHost, h->stream);
        hipStreamSynchronize(h->stream);
        std::vector<int32_t> hl;
        for (int b = 0; b < nb; b++)
            if (hact[b]) hl.push_back(b);
        for (int64_t it = 0; it < limit && !hl.empty(); it += 16) {
            hipMemcpyAsync(blist, hl.data(), sizeof(int32_t) * hl.size(), hipMemcpyHostToDevice, h->stream);
            const int na = (int)hl.size();
            for (int q = 0; q < 16; q++) {
                k_pm_gemv<<<dim3(ggrid.x, na), 256, 0, h->stream>>>(gb.G, gb.goff, h->d_ls, active, v, y, blist);
                k_pm_post<<<na, 256, 0, h->stream>>>(h->d_ls, active, iters, h->xxd, y, v, blist);
            }
            if (hip_check(hipMemcpyAsync(hact.data(), active, sizeof(int32_t) * nb, hipMemcpyDeviceToHost,
                                         h->stream),
                          "power method"))
                break;
            hipStreamSynchronize(h->stream);
            hl.clear();
            for (int b = 0; b < nb; b++)
                if (hact[b]) hl.push_back(b);
            if (it % 2048 == 0) progress("power method", it + 16, (int)hl.size());
        }
    };
    power_loop(maxit);
    if (dshift) apply_dshift();
    // Cholesky of dI - G (blocks with dfail != 0 are skipped)
    const int T = (P.blocksize + TILE - 1) / TILE;
    auto factor = [&]() {
        k_form_m<<<dim3(64, nb), 256, 0, h->stream>>>(gb.G, gb.goff, h->d_ls, h->xxd, dfail);
        for (int k = 0; k < T; k++) {
            k_chol_diag<<<nb, 256, 0, h->stream>>>(gb.G, gb.goff, h->d_ls, k, dfail);
            if (T - k - 1 > 0) {
                k_chol_panel<<<dim3(T - k - 1, nb), 256, 0, h->stream>>>(gb.G, gb.goff, h->d_ls, k, dfail);
                int R = T - k - 1;
                k_chol_update<<<dim3(R * (R + 1) / 2, nb), 256, 0, h->stream>>>(gb.G, gb.goff, h->d_ls, k,
                                                                               dfail);
            }
        }
    };
    factor();
    if (lanczos) {
        // a block whose Lanczos bound left dI - G indefinite gets the
        // reference's power iteration: rebuild its Gram, iterate, refactor
        std::vector<int32_t> hf(nb), mask(nb);
        hipMemcpyAsync(hf.data(), dfail, sizeof(int32_t) * nb, hipMemcpyDeviceToHost, h->stream);
        hipStreamSynchronize(h->stream);
        int nretry = 0;
        for (int b = 0; b < nb; b++) {
            mask[b] = hf[b] != 0;
            nretry += mask[b];
        }
        if (nretry) {
            progress("lanczos bound too low, power method for the failed blocks", 0, nretry);
            for (int b = 0; b < nb; b++)
                if (mask[b]) launch_gram(h, gb, freq_pos(h), b);
            hipMemcpyAsync(active, mask.data(), sizeof(int32_t) * nb, hipMemcpyHostToDevice, h->stream);
            k_start_counter<<<(P.p + 255) / 256, 256, 0, h->stream>>>(P.seed, P.p, y);
            k_pm_normalize<<<nb, 256, 0, h->stream>>>(h->d_ls, y, v);
            power_loop(1000LL * P.blocksize);
            if (dshift && mask[shift_block]) apply_dshift();
            std::vector<int32_t> skip(nb);
            for (int b = 0; b < nb; b++) skip[b] = mask[b] ? 0 : -1;
            hipMemcpyAsync(dfail, skip.data(), sizeof(int32_t) * nb, hipMemcpyHostToDevice, h->stream);
            factor();
            // blocks factorised in the first pass are fine
            k_clear_skip<<<1, 256, 0, h->stream>>>(dfail, nb);
        }
    }
    k_pack_tri<<<dim3(P.blocksize, nb), 256, 0, h->stream>>>(gb.G, gb.goff, h->d_ls, h->d_tri_off, h->U);
    progress("cholesky done", 0, 0);
    rc = hip_check(hipGetLastError(), "cholesky");
    std::vector<int32_t> hfail(nb), hit(nb);
    std::vector<double> hxxd(nb);
    if (!rc) {
        hipMemcpyAsync(hfail.data(), dfail, sizeof(int32_t) * nb, hipMemcpyDeviceToHost, h->stream);
        hipMemcpyAsync(hit.data(), iters, sizeof(int32_t) * nb, hipMemcpyDeviceToHost, h->stream);
        hipMemcpyAsync(hxxd.data(), h->xxd, sizeof(double) * nb, hipMemcpyDeviceToHost, h->stream);
        rc = hip_check(hipStreamSynchronize(h->stream), "build_aug");
    }
    std::vector<int> failed;
    for (int b = 0; b < nb; b++)
        if (hfail[b] != 0) failed.push_back(b);
    if (fail_block) *fail_block = 0;
    if (fail_col) *fail_col = 0;
    if (!rc && !failed.empty()) {
        // both paths: the mixed path sweeps these blocks with the FP64 block
        // kernels (their factors are dense, with skipped SNPs)
        progress("pivoted-Cholesky fallback", 0, (int)failed.size());
        rc = psd_fallback(h, gb, failed, hxxd);
    }
    if (!rc && h->Uq0) rc = batch_tiles(h);
    const int rs = hip_check(hipStreamSynchronize(h->stream), "build_aug");
    if (!rc) rc = rs;
    gram_free(gb);
    if (rc) return rc;
    if (xxd_out) std::copy(hxxd.begin(), hxxd.end(), xxd_out);
    // naind = blocksize - naless (soda.f90:440, 460, 511-514): naless is 0
    // once a block factorises; otherwise the least blocksize - n_b + nskipped
    int naless = P.p, nsk = 0;
    for (int b = 0; b < nb; b++) {
        const int p0 = h->ls[b], nbb = h->ls[b + 1] - p0;
        if (hfail[b] != 0) {
            for (int q : h->skip_list) nsk += (q >= p0 && q < p0 + nbb);
        } else {
            naless = 0;
        }
        naless = std::min(naless, P.blocksize - nbb + nsk);
    }
    h->par.naind = P.blocksize - naless;
    h->pm_iters_max = 0;
    for (int b = 0; b < nb; b++) h->pm_iters_max = std::max(h->pm_iters_max, (double)hit[b]);
    h->aug_built = true;
    h->setup_seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    return SB_OK;
}

}  // namespace sb

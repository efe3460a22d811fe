// Mixed-precision, chain-batched block products of the SODA sweep.
//
// C chains share one pass over a block's genotypes.  Each chain's residual
// r_c is written as a fixed-point integer q = rint(r_c * 2^e_c) with
// |q| < 2^(8L-1), split into L balanced radix-256 digits (int8), and
//   sum_i x_ij r_ic  ~=  2^-e_c * sum_k 256^k * T_jck,   T_jck = sum_i x_ij d_ick
// where every T is an exact int32 sum from int8 MFMA
// (v_mfma_i32_16x16x64_i8: one 32-bit word of sixteen 2-bit genotypes is
// one lane's 16-byte operand).  The only error is the rounding of r to
// 8L-1 (or, at the top of a binade, 8L-2) bits of its largest entry.  The
// residual update r_c += Z_s delta_c is the same int8 product the other way
// round (individual-major genotypes times the limbs of delta / s), with each
// SNP's rounding remainder carried into its next update.  The augmented
// parts U_s' R_a and U_s diag(s) c~ run on the same int8 MFMA, with U as four
// balanced limbs per entry (31 bits of its column's or row's largest entry),
// as trailing workgroups of the genotype launches.  Chains run in tiles of
// 16; 49-64 chains run as two chain groups of 32 in each launch.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <type_traits>

#include "sb_device.h"
#include "sb_draw.h"
#include "sb_internal.h"

namespace sb {

namespace {
constexpr int NT = 256;
typedef int v4i __attribute__((ext_vector_type(4)));
constexpr uint32_t M2 = 0x03030303u;
constexpr int LIMBS = 3;  // int8 limbs per fixed-point value (23 bits)

// One 32-bit word of 2-bit counts as a 16-byte MFMA operand: dword m holds
// the counts of positions m, m+4, m+8, m+12 (bytes 0..3).  Hence the limb
// slot order {0,4,8,12, 1,5,9,13, 2,6,10,14, 3,7,11,15} of the other operand.
__device__ __forceinline__ v4i unpack_word(uint32_t w)
{
    return v4i{(int)(w & M2), (int)((w >> 2) & M2), (int)((w >> 4) & M2), (int)((w >> 6) & M2)};
}
// The MFMA operands of sub-step s of IT genotype tiles (word s of each
// tile's 4-word lane operand), two tiles at a time through 64-bit shifts
// (one shift serves both words): 11 instead of 14 VALU per pair.
template <int IT>
__device__ __forceinline__ void unpack_tiles(const v4i (&a)[IT], int s, v4i (&av)[IT])
{
    constexpr uint64_t M2L = 0x0303030303030303ull;
#pragma unroll
    for (int it = 0; it + 1 < IT; it += 2) {
        const uint64_t w = ((uint64_t)(uint32_t)a[it + 1][s] << 32) | (uint32_t)a[it][s];
        const uint64_t d0 = w & M2L, d1 = (w >> 2) & M2L, d2 = (w >> 4) & M2L, d3 = (w >> 6) & M2L;
        av[it] = v4i{(int)(uint32_t)d0, (int)(uint32_t)d1, (int)(uint32_t)d2, (int)(uint32_t)d3};
        av[it + 1] = v4i{(int)(d0 >> 32), (int)(d1 >> 32), (int)(d2 >> 32), (int)(d3 >> 32)};
    }
    if constexpr (IT % 2 == 1) av[IT - 1] = unpack_word((uint32_t)a[IT - 1][s]);
}
__device__ __forceinline__ v4i as_v4i(const uint4 &u) { return v4i{(int)u.x, (int)u.y, (int)u.z, (int)u.w}; }
__device__ __forceinline__ uint32_t word_of(const uint4 &u, int s)
{
    return s == 0 ? u.x : (s == 1 ? u.y : (s == 2 ? u.z : u.w));
}
__device__ __forceinline__ int slot_of(int t) { return 4 * (t & 3) + (t >> 2); }

// 64-bit values moved across lanes by DPP (two 32-bit halves): CTRL 0xB1 =
// quad_perm [1,0,3,2] (lane ^ 1), 0x4E = quad_perm [2,3,0,1] (lane ^ 2),
// 0x124 / 0x128 = row_ror 4 / 8 (rotate right within the 16-lane row)
template <int CTRL>
__device__ __forceinline__ uint64_t dpp_u64(uint64_t v)
{
    const int lo = __builtin_amdgcn_update_dpp((int)(uint32_t)v, (int)(uint32_t)v, CTRL, 0xf, 0xf, false);
    const int hi = __builtin_amdgcn_update_dpp((int)(uint32_t)(v >> 32), (int)(uint32_t)(v >> 32), CTRL, 0xf, 0xf,
                                               false);
    return ((uint64_t)(uint32_t)hi << 32) | (uint32_t)lo;
}
__device__ __forceinline__ uint64_t readlane_u64(uint64_t v, int l)
{
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)v, l);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(v >> 32), l);
    return ((uint64_t)hi << 32) | lo;
}

// Sum (double) and maximum (unsigned 64-bit) over the 64 lanes, as VALU
// work: quad and row steps by DPP (every lane of a 16-lane row then holds
// the row's total), the four row totals read from lanes 0, 16, 32, 48 and
// combined in that order.  The result is wave-uniform and deterministic.
__device__ __forceinline__ void wave_sum_max(double &sm, uint64_t &mx)
{
#define SB_STEP(CTRL)                                                              \
    {                                                                              \
        sm += __longlong_as_double((long long)dpp_u64<CTRL>((uint64_t)__double_as_longlong(sm))); \
        const uint64_t o = dpp_u64<CTRL>(mx);                                      \
        mx = o > mx ? o : mx;                                                      \
    }
    SB_STEP(0xB1)
    SB_STEP(0x4E)
    SB_STEP(0x124)
    SB_STEP(0x128)
#undef SB_STEP
    const uint64_t s0 = readlane_u64((uint64_t)__double_as_longlong(sm), 0);
    const uint64_t s1 = readlane_u64((uint64_t)__double_as_longlong(sm), 16);
    const uint64_t s2 = readlane_u64((uint64_t)__double_as_longlong(sm), 32);
    const uint64_t s3 = readlane_u64((uint64_t)__double_as_longlong(sm), 48);
    sm = ((__longlong_as_double((long long)s0) + __longlong_as_double((long long)s1)) +
          __longlong_as_double((long long)s2)) + __longlong_as_double((long long)s3);
    uint64_t m = readlane_u64(mx, 0);
    const uint64_t m1 = readlane_u64(mx, 16), m2 = readlane_u64(mx, 32), m3 = readlane_u64(mx, 48);
    m = m1 > m ? m1 : m;
    m = m2 > m ? m2 : m;
    m = m3 > m ? m3 : m;
    mx = m;
}

// Limb-major operand rows: limb k of chain c is row k * CP + c, CP = 16 *
// ntc (chains padded to whole 16-column tiles), so NTILE = LIMBS * ntc
// column tiles and the LIMBS limbs of a chain sit in the same lane of
// column tiles tc, ntc + tc, 2 ntc + tc.  With four chain tiles (49-64
// chains) the rows form two chain groups of 32 (group g: chains 32 g ..
// 32 g + 31, rows 96 g + 32 k + c mod 32), so that the products run as two
// groups of six column tiles each.
constexpr int CHAIN_GROUP = 32;
constexpr int chain_groups(int ntc) { return ntc == 4 ? 2 : 1; }
constexpr int limb_row(int k, int c, int ntc)
{
    return ntc == 4 ? (c / CHAIN_GROUP) * (LIMBS * CHAIN_GROUP) + k * CHAIN_GROUP + c % CHAIN_GROUP
                    : k * 16 * ntc + c;
}
static_assert(LIMBS == 3, "the products combine three limbs per chain");

// Row tiles per wave of the int8 products: four while the accumulators of
// NTILE column tiles fit in 96 registers, else two.
constexpr int prod_rw(int ntile) { return ntile <= 6 ? 4 : 2; }

// (16-column limb tiles, 16-chain tiles) of LIMBS limbs per chain
template <class F>
int with_tiles(int ntile, int ntc, F &&f)
{
    using std::integral_constant;
#define SB_TILE_CASE(b)                                                                   \
    if (ntc == b && ntile == LIMBS * b) {                                                 \
        f(integral_constant<int, LIMBS * b>{}, integral_constant<int, b>{});              \
        return 0;                                                                         \
    }
    SB_TILE_CASE(1)
    SB_TILE_CASE(2)
    SB_TILE_CASE(3)
#undef SB_TILE_CASE
    if (ntc == 4 && ntile == LIMBS * 4) {   // two chain groups of two tiles
        f(integral_constant<int, LIMBS * 2>{}, integral_constant<int, 2>{});
        return 0;
    }
    return SB_ERR_ARG;
}

}  // namespace

// per chain, per workgroup: max |r| over [0, n), as the bit pattern of a
// non-negative double (which orders like the value)
__global__ __launch_bounds__(NT) void k_rstat(Par P, const double *r, uint64_t *rmax)
{
    __shared__ unsigned long long mx[NT / WAVE];
    const int c = blockIdx.z;
    r += (int64_t)c * P.ldr;
    unsigned long long m = 0;
    for (int i = blockIdx.x * NT + threadIdx.x; i < P.n; i += gridDim.x * NT) {
        unsigned long long b = (unsigned long long)__double_as_longlong(fabs(r[i]));
        m = b > m ? b : m;
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        unsigned long long o = __shfl_xor(m, off, WAVE);
        m = o > m ? o : m;
    }
    if ((threadIdx.x & 63) == 0) mx[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < NT / WAVE; w++) m = mx[w] > m ? mx[w] : m;
        rmax[(int64_t)c * gridDim.x + blockIdx.x] = m;
    }
}

// Fixed-point exponent e from a maximum max|v| given as the bit pattern of
// a non-negative double: the largest e with rint(max|v| 2^e) <= 127 S,
// S = (256^L - 1) / 255.  L balanced digits in [-128, 127] represent
// exactly the integers of [-128 S, 127 S], so every q = rint(v 2^e) has
// its L digits; e is 8L - 1 - E or, when max|v| sits at the top of its
// binade, 8L - 2 - E (max|v| = f 2^E, 0.5 <= f < 1): 23 or 22 bits of the
// maximum for three limbs.
__device__ __forceinline__ int fixed_exponent(unsigned long long m, int L)
{
    double mv = __longlong_as_double((long long)m);
    if (mv == 0.0) return 0;
    int E;
    frexp(mv, &E);              // mv = f * 2^E, 0.5 <= f < 1
    const double top = 127.0 * ((ldexp(1.0, 8 * L) - 1.0) / 255.0) - 1.0;
    int e = 8 * L - 1 - E;
    if (ldexp(mv, e) > top) e -= 1;
    return e;
}

// Balanced radix-256 digits of q = rint(r * 2^e) in the limb-major rows
// limb_row(k, c), written as MFMA B-operand blocks: Q[((S * NTILE + t) * 4 +
// s) * 64 + lane] (16 bytes each) holds, for lane = l16 + 16 g, the 16 slots
// (order of unpack_word) of row 16 t + l16 at word 16 S + 4 g + s.  One
// workgroup per (super-step S, chain tile ct): thread (wave s, lane l16 +
// 16 g) takes word 16 S + 4 g + s of chain 16 ct + l16, so each wave writes
// whole 1 KB blocks (one per limb); chains past nchains get zero digits.
// Each workgroup reduces the nwg partial maxima of |r| of its 16 chains to
// their exponents e (stored by the workgroups of S = 0) and leaves the exact
// integer sum of each chain's q over its 256 individuals (qsum[c * nq + S],
// nq = super-steps), so that the centring term uses sum_i q_i 2^-e, the sum
// of the very values the products see.
__device__ __forceinline__ void quant_wg(const Par &P, int ntile, const double *r, const uint64_t *rmax, int nwg,
                                         uint4 *Q, int *expo, long long *qsum, int nq, int S, int ct)
{
    constexpr int W = NT / WAVE;
    static_assert(W == 4, "one wave per sub-step s of a super-step");
    __shared__ unsigned long long ml[W][16];
    __shared__ int e_s[16];
    __shared__ long long ql[W][16];
    const int C = P.nchains, ntc = ntile / LIMBS;
    const int lane = threadIdx.x & (WAVE - 1), wave = threadIdx.x >> 6;
    const int l16 = threadIdx.x & 15, j = threadIdx.x >> 4;   // j = 0..15
    const int c = 16 * ct + l16, cc = min(c, C - 1);
    // chain l16's maximum over its partials: thread (j, l16) takes j, j + 16, ...
    unsigned long long m = 0;
    for (int w0 = j; w0 < nwg; w0 += 16 * 4) {
        unsigned long long v[4];
#pragma unroll
        for (int u = 0; u < 4; u++) {
            const int w = w0 + 16 * u;
            v[u] = w < nwg ? rmax[(int64_t)cc * nwg + w] : 0ull;
        }
#pragma unroll
        for (int u = 0; u < 4; u++) m = v[u] > m ? v[u] : m;
    }
#pragma unroll
    for (int off = 16; off < WAVE; off <<= 1) {
        const unsigned long long o = __shfl_xor(m, off, WAVE);
        m = o > m ? o : m;
    }
    if (lane < 16) ml[wave][lane] = m;
    __syncthreads();
    if (threadIdx.x < 16) {
        unsigned long long mm = ml[0][threadIdx.x];
#pragma unroll
        for (int w = 1; w < W; w++) mm = ml[w][threadIdx.x] > mm ? ml[w][threadIdx.x] : mm;
        const int e = fixed_exponent(mm, LIMBS);
        e_s[threadIdx.x] = e;
        if (S == 0 && 16 * ct + (int)threadIdx.x < C) expo[16 * ct + threadIdx.x] = e;
    }
    __syncthreads();
    const int e = e_s[l16];
    const int g = lane >> 4, s = wave;
    const int w = 16 * S + 4 * g + s;   // < ldw: ldw is a whole number of super-steps
    int32_t q[16];
    long long qs = 0;
    if (c < C) {
        const double *rc = r + (int64_t)c * P.ldr + (int64_t)w * 16;
#pragma unroll
        for (int t = 0; t < 16; t++) {
            q[t] = (int32_t)rint(ldexp(rc[t], e));
            qs += q[t];
        }
    } else {
#pragma unroll
        for (int t = 0; t < 16; t++) q[t] = 0;
    }
    for (int k = 0; k < LIMBS; k++) {
        uint32_t out[4];
#pragma unroll
        for (int m4 = 0; m4 < 4; m4++) {
            uint32_t v = 0;
#pragma unroll
            for (int b = 0; b < 4; b++) {
                const int t = 4 * b + m4;                    // individual within the word
                const int8_t d = (int8_t)(q[t] & 0xFF);      // balanced digit in [-128, 127]
                v |= ((uint32_t)(uint8_t)d) << (8 * b);
            }
            out[m4] = v;
        }
#pragma unroll
        for (int t = 0; t < 16; t++) q[t] = (q[t] - (int32_t)(int8_t)(q[t] & 0xFF)) >> 8;
        const int row = limb_row(k, c, ntc);
        Q[(((int64_t)S * ntile + (row >> 4)) * 4 + s) * 64 + (row & 15) + 16 * g] =
            make_uint4(out[0], out[1], out[2], out[3]);
    }
    // exact: |q| < 2^23 and at most ldw * 16 terms; the 16 threads of a
    // chain (four lanes g per wave, four waves)
#pragma unroll
    for (int off = 16; off < WAVE; off <<= 1) qs += __shfl_xor(qs, off, WAVE);
    if (lane < 16) ql[wave][lane] = qs;
    __syncthreads();
    if (threadIdx.x < 16 && 16 * ct + (int)threadIdx.x < C) {
        long long t = ql[0][threadIdx.x];
#pragma unroll
        for (int wv = 1; wv < W; wv++) t += ql[wv][threadIdx.x];
        qsum[(int64_t)(16 * ct + threadIdx.x) * nq + S] = t;
    }
}

__global__ __launch_bounds__(NT) void k_quant(Par P, int ntile, const double *r, const uint64_t *rmax, int nwg,
                                              uint4 *Q, int *expo, long long *qsum)
{
    quant_wg(P, ntile, r, rmax, nwg, Q, expo, qsum, gridDim.x, blockIdx.x, blockIdx.z);
}

// The int8 main loop of the LDS-shared products and updates:
//   acc[it][t] += sum over super-steps S in [S0, S1) of A(it, S) x B(S, t)
// A(it, S): the wave's genotype tile it of super-step S, one 1 KB load per
// wave at abase + (toff[it] + S) * 64 (4 words of 16 2-bit genotypes per
// lane); B(S, .): the super-step's limb slab of NTILE x 4 blocks of 1 KB at
// B + S * sstr * 64 (sstr = blocks per super-step of all chain groups),
// shared by the four waves through LDS.
// Pipelined so that the loads of a super-step are in flight while the
// previous one computes: the genotype tiles of step S + 1 are loaded at the
// start of step S, and each wave stages its quarter of slab S + 2 in
// registers after writing slab S + 1 (loaded one step earlier) to the other
// half of a two-slab LDS ring; one barrier per super-step publishes it (a
// barrier with an LDS-only release, so it waits for no global load).  Ends
// with a barrier: the ring is free.
template <int NTILE, int IT>
__device__ __forceinline__ void limb_mainloop(const uint4 *__restrict__ abase, const int64_t (&toff)[IT],
                                              const uint4 *__restrict__ B, int sstr, int S0, int S1,
                                              uint4 *lds, v4i (&out)[IT][NTILE])
{
    v4i acc[IT][NTILE];
#pragma unroll
    for (int it = 0; it < IT; it++)
#pragma unroll
        for (int t = 0; t < NTILE; t++) acc[it][t] = v4i{0, 0, 0, 0};
    constexpr int W = NT / WAVE;
    constexpr int SLAB = NTILE * 4, NB = SLAB / W;
    static_assert(SLAB % W == 0, "limb blocks per super-step split evenly over the waves");
    const int wave = __builtin_amdgcn_readfirstlane((int)threadIdx.x >> 6), lane = threadIdx.x & 63;
    if (S0 >= S1) {
        __syncthreads();
#pragma unroll
        for (int it = 0; it < IT; it++)
#pragma unroll
            for (int t = 0; t < NTILE; t++) out[it][t] = acc[it][t];
        return;
    }
    const int Sl = S1 - 1;
    // operands as native 4 x 32-bit vectors (register arrays of HIP's uint4
    // struct that a loop carries are not always kept in registers)
    const v4i *Bw = reinterpret_cast<const v4i *>(B) + (int64_t)wave * 64 + lane;   // blocks wave + W j
    const v4i *Av = reinterpret_cast<const v4i *>(abase);
    v4i a[IT], an[IT], bs[NB];
#pragma unroll
    for (int j = 0; j < NB; j++) bs[j] = Bw[((int64_t)S0 * sstr + W * j) * 64];
#pragma unroll
    for (int it = 0; it < IT; it++) a[it] = Av[(toff[it] + S0) * 64];
    v4i *Lw = reinterpret_cast<v4i *>(lds) + wave * 64 + lane;
#pragma unroll
    for (int j = 0; j < NB; j++) Lw[W * j * 64] = bs[j];
    {
        const int Sb = min(S0 + 1, Sl);
#pragma unroll
        for (int j = 0; j < NB; j++) bs[j] = Bw[((int64_t)Sb * sstr + W * j) * 64];
    }
    __syncthreads();
    // step S: A(S) in a, A(S + 1) loading into an; B(S) in LDS half
    // (S - S0) & 1, B(S + 1) loading into bs.  Every load is issued (past the
    // range it re-reads step Sl, an L2 hit), so that the compiler counts the
    // loads in flight exactly and waits for no more than the operand at hand.
    for (int S = S0; S < S1; S++) {
        const int buf = (S - S0) & 1;
        const int Sa = min(S + 1, Sl), Sb = min(S + 2, Sl);
#pragma unroll
        for (int it = 0; it < IT; it++) an[it] = Av[(toff[it] + Sa) * 64];
        const v4i *Bl = reinterpret_cast<const v4i *>(lds) + buf * SLAB * 64 + lane;
#pragma unroll
        for (int s = 0; s < 4; s++) {
            v4i b[NTILE];
#pragma unroll
            for (int t = 0; t < NTILE; t++) b[t] = Bl[(t * 4 + s) * 64];
            v4i av[IT];
            unpack_tiles<IT>(a, s, av);
#pragma unroll
            for (int it = 0; it < IT; it++)
#pragma unroll
                for (int t = 0; t < NTILE; t++)
                    acc[it][t] = __builtin_amdgcn_mfma_i32_16x16x64_i8(av[it], b[t], acc[it][t], 0, 0, 0);
        }
        v4i *Ln = Lw + (buf ^ 1) * SLAB * 64;
#pragma unroll
        for (int j = 0; j < NB; j++) Ln[W * j * 64] = bs[j];
#pragma unroll
        for (int j = 0; j < NB; j++) bs[j] = Bw[((int64_t)Sb * sstr + W * j) * 64];
        __syncthreads();
#pragma unroll
        for (int it = 0; it < IT; it++) a[it] = an[it];
    }
#pragma unroll
    for (int it = 0; it < IT; it++)
#pragma unroll
        for (int t = 0; t < NTILE; t++) out[it][t] = acc[it][t];
}

// Products with a block's upper factor U on int8 MFMA, both operands as
// balanced radix-256 limbs with exact int32 sums:
//   mode 0 (rhs):    part[g][c][j] = sum over rows i <= j of group g of U(i, j) r_a(i, c)
//   mode 1 (update): part[g][c][i] = sum over columns j >= i of group g of U(i, j) s_j c~(j, c)
// U is held as ULIMBS limbs of rint(U 2^e) with one exponent per column j
// (mode 0) or per row i of U diag(s) (mode 1), set from that column's or
// row's largest entry (31 bits of it), so every entry keeps its absolute
// precision 2^-31 max.  r_a has three limbs per chain at one exponent per
// chain, from the chain's max |r_a| over the block's rows (each pair
// quantises the rows it needs); mode 1 multiplies with the limbs of
// c~ = delta / s + carry that the genotype update applies (k_dquant), which
// is why its U columns carry s_j.  A group is 64 G rows (mode 0) or
// columns (mode 1), the row split of the partial sums that the draw and the
// r_a reduction add up.  The pairs run as the trailing workgroups of the
// genotype launches (mode 0 after X_s'Q, mode 1 after the residual update),
// where they fill the CUs that the genotype workgroups leave idle.
//
// K steps of one MFMA (64 values of the summed index):
//   mode 0: row group I4 (rows 64 I4 + 16 g + b, byte b of lane group g);
//   mode 1: (super-step ss, sub-step s): columns 16 (16 ss + 4 g + s) + m + 4 b
//           at byte 4 m + b (the slot order of the genotype words and Dq).
// A operand tiles (1 KB, lane l16 + 16 g, 16 bytes), per block:
//   mode 0: Uq0[((off0(J) + I4) * ULIMBS + ku) * 64 + lane], column j = 16 J + l16;
//   mode 1: Uq1[(((off1(I) + ss - I / 16) * 4 + s) * ULIMBS + ku) * 64 + lane], row i = 16 I + l16.
// B operand blocks (limb-major rows, as the genotype products):
//   mode 0: built in LDS from r_a (block (I4, t), lane l16 + 16 g: rows 64 I4 + 16 g + b);
//   mode 1: Dq[((ss * NTILE + t) * 4 + s) * 64 + lane].
constexpr int ULIMBS = 4;   // limbs of U (31 bits)
__host__ __device__ __forceinline__ int64_t uq_off0(int J)
{
    const int64_t a = J >> 2, r = J & 3;   // sum over J' < J of (J' / 4 + 1)
    return (a + 1) * (2 * a + r);
}
// pairs before quad bx (G K steps per group): mode 0, bx / G + 1 row groups per quad
__host__ __device__ __forceinline__ int64_t uq_pairs0(int bx, int G)
{
    const int64_t a = bx / G, r = bx % G;   // sum over x < bx of (x / G + 1)
    return (a + 1) * ((int64_t)G * a / 2 + r);
}
// mode 1: ng - bx / G column groups per quad (ng groups in all)
__host__ __device__ __forceinline__ int64_t uq_pairs1(int bx, int ng, int G)
{
    return (int64_t)bx * (ng + 1) - uq_pairs0(bx, G);
}
// (quad, group) pairs of a block of nbt tiles, nss super-steps
__host__ __device__ __forceinline__ int64_t uq_pairs(int nbt, int nss, int mode, int G)
{
    const int nq4 = (nbt + 3) / 4;
    return mode == 0 ? uq_pairs0(nq4, G) : uq_pairs1(nq4, (nss * 4 + G - 1) / G, G);
}
__host__ __device__ __forceinline__ int64_t uq_off1(int I, int nss)
{
    const int64_t q = I >> 4, r = I & 15;  // sum over I' < I of (nss - I' / 16)
    return (int64_t)I * nss - (8 * q * (q - 1) + r * q);
}

// U products of one block over (quad, group) pairs, pair index k:
//   mode 0: quad bx = output columns J = 4 bx + w (wave w), row group gi:
//           K steps I4 in [G gi, min(G gi + G, bx + 1));
//   mode 1: quad bx = output rows I = 4 bx + w, column group gi: K steps
//           kk = 4 ss + s in [max(G gi, 4 (bx / 4)), min(G gi + G, 4 nss))
// (G = K steps per group: 8 or 16).  The B blocks go to LDS eight K steps
// at a time (so a pair of 8 K steps fills LDS once, between its loads and
// one barrier), and each wave holds the ULIMBS A tiles of four K steps in
// registers (loaded together, the next four's before any barrier).  Waves
// past the block still load and meet the barriers, but do no MFMA and store
// nothing.
constexpr int uq_lds(int ntile) { return 8 * ntile * 64; }   // uint4 of LDS per pair: eight K steps of B

template <int MODE, int NTILE>
__device__ __forceinline__ void uq_pair(const Par &P, const uint4 *__restrict__ Ua, int nbt, int nb, int nss,
                                        const uint4 *__restrict__ Bq, int ntot, int cbase,
                                        const double *__restrict__ ra,
                                        const unsigned long long *__restrict__ amax, const int *__restrict__ ue,
                                        const int *__restrict__ ce, double *__restrict__ part, int k,
                                        int G, uint4 *lds)
{
    constexpr int W = NT / WAVE, NTC = NTILE / LIMBS;
    constexpr int KR = 4, KB = 8;   // K steps per round of A tiles (registers), of B blocks (LDS)
    const int wave = __builtin_amdgcn_readfirstlane((int)threadIdx.x >> 6), lane = threadIdx.x & 63;
    const int g = lane >> 4, l16 = lane & 15;
    int bx, gi, k0, k1;
    if (MODE == 0) {
        // pairs of quad bx: bx / G + 1 groups; uq_pairs0(bx, G) pairs before it
        bx = (int)sqrtf(2.0f * G * (float)k);
        while (bx > 0 && uq_pairs0(bx, G) > k) bx--;
        while (uq_pairs0(bx + 1, G) <= k) bx++;
        gi = (int)(k - uq_pairs0(bx, G));
        k0 = G * gi;
        k1 = min(G * gi + G, bx + 1);
    } else {
        // pairs of quad bx: the column groups (G / 4 super-steps each)
        // from the one holding super-step bx / 4
        const int ng = (nss * 4 + G - 1) / G;
        bx = 0;
        while (uq_pairs1(bx + 1, ng, G) <= k) bx++;
        gi = bx / G + (int)(k - uq_pairs1(bx, ng, G));
        k0 = max(G * gi, 4 * (bx / 4));
        k1 = min(G * gi + G, 4 * nss);
    }
    const int O = 4 * bx + wave;      // the wave's output tile (J or I)
    const bool live = O < nbt;
    const int Oc = min(O, nbt - 1);
    // A tile of K step kk, limb ku
    auto aidx = [&](int kk, int ku) -> int64_t {
        if (MODE == 0) return ((uq_off0(Oc) + kk) * ULIMBS + ku) * 64 + lane;
        return (((uq_off1(Oc, nss) + (kk >> 2) - Oc / 16) * 4 + (kk & 3)) * ULIMBS + ku) * 64 + lane;
    };
    // B block t of K step kk (mode 1: Dq, whose super-steps hold ntot tiles;
    // Bq points at this chain group's first tile)
    auto bidx = [&](int kk, int t) -> int64_t {
        return (((int64_t)(kk >> 2) * ntot + t) * 4 + (kk & 3)) * 64 + lane;
    };
    // The pair's K steps in rounds of four: each wave holds the A tiles of one
    // round in registers (loaded together); the B blocks of eight K steps go
    // to LDS at a time (each wave fills every fourth block, its loads of four
    // K steps before their stores).
    const v4i *Av = reinterpret_cast<const v4i *>(Ua);
    const v4i *Bv = reinterpret_cast<const v4i *>(Bq);
    v4i *L = reinterpret_cast<v4i *>(lds);
    const int nk = k1 - k0;
    v4i a[KR][ULIMBS];
#pragma unroll
    for (int s = 0; s < KR; s++)
#pragma unroll
        for (int u = 0; u < ULIMBS; u++) a[s][u] = Av[aidx(k0 + min(s, nk - 1), u)];
    const int C = P.nchains;
    int ea[NTC];   // mode 0: the exponent of chain cbase + 16 tc + l16's r_a limbs
#pragma unroll
    for (int tc = 0; tc < NTC; tc++)
        ea[tc] = MODE == 0 ? fixed_exponent(amax[min(cbase + 16 * tc + l16, C - 1)], LIMBS) : 0;
    // the B blocks of K steps k0 + KB rb + s, s < KB, into LDS block s
    auto fill = [&](int rb) {
        const int kb = k0 + KB * rb, nkr = min(KB, k1 - kb);
        if (MODE == 0) {
            // B from r_a itself, as LIMBS balanced limbs of rint(r_a 2^e) with
            // e from the chain's max |r_a| over the block's rows (zero past
            // the block's rows and the chains).  Unit (s, g4, tc): lane l16 +
            // 16 q takes rows 64 (kb + s) + 16 g4 + 4 q .. + 3 of chain
            // cbase + 16 tc + l16 (32 contiguous bytes) and writes dword q of
            // its limbs' B entries (lane l16 + 16 g4 of blocks (s, kk NTC + tc)).
            uint32_t *Lw = reinterpret_cast<uint32_t *>(lds);
            const int q = lane >> 4;
            for (int u = wave; u < nkr * 4 * NTC; u += W) {   // wave-uniform
                const int tc = u % NTC, sg = u / NTC, s = sg >> 2, g4 = sg & 3;
                const int c = cbase + 16 * tc + l16, i0 = 64 * (kb + s) + 16 * g4 + 4 * q;
                const double *rp = ra + (int64_t)min(c, C - 1) * P.blocksize + i0;
                double v[4];
                if (c < C && i0 + 4 <= nb && (P.blocksize & 1) == 0) {
                    const double2 x = reinterpret_cast<const double2 *>(rp)[0];
                    const double2 y = reinterpret_cast<const double2 *>(rp)[1];
                    v[0] = x.x;
                    v[1] = x.y;
                    v[2] = y.x;
                    v[3] = y.y;
                } else {
#pragma unroll
                    for (int b = 0; b < 4; b++) v[b] = (c < C && i0 + b < nb) ? rp[b] : 0.0;
                }
                int e = ea[0];
#pragma unroll
                for (int t = 1; t < NTC; t++) e = tc == t ? ea[t] : e;
                uint32_t w[LIMBS] = {};
#pragma unroll
                for (int b = 0; b < 4; b++) {
                    int32_t qv = (int32_t)rint(ldexp(v[b], e));
#pragma unroll
                    for (int kk = 0; kk < LIMBS; kk++) {
                        const int32_t d = (int32_t)(int8_t)(qv & 0xFF);
                        w[kk] |= (uint32_t)(uint8_t)d << (8 * b);
                        qv = (qv - d) >> 8;
                    }
                }
#pragma unroll
                for (int kk = 0; kk < LIMBS; kk++)
                    Lw[((s * NTILE + kk * NTC + tc) * 64 + l16 + 16 * g4) * 4 + q] = w[kk];
            }
        } else {
            constexpr int NBQ = (KR * NTILE + W - 1) / W;   // B blocks per wave and four K steps
#pragma unroll
            for (int h = 0; h < KB / KR; h++) {
                if (KR * h >= nkr) break;   // workgroup-uniform
                v4i bq[NBQ];
#pragma unroll
                for (int j = 0; j < NBQ; j++) {
                    const int q = KR * NTILE * h + wave + W * j, s = min(q / NTILE, nkr - 1), t = q % NTILE;
                    bq[j] = Bv[bidx(kb + s, t)];
                }
#pragma unroll
                for (int j = 0; j < NBQ; j++) {
                    const int q = KR * NTILE * h + wave + W * j;
                    if (wave + W * j < KR * NTILE && q < nkr * NTILE) L[q * 64 + lane] = bq[j];
                }
            }
        }
    };
    // one accumulator per chain tile and weight 256^(ku + kr): the limb
    // products of equal weight are summed exactly in int32 (at most three
    // products of 64 terms of at most 2^14 per K step, G <= 16 K steps)
    constexpr int NWT = ULIMBS + LIMBS - 1;
    v4i acc[NWT][NTC];
#pragma unroll
    for (int w = 0; w < NWT; w++)
#pragma unroll
        for (int tc = 0; tc < NTC; tc++) acc[w][tc] = v4i{0, 0, 0, 0};
    fill(0);
    __syncthreads();
    for (int r = 0; KR * r < nk; r++) {   // workgroup-uniform
        if (r > 0) {
#pragma unroll
            for (int s = 0; s < KR; s++)
#pragma unroll
                for (int u = 0; u < ULIMBS; u++) a[s][u] = Av[aidx(k0 + min(KR * r + s, nk - 1), u)];
            if ((KR * r) % KB == 0) {
                __syncthreads();   // every wave is done with the previous B blocks
                fill(KR * r / KB);
                __syncthreads();
            }
        }
        if (live) {
            const int sb = (KR * r) % KB;   // LDS block of the round's first K step
#pragma unroll
            for (int s = 0; s < KR; s++) {
                if (KR * r + s >= nk) break;
                const v4i *Lb = L + (sb + s) * NTILE * 64 + lane;
#pragma unroll
                for (int kr = 0; kr < LIMBS; kr++)
#pragma unroll
                    for (int tc = 0; tc < NTC; tc++) {
                        const v4i bt = Lb[(kr * NTC + tc) * 64];
#pragma unroll
                        for (int u = 0; u < ULIMBS; u++)
                            acc[u + kr][tc] =
                                __builtin_amdgcn_mfma_i32_16x16x64_i8(a[s][u], bt, acc[u + kr][tc], 0, 0, 0);
                    }
            }
        }
    }
    if (!live) return;
    // D[m][n] of the 16x16 MFMA: m = 4 g + e (column j in mode 0, row i in
    // mode 1), n = l16 (chain 16 tc + l16); the weights combined in FP64
    // (Horner from the top weight), then scaled by 2^-(e_U + e_chain).
#pragma unroll
    for (int tc = 0; tc < NTC; tc++) {
        const int c = cbase + 16 * tc + l16;
        if (c >= C) continue;
        const int ec = MODE == 0 ? ea[tc] : ce[c];
#pragma unroll
        for (int e = 0; e < 4; e++) {
            const int o = 16 * O + 4 * g + e;
            if (o >= nb) continue;
            double S = (double)acc[NWT - 1][tc][e];
#pragma unroll
            for (int w = NWT - 2; w >= 0; w--) S = S * 256.0 + (double)acc[w][tc][e];
            part[((int64_t)gi * C + c) * P.blocksize + o] = ldexp(S, -(ue[o] + ec));
        }
    }
}


// X_s'Q of one block, by (row group, chunk) workgroups, id = rg + nrg * ch:
//   Tp[ch][c][jk] = sum over the chunk's individuals i and the limbs k of
//                   256^k x_(p0+jk, i) q_k(i, c),
// one exact int64 per (chunk, chain, SNP) (columns of length blocksize).
// The four waves take RW consecutive SNP tiles each (a workgroup covers
// 64 RW SNPs) and walk the chunk's super-steps together (limb_mainloop), so
// each super-step's limb slab is read once per workgroup.  Chunk ch covers
// super-steps [ch nS / nch, (ch + 1) nS / nch) of 256 individuals.  A lane
// holds the three limbs of its chains in the same accumulator position of
// column tiles tc, ntc + tc, 2 ntc + tc and forms the int64 value itself.
// Tiles past the block re-read its last tile and store nothing.  The
// workgroups after the nrg nch of X_s'Q are the pairs of U_s'r_a (uq_pair,
// mode 0), which quantise r_a themselves with the chain exponents from
// amax (max |r_a| bits per chain over the block's rows); workgroup 0 also
// clears anext, where the r_a reduction after this block's update collects
// the next block's maxima.
// Register budget for three waves per SIMD up to 12 accumulator tiles per
// wave (RW x NTILE), two beyond: the compiler then keeps the accumulators in
// VGPRs, without spills, and three workgroups share a CU's LDS.
template <int NTILE, int RW>
__global__ __launch_bounds__(NT, RW * NTILE <= 12 ? 3 : 2) void k_prod(Par P, const uint4 *__restrict__ gsw, int nbt, int nb,
                                             const uint4 *__restrict__ Q, int nch, long long *__restrict__ Tp,
                                             int64_t tstride, int nrg, int ngrp, int npair,
                                             const uint4 *__restrict__ Ua, const double *__restrict__ ra,
                                             const unsigned long long *__restrict__ amax,
                                             unsigned long long *__restrict__ anext, const int *__restrict__ ue,
                                             double *__restrict__ upart, int ug)
{
    constexpr int W = NT / WAVE, SLAB = NTILE * 4, NTC = NTILE / LIMBS;
    static_assert(2 * SLAB * 64 >= uq_lds(NTILE), "LDS of the U pairs");
    __shared__ uint4 lds[2 * SLAB * 64];
    const int id = blockIdx.x, nx = nrg * nch;
    if (id == 0 && (int)threadIdx.x < P.nchains) anext[threadIdx.x] = 0;
    if (id >= ngrp * nx) {   // U_s'r_a pairs, chain group by chain group
        const int k = id - ngrp * nx, cg = k / npair;
        uq_pair<0, NTILE>(P, Ua, nbt, nb, (nb + 255) / 256, nullptr, 0, cg * 16 * NTC, ra, amax, ue, nullptr,
                          upart, k - cg * npair, ug, lds);
        return;
    }
    const int wave = __builtin_amdgcn_readfirstlane((int)threadIdx.x >> 6), lane = threadIdx.x & 63;
    const int g = lane >> 4, l16 = lane & 15;
    const int cg = id / nx, rg = (id - cg * nx) % nrg, ch = (id - cg * nx) / nrg;   // chain group, row group, chunk
    const int cbase = cg * 16 * NTC;
    const int nS = P.ldw / 16;
    const int S0 = (int)((int64_t)ch * nS / nch), S1 = (int)((int64_t)(ch + 1) * nS / nch);
    const int tile0 = (rg * W + wave) * RW;
    int64_t toff[RW];
#pragma unroll
    for (int it = 0; it < RW; it++) toff[it] = (int64_t)min(tile0 + it, nbt - 1) * nS;
    v4i acc[RW][NTILE];
    limb_mainloop<NTILE, RW>(gsw + lane, toff, Q + (int64_t)cg * SLAB * 64, ngrp * SLAB, S0, S1, lds, acc);
    const int C = P.nchains;
    long long *out = Tp + (int64_t)ch * tstride;
#pragma unroll
    for (int it = 0; it < RW; it++) {
        const int jk0 = 16 * (tile0 + it) + 4 * g;
#pragma unroll
        for (int tc = 0; tc < NTC; tc++) {
            const int c = cbase + 16 * tc + l16;
            if (c >= C) continue;
            long long *o = out + (int64_t)c * P.blocksize + jk0;
#pragma unroll
            for (int e = 0; e < 4; e++)
                if (jk0 + e < nb)
                    o[e] = (long long)acc[it][tc][e] + 256LL * acc[it][NTC + tc][e] +
                           65536LL * acc[it][2 * NTC + tc][e];
        }
    }
}

// Limb exponents of U, per block: ue0[j] from max over i <= j of |U(i, j)|,
// ue1[i] from max over j >= i of |U(i, j) s_j| (ULIMBS limbs, 31 bits).  One
// thread per (index k, block): column k and row k of the packed triangle.
__global__ __launch_bounds__(NT) void k_uexp(const double *U, const int64_t *tri_off, const int32_t *ls,
                                             const double *sdv, int bs, int *ue0, int *ue1)
{
    const int b = blockIdx.y, p0 = ls[b], nb = ls[b + 1] - p0;
    const int k = blockIdx.x * NT + threadIdx.x;
    if (k >= nb) return;
    const double *Ub = U + tri_off[b];
    unsigned long long mc = 0, mr = 0;
    const double *col = Ub + (int64_t)k * (k + 1) / 2;
    for (int i = 0; i <= k; i++) {
        const unsigned long long u = (unsigned long long)__double_as_longlong(fabs(col[i]));
        mc = u > mc ? u : mc;
    }
    for (int j = k; j < nb; j++) {
        const double v = Ub[(int64_t)j * (j + 1) / 2 + k] * sdv[p0 + j];
        const unsigned long long u = (unsigned long long)__double_as_longlong(fabs(v));
        mr = u > mr ? u : mr;
    }
    ue0[(int64_t)b * bs + k] = fixed_exponent(mc, ULIMBS);
    ue1[(int64_t)b * bs + k] = fixed_exponent(mr, ULIMBS);
}

// The ULIMBS limb tiles of one block's U: one thread per (tile entry, lane).
// Mode 0 entries (J, I4), mode 1 entries (I, ss, s); see uq_pair for the byte
// order.  Entries past the triangle or the block are zero.
__global__ __launch_bounds__(NT) void k_pack_uq(const double *U, int64_t tri0, int nb, int nbt, int nss, int p0,
                                                const double *sdv, const int *ue0, const int *ue1, int mode,
                                                uint4 *__restrict__ Uq)
{
    const int64_t tid = (int64_t)blockIdx.x * NT + threadIdx.x;
    const int64_t nent = mode == 0 ? uq_off0(nbt) : uq_off1(nbt, nss) * 4;
    if (tid >= nent * 64) return;
    const int64_t ent = tid >> 6;
    const int lane = (int)(tid & 63), g = lane >> 4, l16 = lane & 15;
    auto u = [&](int i, int j) -> double {
        return (i <= j && j < nb) ? U[tri0 + (int64_t)j * (j + 1) / 2 + i] : 0.0;
    };
    int32_t q[16];
    if (mode == 0) {
        int J = (int)sqrtf(2.0f * (float)ent);
        while (J > 0 && uq_off0(J) > ent) J--;
        while (uq_off0(J + 1) <= ent) J++;
        const int I4 = (int)(ent - uq_off0(J));
        const int j = 16 * J + l16, e = j < nb ? ue0[j] : 0;
#pragma unroll
        for (int b = 0; b < 16; b++) q[b] = (int32_t)rint(ldexp(u(64 * I4 + 16 * g + b, j), e));
    } else {
        const int64_t e2 = ent >> 2;   // (I, ss) entry
        const int s = (int)(ent & 3);
        int I = 0;
        {
            int lo = 0, hi = nbt;   // largest I with uq_off1(I) <= e2
            while (hi - lo > 1) {
                const int mid = (lo + hi) / 2;
                if (uq_off1(mid, nss) <= e2) lo = mid; else hi = mid;
            }
            I = lo;
        }
        const int ss = I / 16 + (int)(e2 - uq_off1(I, nss));
        const int i = 16 * I + l16, e = i < nb ? ue1[i] : 0;
        const int sw = 16 * ss + 4 * g + s;
#pragma unroll
        for (int bb = 0; bb < 16; bb++) {
            const int m = bb >> 2, b = bb & 3;   // byte 4 m + b holds column 16 sw + m + 4 b
            const int j = 16 * sw + m + 4 * b;
            q[bb] = (int32_t)rint(ldexp(j < nb ? u(i, j) * sdv[p0 + j] : 0.0, e));
        }
    }
    for (int ku = 0; ku < ULIMBS; ku++) {
        uint32_t w4[4];
#pragma unroll
        for (int d = 0; d < 4; d++) {
            uint32_t v = 0;
#pragma unroll
            for (int b = 0; b < 4; b++) v |= ((uint32_t)(uint8_t)(int8_t)(q[4 * d + b] & 0xFF)) << (8 * b);
            w4[d] = v;
        }
#pragma unroll
        for (int b = 0; b < 16; b++) q[b] = (q[b] - (int32_t)(int8_t)(q[b] & 0xFF)) >> 8;
        Uq[(ent * ULIMBS + ku) * 64 + lane] = make_uint4(w4[0], w4[1], w4[2], w4[3]);
    }
}

// max over the workgroup of m (any value from threads past the data is 0)
__device__ __forceinline__ unsigned long long wg_max(unsigned long long m)
{
    __shared__ unsigned long long ml[NT / WAVE];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        const unsigned long long o = __shfl_xor(m, off, WAVE);
        m = o > m ? o : m;
    }
    if ((threadIdx.x & 63) == 0) ml[threadIdx.x >> 6] = m;
    __syncthreads();
    m = ml[0];
#pragma unroll
    for (int w = 1; w < NT / WAVE; w++) m = ml[w] > m ? ml[w] : m;
    return m;
}

// r_a(i, c) += sum of the mode-1 partials of the column groups that reach row
// i's tile (groups left of the diagonal wrote nothing), rows i < nb.  With
// amax, also the max |r_a(i, c)| bits over the next block's rows i < nbn
// (atomic max into amax[c], cleared by the launch of the block's products).
__device__ __forceinline__ void ra_reduce_wg(const Par &P, const double *part, int spant, int ngroups, int nb,
                                             double *ra, int bx, int c, int nbn, unsigned long long *amax)
{
    const int i = bx * NT + threadIdx.x;
    const int C = P.nchains;
    const int64_t gstride = (int64_t)C * P.blocksize;
    double *rc = ra + (int64_t)c * P.blocksize;
    double v = 0.0;
    if (i < nb) {
        const double *pp = part + (int64_t)c * P.blocksize + i;
        double s = 0.0;
        for (int g0 = (i >> 4) / spant; g0 < ngroups; g0 += 8) {
            double t[8];
#pragma unroll
            for (int u = 0; u < 8; u++) t[u] = g0 + u < ngroups ? pp[(g0 + u) * gstride] : 0.0;
#pragma unroll
            for (int u = 0; u < 8; u++) s += t[u];
        }
        v = rc[i] + s;
        rc[i] = v;
    } else if (i < nbn) {
        v = rc[i];
    }
    if (amax) {   // workgroup-uniform
        const unsigned long long m = wg_max(i < nbn ? (unsigned long long)__double_as_longlong(fabs(v)) : 0ull);
        if (threadIdx.x == 0) atomicMax(amax + c, m);
    }
}

__global__ __launch_bounds__(NT) void k_ra_reduce(Par P, const double *part, int spant, int ngroups, int nb,
                                                  double *ra)
{
    ra_reduce_wg(P, part, spant, ngroups, nb, ra, blockIdx.x, blockIdx.y, 0, nullptr);
}

// amax[c] = max |r_a(i, c)| bits over rows i < nb, one workgroup per chain
// (r_a set outside the batched update: the first block of a sweep)
__global__ __launch_bounds__(NT) void k_ramax(Par P, const double *ra, int nb, unsigned long long *amax)
{
    const int c = blockIdx.x;
    const double *rc = ra + (int64_t)c * P.blocksize;
    unsigned long long m = 0;
    for (int i0 = 0; i0 < nb; i0 += 16 * NT) {   // 16 loads in flight per thread
        double v[16];
#pragma unroll
        for (int u = 0; u < 16; u++) v[u] = rc[min(i0 + u * NT + (int)threadIdx.x, nb - 1)];
#pragma unroll
        for (int u = 0; u < 16; u++) {
            const unsigned long long b = (unsigned long long)__double_as_longlong(fabs(v[u]));
            m = b > m ? b : m;
        }
    }
    m = wg_max(m);
    if (threadIdx.x == 0) amax[c] = m;
}

// The r_a reduction of this block beside the r limbs of the next block (both
// need only the update just done): workgroups [0, nq * chain tiles) quantise r,
// the rest (nra per chain, covering max(nb, nbn) rows) reduce r_a and
// collect the next block's max |r_a| (nbn rows) into amax.
__global__ __launch_bounds__(NT) void k_rared_quant(Par P, const double *part, int spant, int ngroups, int nb,
                                                    double *ra, int nra, int ntile, const double *r,
                                                    const uint64_t *rmax, int nwg, uint4 *Q, int *expo,
                                                    long long *qsum, int nq, int nbn, unsigned long long *amax)
{
    const int id = blockIdx.x, nqt = nq * (ntile / LIMBS);
    if (id < nqt)
        quant_wg(P, ntile, r, rmax, nwg, Q, expo, qsum, nq, id % nq, id / nq);
    else
        ra_reduce_wg(P, part, spant, ngroups, nb, ra, (id - nqt) % nra, (id - nqt) / nra, nbn, amax);
}

// SNP-major genotype tiles of one block for k_xq: element ((tile * nS + S) *
// 64 + lane), lane = l16 + 16 g, holds words 16 S + 4 g .. + 3 of SNP
// 16 tile + l16 (zero beyond the block).
__global__ void k_swz_geno(const uint32_t *__restrict__ geno, int ldw, int p0, int nb, uint4 *__restrict__ gsw)
{
    const int nS = ldw / 16;
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;  // (S, lane) within the tile
    if (e >= (int64_t)nS * 64) return;
    const int tile = blockIdx.y, S = (int)(e >> 6), lane = (int)(e & 63);
    const int snp = 16 * tile + (lane & 15);
    uint4 v = make_uint4(0, 0, 0, 0);
    if (snp < nb) v = *reinterpret_cast<const uint4 *>(geno + (int64_t)(p0 + snp) * ldw + 16 * S + 4 * (lane >> 4));
    gsw[((int64_t)tile * nS + S) * 64 + lane] = v;
}

// Individual-major genotype tiles of one block for k_xupd: SNP word sw
// holds the 2-bit counts of block positions 16 sw .. 16 sw + 15 of one
// individual (position t at bits 2t).  Element ((itile * nss + ss) * 64 +
// lane), lane = l16 + 16 g, holds words 16 ss + 4 g .. + 3 of individual
// 16 itile + l16 (positions >= nb are 0).  One thread per (16-individual
// word iw = itile, SNP word sw).
__global__ void k_transpose(const uint32_t *__restrict__ geno, int ldw, int p0, int nb, int nss,
                            uint32_t *__restrict__ gt)
{
    const int sw = blockIdx.y;
    const int iw = blockIdx.x * blockDim.x + threadIdx.x;
    if (iw >= ldw) return;
    uint32_t in[16];
#pragma unroll
    for (int t = 0; t < 16; t++) {
        int pos = 16 * sw + t;
        in[t] = pos < nb ? geno[(int64_t)(p0 + pos) * ldw + iw] : 0u;
    }
    const int ss = sw >> 4, g = (sw >> 2) & 3, s = sw & 3;
    uint32_t *dst = gt + (((int64_t)iw * nss + ss) * 64 + 16 * g) * 4 + s;
#pragma unroll
    for (int li = 0; li < 16; li++) {
        uint32_t w = 0;
#pragma unroll
        for (int t = 0; t < 16; t++) w |= ((in[t] >> (2 * li)) & 3u) << (2 * t);
        dst[4 * li] = w;
    }
}

// One thread per (block position jk, chain): c_j (delta_j / s_j plus the
// SNP's carried remainder, from k_draw_b) quantised to L balanced int8
// limbs, q_j = rint(c_j 2^e), written as B-operand blocks like Q:
// Dq[((ss * NTILE + t) * 4 + s) * 64 + l16 + 16 g] for limb row 16 t + l16 and
// SNP word sw = 16 ss + 4 g + s (zero beyond the block).  The 16 positions
// of a word sit in 16 consecutive lanes; dword m of the word's operand holds
// the digits of positions m, m+4, m+8, m+12 (slot order of unpack_word), so
// lane m of the group gathers them by shuffles.  The exponent comes from the
// per-workgroup maxima of |c_j| that k_draw_b left; position 0 stores it.
// What the update applies is c~_j = q_j 2^-e: the remainder c_j - c~_j goes
// back to carry_j for the SNP's next update, and each workgroup leaves its
// part of the centring constant K = sum_j 2 f_j c~_j (kpart), formed from
// the same c~_j, so that r moves by exactly sum_j (x_ij - 2 f_j) c~_j.
// It also clears the RSLOT max |r| slots of each chain that the update
// launched next folds its maxima into.
__global__ __launch_bounds__(NT) void k_dquant(Par P, int ntile, const double *cjv, const double *f2,
                                               int p0, int nb, const uint64_t *dmax, int nwg_d, uint4 *Dq,
                                               int *expo, double *carry, double *kpart, uint64_t *rslot)
{
    __shared__ double lds[NT / WAVE];
    const int c = blockIdx.y;
    const int jk = blockIdx.x * NT + threadIdx.x;   // the grid covers whole words
    if (blockIdx.x == 0 && threadIdx.x < RSLOT) rslot[(int64_t)c * RSLOT + threadIdx.x] = 0;   // for the update
    unsigned long long m = 0;
    for (int w0 = 0; w0 < nwg_d; w0 += 8) {
        unsigned long long v[8];
#pragma unroll
        for (int u = 0; u < 8; u++) v[u] = w0 + u < nwg_d ? dmax[(int64_t)c * nwg_d + w0 + u] : 0ull;
#pragma unroll
        for (int u = 0; u < 8; u++) m = v[u] > m ? v[u] : m;
    }
    const int e = fixed_exponent(m, LIMBS);
    if (jk == 0) expo[c] = e;
    const double cj = jk < nb ? cjv[(int64_t)c * P.blocksize + jk] : 0.0;
    int32_t q = (int32_t)rint(ldexp(cj, e));
    double kq = 0.0;
    if (jk < nb) {
        const double cq = ldexp((double)q, -e);
        carry[(int64_t)c * P.p + p0 + jk] = cj - cq;
        kq = f2[p0 + jk] * cq;
    }
    const int t = jk & 15, sw = jk >> 4;
    const int ss = sw >> 4, g = (sw >> 2) & 3, s = sw & 3;
    const int src = (threadIdx.x & (WAVE - 1) & ~15) + (t & 3);
    for (int k = 0; k < LIMBS; k++) {
        const int32_t d = (int32_t)(int8_t)(q & 0xFF);   // balanced digit in [-128, 127]
        q = (q - d) >> 8;
        const uint32_t byte = (uint32_t)(uint8_t)d;
        const uint32_t v = __shfl(byte, src, WAVE) | (__shfl(byte, src + 4, WAVE) << 8) |
                           (__shfl(byte, src + 8, WAVE) << 16) | (__shfl(byte, src + 12, WAVE) << 24);
        if (t < 4) {
            const int row = limb_row(k, c, ntile / LIMBS);
            reinterpret_cast<uint32_t *>(Dq + (((int64_t)ss * ntile + (row >> 4)) * 4 + s) * 64 + (row & 15) +
                                         16 * g)[t] = v;
        }
    }
    kq = block_sum<NT>(kq, lds);
    if (threadIdx.x == 0) kpart[(int64_t)c * gridDim.x + blockIdx.x] = kq;
}

// Centring constants K_c of the update: every chain's nkp workgroup parts
// from k_dquant.  The update kernels load them into registers before their
// main loop (KPre: up to 4 per thread, the first 4 * NT parts), so that no
// wave waits for them, and sum them in workgroup order into LDS after it
// (kconst_finish; the caller's next barrier publishes kc_s).
struct KPre {
    double v[4];
};
__device__ __forceinline__ KPre kconst_prefetch(const double *__restrict__ kpart, int nkp, int C)
{
    KPre k;
    const int tot = nkp * C;
#pragma unroll
    for (int u = 0; u < 4; u++) {
        const int i = min((int)threadIdx.x + u * NT, tot - 1);
        k.v[u] = kpart[i];
    }
    return k;
}
// kp_s: scratch of at least min(nkp * C, 4 * NT) doubles (free LDS after the
// main loop); kc_s: C doubles
__device__ __forceinline__ void kconst_finish(const KPre &k, const double *__restrict__ kpart, int nkp, int C,
                                              double *kp_s, double *kc_s)
{
    const int tot = nkp * C, held = min(tot, 4 * NT);
#pragma unroll
    for (int u = 0; u < 4; u++) {
        const int i = (int)threadIdx.x + u * NT;
        if (i < held) kp_s[i] = k.v[u];
    }
    __syncthreads();
    if (threadIdx.x < C) {
        double v = 0.0;
        for (int w = 0; w < nkp; w++) {
            const int i = (int)threadIdx.x * nkp + w;
            v += i < held ? kp_s[i] : kpart[i];
        }
        kc_s[threadIdx.x] = v;
    }
}

// r_c += X_s c - K_c for every chain: the int8 MFMA product of the
// individual-major genotype tiles with the limbs of c (exact int32 sums),
// then r_c(i) += (sum_k 256^k T_k(i, c)) 2^-e_c - K_c in FP64.  Workgroup wg
// of ng takes the 16-individual tiles [wg ntr / ng, (wg + 1) ntr / ng) (ntr
// = ceil(n / 16)); its four waves split them into runs of at most RW and walk
// every super-step of the block together (limb_mainloop), so each limb slab
// is read once per workgroup.  ng = ceil(ntr / (4 RW)), so nearly every wave
// has RW tiles (padding tiles would cost MFMA time on their SIMD); the CUs
// that hold fewer update workgroups take the U pairs that follow them in the
// launch (workgroups ng and up, uq_pair mode 1).  Each lane holds the three
// limbs of its chains in one accumulator position and updates four
// consecutive residuals per (tile, chain tile) directly from registers.
// Folds the maximum |r| of its new values into slot wg mod RSLOT of each of
// its chains (rmax[c * RSLOT + slot], atomic max; k_dquant clears the slots
// before the launch) for the next block's quantisation.
template <int NTILE, int RW>
__global__ __launch_bounds__(NT, RW * NTILE <= 12 ? 3 : 2) void k_upd(Par P, const uint4 *__restrict__ gt, int nss, int ntr,
                                            const uint4 *__restrict__ Dq, const int *__restrict__ expo,
                                            const double *__restrict__ kpart, int nkp, double *__restrict__ r,
                                            uint64_t *__restrict__ rmax, int ng, int ngrp, int npair,
                                            const uint4 *__restrict__ Ua, int nbt, int nb,
                                            const int *__restrict__ ue, double *__restrict__ upart, int ug)
{
    constexpr int W = NT / WAVE, SLAB = NTILE * 4, NTC = NTILE / LIMBS;
    static_assert(2 * SLAB * 64 >= uq_lds(NTILE), "LDS of the U pairs");
    __shared__ uint4 lds[2 * SLAB * 64];   // limb slab ring, then the K-constant parts
    __shared__ double kc_s[64];
    __shared__ unsigned long long mx_s[W][16 * NTC];
    if ((int)blockIdx.x >= ngrp * ng) {
        // the workgroups after the update's: pairs of U_s diag(s) c~ (same
        // limbs Dq and exponents), chain group by chain group, filling the
        // CUs the update leaves idle
        const int k = (int)blockIdx.x - ngrp * ng, cg = k / npair;
        uq_pair<1, NTILE>(P, Ua, nbt, nb, nss, Dq + (int64_t)cg * SLAB * 64, ngrp * NTILE, cg * 16 * NTC, nullptr,
                          nullptr, ue, expo, upart, k - cg * npair, ug, lds);
        return;
    }
    const KPre kpre = kconst_prefetch(kpart, nkp, P.nchains);
    const int wave = __builtin_amdgcn_readfirstlane((int)threadIdx.x >> 6), lane = threadIdx.x & 63;
    const int g = lane >> 4, l16 = lane & 15;
    const int cg = (int)blockIdx.x / ng, wg = (int)blockIdx.x - cg * ng;   // chain group, row workgroup
    const int cbase = cg * 16 * NTC;
    const int t0 = (int)((int64_t)wg * ntr / ng), T = (int)((int64_t)(wg + 1) * ntr / ng) - t0;
    const int w0 = t0 + wave * T / W, nt = t0 + (wave + 1) * T / W - w0;   // nt <= RW
    int64_t toff[RW];
#pragma unroll
    for (int it = 0; it < RW; it++) toff[it] = (int64_t)min(w0 + it, ntr - 1) * nss;
    v4i acc[RW][NTILE];
    limb_mainloop<NTILE, RW>(gt + lane, toff, Dq + (int64_t)cg * SLAB * 64, ngrp * SLAB, 0, nss, lds, acc);
    kconst_finish(kpre, kpart, nkp, P.nchains, reinterpret_cast<double *>(lds), kc_s);
    __syncthreads();
    const int C = P.nchains;
    double sc[NTC], kc[NTC];
    unsigned long long mx[NTC];
#pragma unroll
    for (int tc = 0; tc < NTC; tc++) {
        const int c = min(cbase + 16 * tc + l16, C - 1);
        sc[tc] = ldexp(1.0, -expo[c]);   // S * sc == ldexp(S, -e) exactly
        kc[tc] = kc_s[c];
        mx[tc] = 0;
    }
#pragma unroll
    for (int it = 0; it < RW; it++) {
        if (it >= nt) break;
        const int i0 = 16 * (w0 + it) + 4 * g;
        double2 rv[NTC][2];
#pragma unroll
        for (int tc = 0; tc < NTC; tc++) {
            const double2 *rp = reinterpret_cast<const double2 *>(
                r + (int64_t)min(cbase + 16 * tc + l16, C - 1) * P.ldr + i0);
            rv[tc][0] = rp[0];
            rv[tc][1] = rp[1];
        }
#pragma unroll
        for (int tc = 0; tc < NTC; tc++) {
            const int c = cbase + 16 * tc + l16;
            if (c >= C) continue;
            double *rp = r + (int64_t)c * P.ldr + i0;
#pragma unroll
            for (int e = 0; e < 4; e++) {
                if (i0 + e >= P.n) break;
                const long long S = (long long)acc[it][tc][e] + 256LL * acc[it][NTC + tc][e] +
                                    65536LL * acc[it][2 * NTC + tc][e];
                const double old = e == 0 ? rv[tc][0].x : e == 1 ? rv[tc][0].y : e == 2 ? rv[tc][1].x : rv[tc][1].y;
                const double nv = old + ((double)S * sc[tc] - kc[tc]);
                rp[e] = nv;
                const unsigned long long b = (unsigned long long)__double_as_longlong(fabs(nv));
                mx[tc] = b > mx[tc] ? b : mx[tc];
            }
        }
    }
    // the four lanes of a chain (g = 0..3), then the four waves
#pragma unroll
    for (int tc = 0; tc < NTC; tc++) {
#pragma unroll
        for (int off = 16; off < WAVE; off <<= 1) {
            const unsigned long long o = __shfl_xor(mx[tc], off, WAVE);
            mx[tc] = o > mx[tc] ? o : mx[tc];
        }
        if (g == 0) mx_s[wave][16 * tc + l16] = mx[tc];
    }
    __syncthreads();
    if ((int)threadIdx.x < min(16 * NTC, C - cbase)) {
        unsigned long long m = mx_s[0][threadIdx.x];
#pragma unroll
        for (int w = 1; w < W; w++) m = mx_s[w][threadIdx.x] > m ? mx_s[w][threadIdx.x] : m;
        atomicMax(reinterpret_cast<unsigned long long *>(rmax) + (int64_t)(cbase + threadIdx.x) * RSLOT + wg % RSLOT,
                  (unsigned long long)m);
    }
}

// One thread per (SNP, chain): exact limb combination of the chunk
// partials and the U_s'r_a group partials, rhs, draw (Philox, counter
// (iteration, SNP id, tag)).  The residual update applies
// c_j = delta_j / s_j + carry_j, where carry_j is what the last
// quantisation of SNP j's update left unapplied (error feedback, so the
// rounding of the fixed-point update never accumulates in r); c_j goes to
// cj_out for k_dquant, and per workgroup max |c_j| sets its exponent.
constexpr int DCH = 8;   // chunk partials unrolled in k_draw_b
constexpr int DRG = 8;   // r_a row-group partials unrolled in k_draw_b (4,096 rows)
template <int NM>
__global__ __launch_bounds__(NT) void k_draw_b(Par P, const Scal *sc, const long long *Tp,
                                               int64_t tstride, int nch, const int *expo, const long long *qsum, int nq,
                                               const double *rapart, int spant, const double *f2,
                                               const double *sdv, const int32_t *grp,
                                               const int32_t *snp_of_pos, int p0, int nb, double zz,
                                               const int *it, double *beta, int32_t *cls, double *vara_s,
                                               double *delta, uint8_t *chg, double *rhs_out, uint64_t *dmax,
                                               const double *carry, double *cj_out)
{
    __shared__ unsigned long long mx[NT / WAVE];
    __shared__ MixConst mc;
    constexpr int L = LIMBS;
    const int jk = blockIdx.x * NT + threadIdx.x;
    const int c = blockIdx.y;
    const int C = P.nchains;
    const bool mixm = P.method != SB_BAYES_A && P.method != SB_BAYES_B;
    double cj = 0.0;
    {
        // Every load of the thread is issued before the first use: the SNP
        // state, the chunk partials (up to DCH chunks unrolled) and the r_a
        // partials (up to DRG row groups).  Loads are unconditional (clamped
        // index, also past the block) and masked afterwards: a guarded load
        // compiles to a branch and a wait per load.  The per-chain mixture
        // constants are computed while they are in flight.
        const int jkc = min(jk, nb - 1);
        const int pos = p0 + jkc;
        const int64_t o = (int64_t)c * P.p;
        const bool ab = P.method == SB_BAYES_A || P.method == SB_BAYES_B;
        const double s = sdv[pos], f2p = f2[pos];
        const double gk = beta[o + pos];
        const int cold = P.method == SB_BAYES_A ? 2 : cls[o + pos];
        const double vs = ab ? vara_s[o + pos] : 0.0;
        const int gp = grp[pos], sid = snp_of_pos[pos];
        const double cy = carry[o + pos];
        const long long *tcol = Tp + (int64_t)c * P.blocksize + jkc;
        long long v[DCH];
#pragma unroll
        for (int u = 0; u < DCH; u++) v[u] = tcol[(int64_t)min(u, nch - 1) * tstride];
        // row groups entirely below this column's diagonal tile hold no partial
        const int ng = (jkc >> 4) / spant + 1;
        const double *rp = rapart + (int64_t)c * P.blocksize + jkc;
        const int64_t gstride = (int64_t)C * P.blocksize;
        double ra[DRG];
#pragma unroll
        for (int u = 0; u < DRG; u++) ra[u] = rp[(int64_t)min(u, ng - 1) * gstride];
        // sum r of the quantised residuals: the exact integer sum of the
        // quantisation workgroups' sums of q, formed by every wave (one part
        // per lane, then a butterfly; integer, so in any order)
        long long Sq = 0;
        for (int u0 = 0; u0 < nq; u0 += 4 * WAVE) {   // four loads in flight per lane
            long long v[4];
#pragma unroll
            for (int i = 0; i < 4; i++) {
                const int u = u0 + i * WAVE + (threadIdx.x & (WAVE - 1));
                v[i] = qsum[(int64_t)c * nq + min(u, nq - 1)];
                v[i] = u < nq ? v[i] : 0;
            }
#pragma unroll
            for (int i = 0; i < 4; i++) Sq += v[i];
        }
        if (mixm && !rhs_out) mix_const(P, sc + c, zz, mc);   // workgroup-uniform
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) Sq += __shfl_xor(Sq, off, WAVE);   // every lane, before any exits
        if (jk >= nb) goto reduce;
        long long S = 0;
#pragma unroll
        for (int u = 0; u < DCH; u++) S += u < nch ? v[u] : 0;
        for (int ch = DCH; ch < nch; ch++) S += tcol[(int64_t)ch * tstride];   // more chunks than unrolled
        const double sx = ldexp((double)S, -expo[c]), rt = ldexp((double)Sq, -expo[c]);
        double rhs = s > 0.0 ? (sx - f2p * rt) / s : 0.0;
#pragma unroll
        for (int u = 0; u < DRG; u++) rhs += u < ng ? ra[u] : 0.0;
        for (int g = DRG; g < ng; g++) rhs += rp[(int64_t)g * gstride];
        if (rhs_out) {  // diagnostics: the products only, no draw
            rhs_out[(int64_t)c * nb + jk] = rhs;
        } else {
            const Scal *scc = sc + c;
            PhiloxSrc src(chain_key(P, c), (uint32_t)it[0], (uint32_t)sid, TAG_SNP);
            const SnpDraw d = mixm ? draw_mix_fast<NM>(P, scc, src, mc, rhs, zz, gp, gk, cold)
                                   : draw_core<PhiloxSrc, NM>(P, scc, src, rhs, zz, zz / scc->vare, gp, gk, cold, vs);
            beta[o + pos] = d.beta;
            if (P.method != SB_BAYES_A) cls[o + pos] = d.cls;
            if (ab) vara_s[o + pos] = d.vs;
            delta[o + pos] = d.delta;
            chg[o + pos] = d.chg;
            cj = (s > 0.0 ? d.delta / s : 0.0) + cy;
            cj_out[(int64_t)c * P.blocksize + jk] = cj;
        }
    }
reduce:
    if (rhs_out) return;
    unsigned long long m = (unsigned long long)__double_as_longlong(fabs(cj));
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        unsigned long long o = __shfl_xor(m, off, WAVE);
        m = o > m ? o : m;
    }
    if ((threadIdx.x & 63) == 0) mx[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < NT / WAVE; w++) m = mx[w] > m ? mx[w] : m;
        dmax[(int64_t)c * gridDim.x + blockIdx.x] = m;
    }
}

// ----------------------------------------------------------------- host --

// SNP words of a block rounded up to whole super-steps (16 words)
int batch_words(int nb) { return ((nb + 255) / 256) * 16; }

// Row tiles per wave of the int8 products: rw_sel where the accumulators
// allow (2 or 4), else 2.
template <int NTL, class F>
void with_rw(int rw_sel, F &&f)
{
    using std::integral_constant;
    if constexpr (prod_rw(NTL) == 4) {
        if (rw_sel == 4) {
            f(integral_constant<int, 4>{});
            return;
        }
    }
    f(integral_constant<int, 2>{});
}

// Row groups (of 64 rw SNPs) of the handle's full-size block
int prod_groups(const sb_handle *h, int rw)
{
    const int nbt = (h->par.blocksize + 15) / 16;
    return std::max(1, (nbt + 4 * rw - 1) / (4 * rw));
}

// K steps (64 rows or columns each) per group of the U pairs: 16 with two
// chain groups, whose genotype workgroups keep the CUs busy while the
// longer pairs run (half the partial sums), else 8, so that no pair
// outlasts the genotype workgroups of its launch
int batch_uq_group(int nchains) { return chain_groups((nchains + 15) / 16) >= 2 ? 16 : 8; }

// X_s'Q chunks for about q / 4 workgroups per CU (chain groups x row groups x
// chunks), each chunk at least two super-steps when the individuals allow
int prod_chunks(const sb_handle *h, int rw, int q)
{
    const int nS = h->par.ldw / 16, nrg = prod_groups(h, rw) * chain_groups((h->par.nchains + 15) / 16);
    const int want = std::max(1, (q * h->ncu / 4 + nrg / 2) / nrg);
    return std::max(1, std::min(want, std::max(1, nS / 2)));
}

// Update workgroups for row tiles per wave rw: ceil(tiles / (4 rw)), so each
// takes at most 4 rw tiles of 16 individuals
int upd_groups(const sb_handle *h, int rw)
{
    const int ntr = (h->par.n + 15) / 16;
    return std::max(1, (ntr + 4 * rw - 1) / (4 * rw));
}

// X_s'Q of block b with row tiles per wave rw and nch chunks, and (with_u)
// the pairs of the block's U_s'r_a (one launch, through the handle's kernel
// clock)
int prod_launch(sb_handle *h, int b, int rw_sel, int nch, bool with_u)
{
    const Par &P = h->par;
    hipStream_t st = h->stream;
    const int p0 = h->ls[b], nb = h->ls[b + 1] - p0, nbt = (nb + 15) / 16;
    const int C = P.nchains;
    const int64_t nup = with_u ? uq_pairs(nbt, batch_words(nb) / 16, 0, h->uq_group) : 0;   // U_s'r_a pairs
    unsigned long long *amax = h->b_ramax + (int64_t)(b & 1) * C, *anext = h->b_ramax + (int64_t)((b + 1) & 1) * C;
    const uint4 *Q = reinterpret_cast<const uint4 *>(h->b_Q);
    const int64_t tstride = (int64_t)P.blocksize * P.nchains;   // int64 chunk partials per chunk
    const uint4 *gsw = reinterpret_cast<const uint4 *>(h->geno_sw) + h->gs_off[b];
    const int ngrp = chain_groups((C + 15) / 16);
    KClock &kc = h->kc;
    return with_tiles(h->b_ntile, (C + 15) / 16, [&](auto ntl, auto) {
        constexpr int NTL = decltype(ntl)::value;
        with_rw<NTL>(rw_sel, [&](auto rw) {
            constexpr int RW = decltype(rw)::value;
            const int nrg = (nbt + 4 * RW - 1) / (4 * RW);
            launch(kc, k_prod<NTL, RW>, dim3(ngrp * (nrg * nch + (unsigned)nup)), dim3(NT), 0, st, P, gsw, nbt, nb,
                   Q, nch, reinterpret_cast<long long *>(h->b_T), tstride, nrg, ngrp, (int)nup,
                   reinterpret_cast<const uint4 *>(h->Uq0) + h->uq0_off[b], h->ra, amax, anext,
                   h->ue0 + (int64_t)b * P.blocksize, h->b_rapart, h->uq_group);
        });
    });
}

// r += X_s C - K of block b with row tiles per wave rw, on the residuals r
// (the handle's, or a scratch copy while tuning); returns the number of
// max |r| partials it leaves per chain through *ng_out
int upd_launch(sb_handle *h, int b, int rw_sel, double *r, int *ng_out, bool with_u)
{
    const Par &P = h->par;
    hipStream_t st = h->stream;
    const int p0 = h->ls[b], nb = h->ls[b + 1] - p0, nbt = (nb + 15) / 16;
    const int64_t nup = with_u ? uq_pairs(nbt, batch_words(nb) / 16, 1, h->uq_group) : 0;   // U_s diag(s) c~ pairs
    const int C = P.nchains;
    const int nsw = batch_words(nb), ntr = (P.n + 15) / 16;
    const uint4 *Dq = reinterpret_cast<const uint4 *>(h->b_Dq);
    const uint4 *gt = reinterpret_cast<const uint4 *>(h->geno_t) + h->gt_off[b];
    const int ngrp = chain_groups((C + 15) / 16);
    KClock &kc = h->kc;
    return with_tiles(h->b_ntile, (C + 15) / 16, [&](auto ntl, auto) {
        constexpr int NTL = decltype(ntl)::value;
        with_rw<NTL>(rw_sel, [&](auto rw) {
            constexpr int RW = decltype(rw)::value;
            const int ng = upd_groups(h, RW);
            *ng_out = ng;
            launch(kc, k_upd<NTL, RW>, dim3(ngrp * (ng + (unsigned)nup)), dim3(NT), 0, st, P, gt, nsw / 16, ntr,
                   Dq, h->b_expo2, h->b_kpart, (nb + NT - 1) / NT, r, h->b_rmax, ng, ngrp, (int)nup,
                   reinterpret_cast<const uint4 *>(h->Uq1) + h->uq1_off[b], nbt, nb, h->ue1 + (int64_t)b * P.blocksize,
                   h->b_uapart, h->uq_group);
        });
    });
}

int batch_block(sb_handle *h, int b, double zz, int rep)
{
    (void)zz;
    (void)rep;
    const Par &P = h->par;
    hipStream_t st = h->stream;
    const int p0 = h->ls[b], nb = h->ls[b + 1] - p0, nbt = (nb + 15) / 16;
    const int C = P.nchains, L = h->limbs;
    const int nwg_r = h->nwg_r;
    uint4 *Q = reinterpret_cast<uint4 *>(h->b_Q);
    KClock &kc = h->kc;
    const double n = P.n, cl = (double)C * L, gbytes = (double)nb * ((n + 3.0) / 4.0);
    const double tbytes = 1024.0 * ULIMBS * (double)uq_off0(nbt);   // U limb tiles of the block (mode 0)
    if (h->q_block != b || h->rstat_nwg == 0) {
        // r changed outside the batched update (first block of a sweep)
        kc.begin(st);
        if (h->rstat_nwg == 0) {
            launch(kc, k_rstat, dim3(nwg_r, 1, C), dim3(NT), 0, st, P, h->r, h->b_rmax);
            h->rstat_nwg = nwg_r;
        }
        kc.last(KS_RQUANT, 16.0 * C * n + cl * n);
        launch(kc, k_quant, dim3(h->b_nq, 1, h->b_ntile / LIMBS), dim3(NT), 0, st, P, h->b_ntile, h->r, h->b_rmax,
               h->rstat_nwg, Q, h->b_expo, h->b_qsum);
    }
    h->q_block = -1;
    if (h->ra_block != b) {
        // r_a set outside the batched update (drawn for the sweep): the
        // chain maxima of its limbs (once per sweep, not timed)
        launch(kc, k_ramax, dim3(C), dim3(NT), 0, st, P, h->ra, nb, h->b_ramax + (int64_t)(b & 1) * C);
    }
    h->ra_block = -1;
    kc.begin(st);
    kc.last(KS_XQ, gbytes + cl * n + 8.0 * C * nb * h->b_nch + tbytes + 16.0 * C * nb);
    const int rc = prod_launch(h, b, h->b_rw, h->b_nch, true);   // with U_s'r_a
    if (rc) return fail(rc, "unsupported chain count");
    return hip_check(hipGetLastError(), "batch products");
}

int batch_update(sb_handle *h, int b)
{
    const Par &P = h->par;
    hipStream_t st = h->stream;
    const int p0 = h->ls[b], nb = h->ls[b + 1] - p0, nbt = (nb + 15) / 16;
    const int C = P.nchains, L = h->limbs;
    const int nsw = batch_words(nb), nwg_d = (nb + NT - 1) / NT;
    uint4 *Dq = reinterpret_cast<uint4 *>(h->b_Dq);
    KClock &kc = h->kc;
    const double n = P.n, cl = (double)C * L, gbytes = (double)nb * ((n + 3.0) / 4.0);
    const double tbytes = 4096.0 * ULIMBS * (double)uq_off1(nbt, nsw / 16);   // U limb tiles (mode 1)
    kc.begin(st);
    kc.last(KS_DQUANT, 16.0 * C * nb + cl * nb);
    launch(kc, k_dquant, dim3(nsw * 16 / NT, C), dim3(NT), 0, st, P, h->b_ntile, h->b_cj, h->f2, p0, nb,
           h->b_dmax, nwg_d, Dq, h->b_expo2, h->b_carry, h->b_kpart, h->b_rmax);
    const int spant = h->rows_per_group / 16, ncg = (nbt + spant - 1) / spant;
    kc.begin(st);
    kc.last(KS_XUPD, gbytes + cl * nb + 16.0 * C * n + tbytes + 12.0 * C * nb);
    int ng = 0;
    const int rc = upd_launch(h, b, h->b_rw, h->r, &ng, true);   // with U_s diag(s) c~
    if (rc) return fail(rc, "unsupported chain count");
    (void)ng;
    h->rstat_nwg = RSLOT;   // max |r| of the new r, in the update's slots
    kc.begin(st);
    if (b + 1 < P.nblocks) {
        // the next block's r limbs and r_a maxima in the same launch
        const int nq = h->b_nq, nbn = h->ls[b + 2] - h->ls[b + 1];
        const int nra = (std::max(nb, nbn) + NT - 1) / NT;
        uint4 *Q = reinterpret_cast<uint4 *>(h->b_Q);
        kc.last(KS_RARED, 8.0 * C * nb * (ncg + 3.0) + 8.0 * C * n + cl * n);
        launch(kc, k_rared_quant, dim3(nq * (h->b_ntile / LIMBS) + nra * C), dim3(NT), 0, st, P, h->b_uapart, spant,
               ncg, nb, h->ra,
               nra, h->b_ntile, h->r, h->b_rmax, h->rstat_nwg, Q, h->b_expo, h->b_qsum, nq, nbn,
               h->b_ramax + (int64_t)((b + 1) & 1) * C);
        h->q_block = b + 1;
        h->ra_block = b + 1;
    } else {
        const int nra = (nb + NT - 1) / NT;
        kc.last(KS_RARED, 8.0 * C * nb * (ncg + 3.0));
        launch(kc, k_ra_reduce, dim3(nra, C), dim3(NT), 0, st, P, h->b_uapart, spant, ncg, nb, h->ra);
    }
    return hip_check(hipGetLastError(), "batch update");
}

// Tile shape of the batched products: the X_s'Q chunk count (0.75 to 2
// workgroups per CU) and the row tiles per wave of the int8 products (2 or
// 4).  Every shape gives bit-identical results (the limb products are exact
// int32 sums and the chunk partials are added as 64-bit integers), so the
// choice affects speed only.  Picked once per handle, before its first
// sweep, by timing the products (k_prod) and update (k_upd, on a scratch
// copy of the residuals) of the first six blocks for every candidate, in
// three interleaved rounds; SODABAYES_KSS (super-steps per chunk) /
// SODABAYES_ROWS_PER_WAVE fix the shape instead.
void tune_rws(const sb_handle *h, std::vector<int> &rws)
{
    rws.assign({2});
    if (prod_rw(h->b_ntile) == 4) rws.push_back(4);
}
// chunk candidates: 0.75 to 3 workgroups per CU (in quarters), and the
// counts whose X_s'Q workgroups and U'r_a pairs together come closest to one,
// two or three workgroups per CU without passing it.  Those often win: at
// n = 50,000, 7 chunks both at 64 chains (448 + 320 workgroups, three per
// CU) and at 32 chains (224 + 288, two per CU), which the quarter steps miss
// at 32 chains.
constexpr int TUNE_Q[] = {3, 4, 5, 6, 7, 8, 10, 12};
void tune_chunks(const sb_handle *h, int rw, std::vector<int> &nchs)
{
    nchs.clear();
    auto add = [&](int c) {
        c = std::max(1, std::min(c, std::max(1, h->par.ldw / 32)));
        if (std::find(nchs.begin(), nchs.end(), c) == nchs.end()) nchs.push_back(c);
    };
    for (int q : TUNE_Q) add(prod_chunks(h, rw, q));
    const int ngrp = chain_groups((h->par.nchains + 15) / 16), nrgt = prod_groups(h, rw) * ngrp;
    const int nbt = (h->par.blocksize + 15) / 16;
    const int64_t npair = uq_pairs(nbt, batch_words(h->par.blocksize) / 16, 0, h->uq_group) * ngrp;
    for (int k = 1; k <= 3; k++) {
        const int64_t c = ((int64_t)k * h->ncu - npair) / nrgt;
        if (c >= 1) add((int)c);
    }
}

// chunks the partial buffer must hold: the most any candidate uses
int batch_max_chunks(const sb_handle *h)
{
    int mx = 1;
    std::vector<int> nchs;
    for (int rw : {2, 4}) {
        tune_chunks(h, rw, nchs);
        for (int c : nchs) mx = std::max(mx, c);
    }
    return std::max(mx, h->b_nch);
}

int batch_autotune(sb_handle *h)
{
    h->tuned = true;
    h->graph_ok = false;   // the tile shape changes the captured launches
    const Par &P = h->par;
    hipStream_t st = h->stream;
    if (!h->tune_r && dev_alloc(h, (void **)&h->tune_r, sizeof(double) * (size_t)P.ldr * P.nchains))
        return SB_ERR_NOMEM;
    hipEvent_t e0, e1;
    hipEventCreate(&e0);
    hipEventCreate(&e1);
    const bool timing = h->kc.on;
    h->kc.on = false;   // the trial launches are not kernel-clock spans
    // each candidate over several different blocks, so that its operands come
    // from HBM as in a sweep (repeating one block would time it from the
    // caches)
    const int nbt = std::min(6, P.nblocks);
    auto time_it = [&](auto &&fn) {
        fn(0);   // warm-up
        hipEventRecord(e0, st);
        for (int b = 0; b < nbt; b++) fn(b);
        hipEventRecord(e1, st);
        hipEventSynchronize(e1);
        float ms = 0.f;
        hipEventElapsedTime(&ms, e0, e1);
        return ms / nbt;
    };
    int brw = h->b_rw, bnch = h->b_nch, rc = 0, ng = 0;
    std::vector<int> rws, nchs;
    std::vector<std::pair<int, int>> cand;   // (rows per wave, chunks)
    tune_rws(h, rws);
    for (int rw : rws) {
        tune_chunks(h, rw, nchs);
        for (int nch : nchs) cand.emplace_back(rw, nch);
    }
    // products and update of each block in turn, as a sweep runs them; the
    // candidates in rounds, each keeping its fastest round, so that a slow
    // moment of the GPU does not decide
    constexpr int ROUNDS = 3;
    std::vector<double> tbest(cand.size(), 1e30);
    for (int rd = 0; rd < ROUNDS; rd++)
        for (size_t i = 0; i < cand.size(); i++) {
            const int rw = cand[i].first, nch = cand[i].second;
            const float t = time_it([&](int b) {
                rc |= prod_launch(h, b, rw, nch, true);
                rc |= upd_launch(h, b, rw, h->tune_r, &ng, true);
            });
            tbest[i] = std::min(tbest[i], (double)t);
        }
    // each chunk adds nchains x blocksize int64 partials that the draw reads
    // back (not in the trial launches): charged at 2.5 TB/s
    double best = 1e30;
    for (size_t i = 0; i < cand.size(); i++) {
        const double t = tbest[i] + cand[i].second * 8.0 * P.nchains * P.blocksize / 2.5e12 * 1e3;
        if (t < best) {
            best = t;
            brw = cand[i].first;
            bnch = cand[i].second;
        }
    }
    h->kc.on = timing;
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    if (rc) return fail(rc, "unsupported chain count");
    h->b_rw = brw;
    h->b_nch = bnch;
    h->kss = (P.ldw / 16 + bnch - 1) / bnch;
    h->tune_us = best * 1e3;
    h->rstat_nwg = 0;   // the trial updates overwrote the residual statistics
    h->q_block = -1;
    h->ra_block = -1;
    return hip_check(hipGetLastError(), "tile-shape tuning");
}

// Read-only pass over a layout: every wave reads contiguous 1 KB (one uint4
// per lane) per load, four loads in flight per lane, grid-stride, as the
// product kernels read their operand tiles.  The XOR of everything read is
// stored only if it equals a value it cannot take when `never` is 0, which
// keeps the loads alive.
__global__ __launch_bounds__(NT) void k_stream(const uint4 *__restrict__ p, int64_t n, uint4 *sink, int never)
{
    uint4 acc = make_uint4(0, 0, 0, 0);
    const int64_t stride = (int64_t)gridDim.x * NT * 4;
    for (int64_t base = (int64_t)blockIdx.x * NT * 4 + threadIdx.x; base < n; base += stride) {
        uint4 v[4];
#pragma unroll
        for (int u = 0; u < 4; u++) {
            const int64_t i = min(base + (int64_t)u * NT, n - 1);
            v[u] = p[i];
        }
#pragma unroll
        for (int u = 0; u < 4; u++) {
            acc.x ^= v[u].x;
            acc.y ^= v[u].y;
            acc.z ^= v[u].z;
            acc.w ^= v[u].w;
        }
    }
    if (never && acc.x == 0x9E3779B9u && acc.y == 0x7F4A7C15u) sink[threadIdx.x] = acc;
}

// For each operand layout of the handle's path, the time of one read-only
// pass per block, one launch per block as the sweep launches its kernels,
// over every block (a working set far beyond the caches): out[3 k ..] =
// (bytes per block, microseconds per block, GB/s) for layout k: the batched
// path's geno_sw, geno_t, Uq0, Uq1; the FP64 path's genotype words and
// packed U (then slots 2, 3 are zero).
int layout_stream(sb_handle *h, double *out)
{
    const Par &P = h->par;
    hipStream_t st = h->stream;
    for (int k = 0; k < 12; k++) out[k] = 0.0;
    if (P.nblocks <= 0) return fail(SB_ERR_STATE, "sb_layout_stream: no SODA blocks");
    uint4 *sink = nullptr;
    TmpDev tmp;
    int rc = tmp.alloc(&sink, sizeof(uint4) * NT, "stream sink");
    if (rc) return rc;
    struct Lay {
        const uint4 *base;
        std::vector<int64_t> off;   // uint4 offset of each block, nblocks + 1
    };
    std::vector<Lay> lays;
    const int nb = P.nblocks;
    auto mk = [&](const void *b, auto &&f) {
        Lay l{static_cast<const uint4 *>(b), std::vector<int64_t>(nb + 1)};
        for (int i = 0; i <= nb; i++) l.off[i] = f(i);
        lays.push_back(l);
    };
    if (P.precision == SB_PREC_MIXED) {
        mk(h->geno_sw, [&](int i) { return h->gs_off[i]; });
        mk(h->geno_t, [&](int i) { return h->gt_off[i]; });
        mk(h->Uq0, [&](int i) { return h->uq0_off[i]; });
        mk(h->Uq1, [&](int i) { return h->uq1_off[i]; });
    } else {
        mk(h->geno, [&](int i) { return (int64_t)h->ls[i] * P.ldw / 4; });
        // packed triangles of doubles: block starts rounded down to whole
        // 16-byte elements (a block's last double may be left out)
        mk(h->U, [&](int i) { return h->tri_off[i] / 2; });
    }
    hipEvent_t e0, e1;
    hipEventCreate(&e0);
    hipEventCreate(&e1);
    for (size_t k = 0; k < lays.size() && !rc; k++) {
        const Lay &l = lays[k];
        auto pass = [&]() {
            for (int b = 0; b < nb; b++) {
                const int64_t n = l.off[b + 1] - l.off[b];
                if (n <= 0) continue;
                const int grid = (int)std::min<int64_t>((n + 4 * NT - 1) / (4 * NT), 8192);
                k_stream<<<grid, NT, 0, st>>>(l.base + l.off[b], n, sink, 0);
            }
        };
        pass();   // warm-up (page tables, code)
        hipEventRecord(e0, st);
        pass();
        hipEventRecord(e1, st);
        rc = hip_check(hipEventSynchronize(e1), "layout stream");
        float ms = 0.f;
        hipEventElapsedTime(&ms, e0, e1);
        const double bytes = 16.0 * (double)(l.off[nb] - l.off[0]) / nb, us = 1e3 * ms / nb;
        out[3 * k] = bytes;
        out[3 * k + 1] = us;
        out[3 * k + 2] = us > 0 ? bytes / (us * 1e-6) / 1e9 : 0.0;
    }
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    if (!rc) rc = hip_check(hipGetLastError(), "layout stream");
    return rc;
}

// Every chain's carried quantisation remainders to zero: when r is exact
// (the initial state, a resync r = y - X beta - mu) nothing is left unapplied.
// A fill kernel rather than a memset, so that the call can sit in a captured
// graph.
__global__ void k_zero(double *v, int64_t n)
{
    for (int64_t k = (int64_t)blockIdx.x * NT + threadIdx.x; k < n; k += (int64_t)gridDim.x * NT) v[k] = 0.0;
}

int batch_clear_carry(sb_handle *h)
{
    if (!h->b_carry) return SB_OK;
    const int64_t n = (int64_t)h->par.nchains * h->par.p;
    k_zero<<<(int)std::min<int64_t>((n + NT - 1) / NT, 4096), NT, 0, h->stream>>>(h->b_carry, n);
    return hip_check(hipGetLastError(), "clear carry");
}

// Genotype operand layouts of the batched path, built once at sb_create.
int batch_transpose(sb_handle *h)
{
    const Par &P = h->par;
    const int nS = P.ldw / 16;
    for (int b = 0; b < P.nblocks; b++) {
        const int p0 = h->ls[b], nb = h->ls[b + 1] - p0, nbt = (nb + 15) / 16;
        const int nsw = batch_words(nb);
        k_transpose<<<dim3((P.ldw + NT - 1) / NT, nsw), NT, 0, h->stream>>>(
            h->geno, P.ldw, p0, nb, nsw / 16, h->geno_t + 4 * h->gt_off[b]);
        k_swz_geno<<<dim3((nS * 64 + NT - 1) / NT, nbt), NT, 0, h->stream>>>(
            h->geno, P.ldw, p0, nb, reinterpret_cast<uint4 *>(h->geno_sw) + h->gs_off[b]);
    }
    int rc = hip_check(hipGetLastError(), "genotype layouts");
    if (!rc) rc = hip_check(hipStreamSynchronize(h->stream), "genotype layouts");
    return rc;
}

// U limb tiles of every block (after the factorisation has filled h->U):
// the exponents of every column and row, then the two tile sets.
int batch_tiles(sb_handle *h)
{
    const Par &P = h->par;
    hipStream_t st = h->stream;
    k_uexp<<<dim3((P.blocksize + NT - 1) / NT, P.nblocks), NT, 0, st>>>(h->U, h->d_tri_off, h->d_ls, h->sdv,
                                                                        P.blocksize, h->ue0, h->ue1);
    for (int b = 0; b < P.nblocks; b++) {
        const int p0 = h->ls[b], nb = h->ls[b + 1] - p0, nbt = (nb + 15) / 16, nss = batch_words(nb) / 16;
        const int *ue0 = h->ue0 + (int64_t)b * P.blocksize, *ue1 = h->ue1 + (int64_t)b * P.blocksize;
        for (int mode = 0; mode < 2; mode++) {
            const int64_t nent = mode == 0 ? uq_off0(nbt) : uq_off1(nbt, nss) * 4;
            uint4 *dst = reinterpret_cast<uint4 *>(mode == 0 ? h->Uq0 : h->Uq1) +
                         (mode == 0 ? h->uq0_off[b] : h->uq1_off[b]);
            k_pack_uq<<<(unsigned)((nent * 64 + NT - 1) / NT), NT, 0, st>>>(h->U, h->tri_off[b], nb, nbt, nss, p0,
                                                                            h->sdv, ue0, ue1, mode, dst);
        }
    }
    int rc = hip_check(hipGetLastError(), "U limb tiles");
    if (!rc) rc = hip_check(hipStreamSynchronize(st), "U limb tiles");
    return rc;
}

// uint4 elements of one block's U limb tiles (mode 0 or 1)
int64_t batch_uq_elems(int nb, int mode)
{
    const int nbt = (nb + 15) / 16, nss = batch_words(nb) / 16;
    return (mode == 0 ? uq_off0(nbt) : uq_off1(nbt, nss) * 4) * ULIMBS * 64;
}

int batch_draw(sb_handle *h, int b, double zz, double *delta, uint8_t *chg, double *rhs_out)
{
    const Par &P = h->par;
    const int p0 = h->ls[b], nb = h->ls[b + 1] - p0;
    const int64_t tstride = (int64_t)P.blocksize * P.nchains;   // int64 chunk partials per chunk
    const int nch = h->b_nch;
    const double C = P.nchains;
    h->kc.begin(h->stream);
    // chunk partials (one int64 each), r_a partials, effect state read and written
    h->kc.last(KS_DRAWB, 8.0 * nch * C * nb + 8.0 * C * nb * ((nb + 255) / 256 / 2.0 + 1.0) + 60.0 * C * nb);
    with_nmix(P.nmix, [&](auto nm) {
        launch(h->kc, k_draw_b<decltype(nm)::value>, dim3((nb + NT - 1) / NT, P.nchains), dim3(NT), 0, h->stream,
               P, h->scal, reinterpret_cast<const long long *>(h->b_T), tstride, nch, h->b_expo, h->b_qsum, h->b_nq,
               h->b_rapart, h->rows_per_group / 16, h->f2, h->sdv, h->grp, h->d_pos_by_id + P.p, p0, nb, zz, h->d_it, h->beta, h->cls, h->vara_s,
               delta, chg, rhs_out, h->b_dmax, h->b_carry, h->b_cj);
    });
    return hip_check(hipGetLastError(), "batch draw");
}

}  // namespace sb

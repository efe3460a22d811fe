// Device helpers of the SODA sampler: the reference's xoroshiro128+ stream
// and its distributions (parallelrng.f90), a counter-based Philox4x32-10
// generator for the parallel mode, and wave64 reductions.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace sb {

constexpr int WAVE = 64;

// parallelrng.f90:20 -- PI is a float32 literal widened to float64
__device__ constexpr double PI_F32 = 3.1415927410125732421875;

// ---------------------------------------------------------------- replay --

struct Xoro {
    uint64_t s1, s2;
};

__device__ __forceinline__ uint64_t rotl64(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }

// rng_xoroshiro128plus, parallelrng.f90:100-123 (rotations 55/14/36, s1+s2)
__device__ __forceinline__ uint64_t xoro_next(Xoro &s)
{
    uint64_t a = s.s1, b = s.s2;
    uint64_t r = a + b;
    b ^= a;
    s.s1 = rotl64(a, 55) ^ b ^ (b << 14);
    s.s2 = rotl64(b, 36);
    return r;
}

// rng_int64_to_real64, parallelrng.f90:242-249: 52 high bits as [1,2) - 1
__device__ __forceinline__ double bits_to_unit(uint64_t x)
{
    return __longlong_as_double((long long)(0x3FF0000000000000ULL | (x >> 12))) - 1.0;
}

__device__ __forceinline__ double xoro_uniform(Xoro &s) { return bits_to_unit(xoro_next(s)); }

// rand_normal, parallelrng.f90:264-277: one normal from two uniforms
// (sin branch), sd <= 0 returns the mean without consuming anything.
__device__ __forceinline__ double box_muller(double mean, double sd, double u1, double u2)
{
    double r = sqrt(-2.0 * log(u1));
    double theta = __dmul_rn(2.0 * PI_F32, u2);
    return __dadd_rn(mean, __dmul_rn(__dmul_rn(sd, r), sin(theta)));
}

// Box-Muller with the exact pi (sinpi): the counter-based sampler's normal
// variates (REPLAY keeps box_muller and the reference's f32 pi).
__device__ __forceinline__ double box_muller_pi(double mean, double sd, double u1, double u2)
{
    return mean + sd * (sqrt(-2.0 * log(u1)) * sinpi(2.0 * u2));
}

// Uniform source concept: struct with `double next()`.
struct XoroSrc {
    Xoro st;
    __device__ double next() { return xoro_uniform(st); }
};

// Philox4x32-10 (Salmon et al., SC'11) -- counter (c0..c3), key (k0, k1).
__device__ __forceinline__ void philox4x32(uint32_t c[4], uint32_t k0, uint32_t k1)
{
    constexpr uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u;
    constexpr uint32_t W0 = 0x9E3779B9u, W1 = 0xBB67AE85u;
#pragma unroll
    for (int r = 0; r < 10; r++) {
        uint32_t hi0 = __umulhi(M0, c[0]), lo0 = M0 * c[0];
        uint32_t hi1 = __umulhi(M1, c[2]), lo1 = M1 * c[2];
        uint32_t n0 = hi1 ^ c[1] ^ k0, n1 = lo1, n2 = hi0 ^ c[3] ^ k1, n3 = lo0;
        c[0] = n0; c[1] = n1; c[2] = n2; c[3] = n3;
        k0 += W0;
        k1 += W1;
    }
}

// Counter-based uniform source: one SNP's private stream
// (key = seed ^ chain, counter = (iteration, SNP, tag, sequence)).
// Each Philox block gives two uniforms, words (0, 1) then (2, 3); they are
// kept as two 64-bit values (no indexed array, which would leave registers).
struct PhiloxSrc {
    uint32_t k0, k1, c0, c1, c2, seq;
    uint64_t x0, x1;
    int avail;
    __device__ PhiloxSrc(uint64_t key, uint32_t iter, uint32_t snp, uint32_t tag)
        : k0((uint32_t)key), k1((uint32_t)(key >> 32)), c0(iter), c1(snp), c2(tag), seq(0), x0(0), x1(0),
          avail(0) {}
    __device__ double next()
    {
        if (avail == 0) {
            uint32_t b[4] = {c0, c1, c2, seq++};
            philox4x32(b, k0, k1);
            x0 = ((uint64_t)b[0] << 32) | b[1];
            x1 = ((uint64_t)b[2] << 32) | b[3];
            avail = 2;
        }
        const uint64_t x = avail == 2 ? x0 : x1;
        avail--;
        return bits_to_unit(x);
    }
};

template <class Src>
__device__ double draw_normal(Src &s, double mean, double sd)
{
    if (sd <= 0.0) return mean;
    double u1 = s.next();
    double u2 = s.next();
    return box_muller(mean, sd, u1, u2);
}

// rand_gamma, parallelrng.f90:294-333 (Marsaglia & Tsang 2000, with the
// shape < 1 boost g(shape+1) * u^(1/shape))
template <class Src>
__device__ double draw_gamma(Src &s, double shape, double scale)
{
    bool small = shape < 1.0;
    double sh = small ? shape + 1.0 : shape;
    double d = sh - 1.0 / 3.0;
    double c = 1.0 / sqrt(9.0 * d);
    double v;
    for (;;) {
        double x = draw_normal(s, 0.0, 1.0);
        v = __dadd_rn(1.0, __dmul_rn(c, x));
        while (v <= 0.0) {
            x = draw_normal(s, 0.0, 1.0);
            v = __dadd_rn(1.0, __dmul_rn(c, x));
        }
        v = __dmul_rn(__dmul_rn(v, v), v);
        double u = s.next();
        double xsq = __dmul_rn(x, x);
        if ((u < __dsub_rn(1.0, __dmul_rn(__dmul_rn(0.0331, xsq), xsq))) ||
            (log(u) < __dadd_rn(__dmul_rn(0.5, xsq),
                                __dmul_rn(d, __dadd_rn(__dsub_rn(1.0, v), log(v))))))
            break;
    }
    if (small) {
        // g = rand_gamma(shape+1, 1) = (1*d)*v, then scale*g*w**(1/shape)
        double g = __dmul_rn(d, v);
        double w = s.next();
        return __dmul_rn(__dmul_rn(scale, g), pow(w, 1.0 / shape));
    }
    return __dmul_rn(__dmul_rn(scale, d), v);
}

// rand_scaled_inverse_chi_square via rand_inverse_gamma, parallelrng.f90:348-362
template <class Src>
__device__ double draw_sinvchi2(Src &s, double dof, double scale)
{
    double shape = __dmul_rn(0.5, dof);
    double sc = __dmul_rn(__dmul_rn(0.5, dof), scale);
    return 1.0 / draw_gamma(s, shape, 1.0 / sc);
}

// ------------------------------------------------------------- reduction --

__device__ __forceinline__ double wave_sum(double v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, WAVE);
    return v;
}

__device__ __forceinline__ int wave_sum_i(int v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, WAVE);
    return v;
}

// Workgroups are handed to the 8 XCDs round-robin (workgroup L on XCD
// L % 8), and each XCD has its own L2.  Renumbers the first 8 * (n / 8) of
// n workgroups so that every XCD takes one contiguous range of logical ids:
// neighbouring tiles that share cache lines or operand slices then meet in
// one L2 instead of being fetched by every XCD.  A bijection on [0, n).
__device__ __forceinline__ int xcd_contiguous(int L, int n)
{
    const int m = (n / 8) * 8;
    return L < m ? (L % 8) * (m / 8) + L / 8 : L;
}

// Workgroup id -> (kind, local id) for two horizontally fused grids of n0
// and n1 workgroups: interleaved in runs of 8 (one run per XCD, which take
// workgroups round-robin), then the rest of the larger grid.
__device__ __forceinline__ void split_ids(int id, int n0, int n1, int &kind, int &local)
{
    const int m = (min(n0, n1) / 8) * 8;
    if (id < 2 * m) {
        kind = (id >> 3) & 1;
        local = ((id >> 4) << 3) + (id & 7);
        return;
    }
    id -= 2 * m;
    if (id < n0 - m) {
        kind = 0;
        local = m + id;
    } else {
        kind = 1;
        local = m + id - (n0 - m);
    }
}

// Fixed-order block reduction (deterministic for a fixed blockDim).
template <int NT>
__device__ double block_sum(double v, double *lds)
{
    v = wave_sum(v);
    int lane = threadIdx.x & (WAVE - 1), w = threadIdx.x / WAVE;
    __syncthreads();
    if (lane == 0) lds[w] = v;
    __syncthreads();
    double t = 0.0;
    if (threadIdx.x == 0)
        for (int i = 0; i < NT / WAVE; i++) t += lds[i];
    __syncthreads();
    return t;  // valid in thread 0
}

}  // namespace sb

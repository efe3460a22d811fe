// The single-site sampler (-SODAOff, regular_bayes, models.f90:607-962) on
// the device: every SNP in turn, its conditional from rhs = z'r, then
// r += z (g_old - g_new) at once, so that each SNP sees the previous one's
// update.
//
// The chain is serial in SNPs; what the device can do is keep the serial
// part short.  One workgroup per chain walks the SNPs in windows of SW = 16:
//   * one pass over the individuals applies the previous window's updates
//     to r (SNP by SNP, in the reference's order and rounding) and forms
//     a_j = z_j'r for the 16 SNPs of the window, r as it stood at the
//     window's start;
//   * the draws then run one after another on wave 0 (a mixture's class
//     terms one per lane, its rhs-free terms formed once per window), SNP j
//     using
//     rhs_j = a_j + sum_{l<j} (z_j'z_l) d_l, which is z_j'r after the
//     window's earlier draws (d_l = g_old - g_new), with z_j'z_l from the
//     window's Gram, computed once at setup;
//   * so a window costs one pass over r and two barriers instead of 16 of
//     each.
// rhs is z_j'r up to the order of the sums; the draws and the residual
// updates are the reference's.
//
// Setup builds, per window, the 16 SNPs' two-bit codes interleaved per
// individual (one 32-bit word per individual, SNP j in bits 2j..2j+1), their
// z tables and the packed lower triangle of their Gram.

#include <hip/hip_runtime.h>

#include <algorithm>

#include "sb_device.h"
#include "sb_draw.h"
#include "sb_internal.h"

namespace sb {
namespace {

constexpr int SST = 512;                   // threads per chain (256 VGPRs: the draw does not spill)
constexpr int SW = 16;                     // SNPs per window
constexpr int SWT = SW * (SW + 1) / 2;     // packed Gram entries per window
constexpr int GT = 256;                    // threads of the setup kernels

// The chain scalars a draw reads (draw_core), staged in LDS for the sweep.
struct ScalLite {
    double vare, scale_va;
    double gp[MAXMIX], vare_gp[MAXMIX];
    double log_p[MAXBIN];
};

__device__ __forceinline__ int tri(int j, int l) { return j * (j + 1) / 2 + l; }   // l <= j

// z tables of each window: wz[m*64 + 4j + x] = z of code x of the window's
// SNP j (0 beyond the last SNP).
__global__ void k_win_ztab(Par P, const double *__restrict__ zlut, const int32_t *list, int nsk, double *wz)
{
    const int m = blockIdx.x, j = threadIdx.x >> 2, x = threadIdx.x & 3;
    const int k = m * SW + j;
    double z = 0.0;
    if (k < nsk && x < 3) z = zlut[4 * (int64_t)list[k] + x];
    wz[(int64_t)m * 64 + threadIdx.x] = z;
}

// Interleaved codes: win[m*ldn + i] = sum_j code(SNP j of window m, i) << 2j.
__global__ __launch_bounds__(GT) void k_win_pack(Par P, const uint32_t *__restrict__ geno, const int32_t *list,
                                                 int nsk, int64_t ldn, uint32_t *win)
{
    const int m = blockIdx.x;
    const int64_t i = (int64_t)blockIdx.y * GT + threadIdx.x;
    if (i >= ldn) return;
    uint32_t w = 0;
    if (i < P.n) {
        const int sh = 2 * (int)(i & 15);
        for (int j = 0; j < SW; j++) {
            const int k = m * SW + j;
            if (k >= nsk) break;
            const uint32_t g = geno[(int64_t)list[k] * P.ldw + (i >> 4)];
            w |= ((g >> sh) & 3u) << (2 * j);
        }
    }
    win[(int64_t)m * ldn + i] = w;
}

// Packed lower triangle of each window's Gram z_j'z_l (l <= j); the diagonal
// is xpx (models.f90:676).  Chunks of GT individuals are decoded into LDS,
// then thread q < SWT sums its pair over the chunk.
__global__ __launch_bounds__(GT) void k_win_gram(Par P, const uint32_t *__restrict__ win, int64_t ldn,
                                                 const double *__restrict__ wz, double *wgram)
{
    __shared__ double zs[SW][GT + 1];
    __shared__ double zt[64];
    const int m = blockIdx.x, t = threadIdx.x;
    if (t < 64) zt[t] = wz[(int64_t)m * 64 + t];
    int jq = 0, lq = 0;
    if (t < SWT) {
        while (tri(jq + 1, 0) <= t) jq++;
        lq = t - tri(jq, 0);
    }
    double acc = 0.0;
    __syncthreads();
    for (int64_t i0 = 0; i0 < P.n; i0 += GT) {
        const int64_t i = i0 + t;
        const uint32_t w = i < P.n ? win[(int64_t)m * ldn + i] : 0u;
#pragma unroll
        for (int j = 0; j < SW; j++) {
            const uint32_t x = (w >> (2 * j)) & 3u;
            const double z = x == 0 ? zt[4 * j] : (x == 1 ? zt[4 * j + 1] : zt[4 * j + 2]);
            zs[j][t] = i < P.n ? z : 0.0;
        }
        __syncthreads();
        if (t < SWT)
            for (int u = 0; u < GT; u++) acc = fma(zs[jq][u], zs[lq][u], acc);
        __syncthreads();
    }
    if (t < SWT) wgram[(int64_t)m * SWT + t] = acc;
}

// Wave sum of v[0..SW-1], transposed: each exchange halves the values a lane
// keeps (8 + 4 + 2 + 1 shuffles), then two more within groups of four lanes.
// Lane L ends with the sum of value idx = 8 b5 + 4 b4 + 2 b3 + b2 (bits of
// L); the lanes with L % 4 == 0 hold the SW sums.
__device__ __forceinline__ double wave_sum16(const double (&v)[SW], int lane, int &idx)
{
    const bool h5 = lane & 32, h4 = lane & 16, h3 = lane & 8, h2 = lane & 4;
    double a[8], b[4], c[2];
#pragma unroll
    for (int i = 0; i < 8; i++) a[i] = (h5 ? v[8 + i] : v[i]) + __shfl_xor(h5 ? v[i] : v[8 + i], 32, WAVE);
#pragma unroll
    for (int i = 0; i < 4; i++) b[i] = (h4 ? a[4 + i] : a[i]) + __shfl_xor(h4 ? a[i] : a[4 + i], 16, WAVE);
#pragma unroll
    for (int i = 0; i < 2; i++) c[i] = (h3 ? b[2 + i] : b[i]) + __shfl_xor(h3 ? b[i] : b[2 + i], 8, WAVE);
    double x = (h2 ? c[1] : c[0]) + __shfl_xor(h2 ? c[0] : c[1], 4, WAVE);
    x += __shfl_xor(x, 2, WAVE);
    x += __shfl_xor(x, 1, WAVE);
    idx = (h5 ? 8 : 0) + (h4 ? 4 : 0) + (h3 ? 2 : 0) + (h2 ? 1 : 0);
    return x;
}

// The terms of a mixture draw (models.f90:810-811, 852-853) that do not
// depend on rhs, for SNP j of a window and class k >= 1, held by lane
// 4j + k of wave 0: logdetV = log(gp_k zz/vare + 1), the denominator
// zz + vare/gp_k of uhat, v1 = zz + vare/gp_k and sd = sqrt(vare/v1), each
// formed as the reference forms it.
struct MixWin {
    double lg, dn, v1, sd;
};

__device__ __forceinline__ MixWin mix_window(const Par &P, const ScalLite &ss, const double *gmw, int ccur, int lane)
{
    MixWin w{0.0, 1.0, 1.0, 0.0};
    const int j = lane >> 2, k = lane & 3;
    if (j < ccur && k >= 1 && k < P.nmix) {
        const double zz = gmw[tri(j, j)];
        w.lg = log(ss.gp[k] * (zz / ss.vare) + 1.0);
        w.dn = zz + ss.vare_gp[k];
        w.v1 = zz + ss.vare / ss.gp[k];
        w.sd = sqrt(ss.vare / w.v1);
    }
    return w;
}

// draw_core's mixture branch for at most four classes, on every lane of wave
// 0 with identical results (each lane keeps the stream in step): lane k
// forms class k's term and probability (mix_probs' operations, branch-free
// so that the exponentials and the speculative normal below are independent
// instruction chains), and the normal variate's transcendental part is
// formed from the stream's 2nd and 3rd uniforms while the class is drawn;
// it is used only when the class is not 1 (then those are the uniforms
// draw_normal takes).  Returns the new effect; `flag` is the new class.
template <class Src>
__device__ __forceinline__ double draw_mix4(const Par &P, const ScalLite &ss, Src &src, const MixWin &mw, int j,
                                            double rhs0, double zz, int g, double gk, int cold, int lane, int &flag)
{
    constexpr bool EXACT = std::is_same<Src, XoroSrc>::value;
    const int nmix = P.nmix;
    const double vare = ss.vare;
    double rhs = rhs0;
    if (cold > 1) rhs = rhs + zz * gk;
    const int kl = lane < nmix ? lane : 0;
    const double lg = __shfl(mw.lg, 4 * j + kl, WAVE), dn = __shfl(mw.dn, 4 * j + kl, WAVE);
    const double lp = ss.log_p[bin_of(P, g, kl)];
    const double uhat = rhs / dn;
    const double sv = kl > 0 ? -0.5 * (lg - (rhs * uhat / vare)) + lp : lp;
    double s[4];
#pragma unroll
    for (int k = 0; k < 4; k++) s[k] = __shfl(sv, k, WAVE);
    Src pk = src;
    (void)pk.next();
    const double u1 = pk.next(), u2 = pk.next();
    const double rr = sqrt(-2.0 * log(u1));
    const double sn = sin(__dmul_rn(2.0 * PI_F32, u2));
    double st;
    if (EXACT) {
        // 1 / (1 + sum_{j != k} exp(s_j - s_k)) with the +-700 cut-offs
        double sk = 0.0;
        bool over = false;
#pragma unroll
        for (int jj = 0; jj < 4; jj++) {
            const double cl = s[jj] - sv;
            const double e = exp(cl);
            const bool valid = jj < nmix && jj != lane;
            over = over || (valid && cl > 700);
            sk = sk + ((valid && cl >= -700 && cl <= 700) ? e : 0.0);
        }
        st = over ? 0.0 : 1.0 / (1.0 + sk);
    } else {
        double mx = s[0];
#pragma unroll
        for (int k = 1; k < 4; k++)
            if (k < nmix) mx = fmax(mx, s[k]);
        st = exp(sv - mx);
    }
    if (lane >= nmix) st = 0.0;
    double stemp[4];
#pragma unroll
    for (int k = 0; k < 4; k++) stemp[k] = __shfl(st, k, WAVE);
    if (!EXACT) {
        double tot = 0.0;
#pragma unroll
        for (int k = 0; k < 4; k++) tot += stemp[k];
        const double inv = 1.0 / tot;
#pragma unroll
        for (int k = 0; k < 4; k++) stemp[k] = stemp[k] * inv;
    }
    const double u = src.next();
    double acc = 0.0;
    flag = 1;
#pragma unroll
    for (int k = 0; k < 4; k++) {
        if (k >= nmix) break;
        acc = acc + stemp[k];
        if (u < acc) {
            flag = k + 1;
            break;
        }
    }
    double gn = 0.0;
    if (flag > 1) {
        const double v1 = __shfl(mw.v1, 4 * j + flag - 1, WAVE), sd = __shfl(mw.sd, 4 * j + flag - 1, WAVE);
        const double mean = rhs / v1;
        if (sd <= 0.0) {
            gn = mean;
        } else {
            (void)src.next();
            (void)src.next();
            gn = __dadd_rn(mean, __dmul_rn(__dmul_rn(sd, rr), sn));   // box_muller
        }
    }
    return gn;
}

// What one window's draws need, fetched by wave 1 a round ahead: lane l
// holds Gram entries l, l+64, l+128, z-table entry l and, for l < SW, the
// state of SNP l.
struct WinRegs {
    double g0, g1, g2, z, b, v;
    int pos, grp, id, cl;
};

__device__ __forceinline__ WinRegs win_load(int m, int lane, const double *wgram, const double *wz,
                                            const int32_t *list, int nsk, const int32_t *grp,
                                            const int32_t *snp_of_pos, const double *beta, const int32_t *cls,
                                            const double *vara_s)
{
    WinRegs w;
    const double *G = wgram + (int64_t)m * SWT;
    w.g0 = G[lane];
    w.g1 = G[lane + 64];
    w.g2 = lane + 128 < SWT ? G[lane + 128] : 0.0;
    w.z = wz[(int64_t)m * 64 + lane];
    const int k = m * SW + lane;
    w.pos = (lane < SW && k < nsk) ? list[k] : -1;
    if (w.pos >= 0) {
        w.grp = grp[w.pos];
        w.id = snp_of_pos[w.pos];
        w.cl = cls[w.pos];
        w.b = beta[w.pos];
        w.v = vara_s[w.pos];
    } else {
        w.grp = w.id = w.cl = 0;
        w.b = w.v = 0.0;
    }
    return w;
}

template <int NM, int E>
__global__ __launch_bounds__(SST) void k_single(Par P, Scal *scg, const uint32_t *__restrict__ win, int64_t ldn,
                                                int nchunk, const double *__restrict__ wz,
                                                const double *__restrict__ wgram, const int32_t *__restrict__ list,
                                                int nsk, const int32_t *__restrict__ grp,
                                                const int32_t *__restrict__ snp_of_pos, int rep, double *r,
                                                double *beta, int32_t *cls, double *vara_s)
{
    __shared__ ScalLite ss;
    __shared__ double part[SW][SST / WAVE];
    __shared__ double zt[2][SW * 4];
    __shared__ double gm[2][SWT];
    __shared__ double mbeta[2][SW], mvs[2][SW];
    __shared__ int mpos[2][SW], mgrp[2][SW], mid[2][SW], mcls[2][SW];
    __shared__ double dl[SW];

    const int c = blockIdx.x;   // one workgroup per chain
    Scal *sc = scg + c;
    r += (int64_t)c * P.ldr;
    beta += (int64_t)c * P.p;
    cls += (int64_t)c * P.p;
    vara_s += (int64_t)c * P.p;
    const int t = threadIdx.x, lane = t & (WAVE - 1), wv = t / WAVE;
    const bool replay = P.rng_mode == SB_RNG_REPLAY;
    const int nwin = (nsk + SW - 1) / SW;

    auto win_store = [&](const WinRegs &w, int s) {
        gm[s][lane] = w.g0;
        gm[s][lane + 64] = w.g1;
        if (lane + 128 < SWT) gm[s][lane + 128] = w.g2;
        zt[s][lane] = w.z;
        if (lane < SW) {
            mpos[s][lane] = w.pos;
            mgrp[s][lane] = w.grp;
            mid[s][lane] = w.id;
            mcls[s][lane] = w.cl;
            mbeta[s][lane] = w.b;
            mvs[s][lane] = w.v;
        }
    };

    if (t == 0) {
        ss.vare = sc->vare;
        ss.scale_va = sc->scale_va;
    }
    if (t < MAXMIX) {
        ss.gp[t] = sc->gp[t];
        ss.vare_gp[t] = sc->vare_gp[t];
    }
    for (int b = t; b < P.nbins; b += SST) ss.log_p[b] = sc->log_p[b];
    if (wv == 1 && nwin > 0)
        win_store(win_load(0, lane, wgram, wz, list, nsk, grp, snp_of_pos, beta, cls, vara_s), 0);
    XoroSrc xs{{sc->rng[0], sc->rng[1]}};   // advanced by thread 0 only
    __syncthreads();

    for (int m = 0; m <= nwin; m++) {
        const int ps = (m + 1) & 1;   // slot of window m-1 (and of window m+1)
        const int cs = m & 1;         // slot of window m
        const int cprev = m > 0 ? min(SW, nsk - (m - 1) * SW) : 0;
        const int ccur = m < nwin ? min(SW, nsk - m * SW) : 0;
        const bool loader = wv == 1 && m + 1 < nwin;
        WinRegs nxt;
        if (loader) nxt = win_load(m + 1, lane, wgram, wz, list, nsk, grp, snp_of_pos, beta, cls, vara_s);

        double acc[SW];
#pragma unroll
        for (int j = 0; j < SW; j++) acc[j] = 0.0;
        for (int ch = 0; ch < nchunk; ch++) {
            const int64_t base = (int64_t)ch * SST * E + t;
            double rv[E];
            uint32_t gw[E];
#pragma unroll
            for (int e = 0; e < E; e++) {
                const int64_t i = base + (int64_t)e * SST;
                rv[e] = i < P.n ? r[i] : 0.0;
            }
            if (cprev > 0) {
                // the previous window's updates, SNP by SNP (models.f90:728, 768, 775, 849, 855)
#pragma unroll
                for (int e = 0; e < E; e++) {
                    const int64_t i = base + (int64_t)e * SST;
                    gw[e] = i < P.n ? win[(int64_t)(m - 1) * ldn + i] : 0u;
                }
#pragma unroll
                for (int j = 0; j < SW; j++) {
                    if (j < cprev) {
                        const double z0 = zt[ps][4 * j], z1 = zt[ps][4 * j + 1], z2 = zt[ps][4 * j + 2];
                        const double d = dl[j];
#pragma unroll
                        for (int e = 0; e < E; e++) {
                            const uint32_t x = (gw[e] >> (2 * j)) & 3u;
                            rv[e] = rv[e] + (x == 0 ? z0 : (x == 1 ? z1 : z2)) * d;
                        }
                    }
                }
#pragma unroll
                for (int e = 0; e < E; e++) {
                    const int64_t i = base + (int64_t)e * SST;
                    if (i < P.n)
                        r[i] = rv[e];
                    else
                        rv[e] = 0.0;   // padding: code 0 gave it z(0) d, which must not enter the products
                }
            }
            if (ccur > 0) {
#pragma unroll
                for (int e = 0; e < E; e++) {
                    const int64_t i = base + (int64_t)e * SST;
                    gw[e] = i < P.n ? win[(int64_t)m * ldn + i] : 0u;
                }
#pragma unroll
                for (int j = 0; j < SW; j++) {
                    if (j < ccur) {
                        const double z0 = zt[cs][4 * j], z1 = zt[cs][4 * j + 1], z2 = zt[cs][4 * j + 2];
#pragma unroll
                        for (int e = 0; e < E; e++) {
                            const uint32_t x = (gw[e] >> (2 * j)) & 3u;
                            acc[j] = fma(x == 0 ? z0 : (x == 1 ? z1 : z2), rv[e], acc[j]);
                        }
                    }
                }
            }
        }
        if (m == nwin) break;

        {
            int idx;
            const double v = wave_sum16(acc, lane, idx);
            if ((lane & 3) == 0) part[idx][wv] = v;
        }
        __syncthreads();
        if (wv == 0) {
            // lane j < SW holds rhs_j; after each draw the later lanes add
            // their Gram term
            double s = 0.0;
            if (lane < SW)
                for (int w = 0; w < SST / WAVE; w++) s += part[lane][w];
            const bool mixture = P.method != SB_BAYES_A && P.method != SB_BAYES_B;
            MixWin mw{};
            if (NM == 4 && mixture) mw = mix_window(P, ss, gm[cs], ccur, lane);
            for (int j = 0; j < ccur; j++) {
                const double sj = __shfl(s, j, WAVE);
                const double zz = gm[cs][tri(j, j)];
                const int pos = mpos[cs][j];
                const int cold = P.method == SB_BAYES_A ? 2 : mcls[cs][j];
                SnpDraw dr{};
                if (NM == 4 && mixture) {
                    int flag;
                    double gn;
                    if (replay) {
                        gn = draw_mix4<XoroSrc>(P, ss, xs, mw, j, sj, zz, mgrp[cs][j], mbeta[cs][j], cold, lane, flag);
                    } else {
                        PhiloxSrc src(chain_key(P, c), (uint32_t)rep, (uint32_t)mid[cs][j], TAG_SNP);
                        gn = draw_mix4<PhiloxSrc>(P, ss, src, mw, j, sj, zz, mgrp[cs][j], mbeta[cs][j], cold, lane,
                                                  flag);
                    }
                    dr.beta = gn;
                    dr.cls = flag;
                    dr.delta = mbeta[cs][j] - gn;
                } else if (mixture) {
                    // class terms on lanes 0..nmix-1, the draws on lane 0
                    if (replay) {
                        dr = draw_mix_wave<XoroSrc, NM>(P, &ss, xs, sj, zz, zz / ss.vare, mgrp[cs][j], mbeta[cs][j],
                                                        cold, lane);
                    } else {
                        PhiloxSrc src(chain_key(P, c), (uint32_t)rep, (uint32_t)mid[cs][j], TAG_SNP);
                        dr = draw_mix_wave<PhiloxSrc, NM>(P, &ss, src, sj, zz, zz / ss.vare, mgrp[cs][j],
                                                          mbeta[cs][j], cold, lane);
                    }
                } else if (lane == 0) {
                    if (replay) {
                        dr = draw_core<XoroSrc, NM>(P, &ss, xs, sj, zz, zz / ss.vare, mgrp[cs][j], mbeta[cs][j],
                                                    cold, mvs[cs][j]);
                    } else {
                        PhiloxSrc src(chain_key(P, c), (uint32_t)rep, (uint32_t)mid[cs][j], TAG_SNP);
                        dr = draw_core<PhiloxSrc, NM>(P, &ss, src, sj, zz, zz / ss.vare, mgrp[cs][j],
                                                      mbeta[cs][j], cold, mvs[cs][j]);
                    }
                }
                double d = 0.0;
                if (lane == 0) {
                    beta[pos] = dr.beta;
                    if (P.method != SB_BAYES_A) cls[pos] = dr.cls;
                    if (P.method == SB_BAYES_A || P.method == SB_BAYES_B) vara_s[pos] = dr.vs;
                    d = dr.delta;
                    dl[j] = d;
                }
                d = __shfl(d, 0, WAVE);
                if (lane > j && lane < SW) s = fma(gm[cs][tri(lane, j)], d, s);
            }
        }
        if (loader) win_store(nxt, ps);
        __syncthreads();
    }
    if (replay && t == 0) {
        sc->rng[0] = xs.st.s1;
        sc->rng[1] = xs.st.s2;
    }
}

template <int NM, int E>
void launch_single(sb_handle *h, int rep, int64_t ldn, int nchunk)
{
    const Par &P = h->par;
    k_single<NM, E><<<P.nchains, SST, 0, h->stream>>>(P, h->scal, h->win, ldn, nchunk, h->wz, h->wgram,
                                                      h->d_skip_list, h->nskipped, h->grp, h->d_pos_by_id + P.p, rep,
                                                      h->r, h->beta, h->cls, h->vara_s);
}

}  // namespace

int single_setup(sb_handle *h)
{
    const Par &P = h->par;
    const int nsk = h->nskipped;
    const int nwin = (nsk + SW - 1) / SW;
    const int64_t ldn = P.ldr;
    h->nwin = nwin;
    int rc = dev_alloc(h, (void **)&h->win, sizeof(uint32_t) * (size_t)nwin * ldn);
    if (!rc) rc = dev_alloc(h, (void **)&h->wz, sizeof(double) * 64 * (size_t)nwin);
    if (!rc) rc = dev_alloc(h, (void **)&h->wgram, sizeof(double) * SWT * (size_t)nwin);
    if (rc) return rc;
    if (nwin == 0) return SB_OK;
    k_win_ztab<<<nwin, 64, 0, h->stream>>>(P, h->zlut, h->d_skip_list, nsk, h->wz);
    k_win_pack<<<dim3(nwin, (unsigned)((ldn + GT - 1) / GT)), GT, 0, h->stream>>>(P, h->geno, h->d_skip_list, nsk,
                                                                                  ldn, h->win);
    k_win_gram<<<nwin, GT, 0, h->stream>>>(P, h->win, ldn, h->wz, h->wgram);
    if (int e = hip_check(hipGetLastError(), "single-site setup kernels")) return e;
    return hip_check(hipStreamSynchronize(h->stream), "single-site setup");
}

int single_sweep(sb_handle *h, int rep)
{
    const Par &P = h->par;
    // individuals per thread and pass: the smallest count that covers n in
    // one pass, up to 8 (then passes of 8 * SST)
    const int64_t need = (P.n + SST - 1) / SST;
    const int E = need <= 6 ? (int)std::max<int64_t>(need, 1) : 8;
    const int nchunk = (int)((P.n + (int64_t)SST * E - 1) / ((int64_t)SST * E));
    const int64_t ldn = P.ldr;
    with_nmix(P.nmix, [&](auto nm) {
        constexpr int NM = decltype(nm)::value;
        switch (E) {
            case 1: launch_single<NM, 1>(h, rep, ldn, nchunk); break;
            case 2: launch_single<NM, 2>(h, rep, ldn, nchunk); break;
            case 3: launch_single<NM, 3>(h, rep, ldn, nchunk); break;
            case 4: launch_single<NM, 4>(h, rep, ldn, nchunk); break;
            case 5: launch_single<NM, 5>(h, rep, ldn, nchunk); break;
            case 6: launch_single<NM, 6>(h, rep, ldn, nchunk); break;
            default: launch_single<NM, 8>(h, rep, ldn, nchunk); break;
        }
    });
    return hip_check(hipGetLastError(), "single-site sweep");
}

}  // namespace sb

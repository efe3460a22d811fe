// Per-SNP conditional draws of the SODA sweep, shared by the FP64 and the
// chain-batched paths (models.f90:145-161 BayesA, 210-255 BayesB,
// 349-412 BayesC/Cpi/R/Rc).
#pragma once
#include <type_traits>

#include "sb_device.h"
#include "sb_internal.h"

namespace sb {

__device__ __forceinline__ int bin_of(const Par &P, int g, int k) { return k * P.ngroup + g; }

// Philox counter word c2 of each kind of draw (COUNTER mode)
enum { TAG_MU = 1, TAG_SNP = 2, TAG_HYP = 3, TAG_RA = 4, TAG_INIT = 5, TAG_IMPUTE = 6 };

// Philox key of chain c (COUNTER mode): chains draw independent streams.
__device__ __forceinline__ uint64_t chain_key(const Par &P, int c)
{
    return P.seed ^ (0x9E3779B97F4A7C15ULL * (uint64_t)(c + 1));
}

// Class probabilities (models.f90:363-390) for BayesC/Cpi/R/Rc.  NM is a
// compile-time bound on the number of classes (>= P.nmix), so that the
// per-class terms stay in registers when it is small.  EXACT follows the
// reference's pairwise form 1 / (1 + sum_j exp(s_j - s_k)) with its +-700
// cut-offs (nmix^2 exponentials; REPLAY reproduces the reference's doubles);
// otherwise the same probabilities as a max-shifted softmax (nmix
// exponentials), for the counter-based sampler.
template <int NM, bool EXACT = true, class S = Scal>
__device__ inline void mix_probs(const Par &P, const S *sc, int g, double rhs, double zz,
                                 double zz_vare, double *stemp)
{
    double s[NM];
    double vare = sc->vare;
    s[0] = sc->log_p[bin_of(P, g, 0)];
#pragma unroll
    for (int k = 1; k < NM; k++) {
        if (k < P.nmix) {
            double logdetV = log(sc->gp[k] * zz_vare + 1.0);
            double uhat = rhs / (zz + sc->vare_gp[k]);
            s[k] = -0.5 * (logdetV - (rhs * uhat / vare)) + sc->log_p[bin_of(P, g, k)];
        }
    }
    if (!EXACT) {
        double mx = s[0];
#pragma unroll
        for (int k = 1; k < NM; k++)
            if (k < P.nmix) mx = fmax(mx, s[k]);
        double tot = 0.0;
#pragma unroll
        for (int k = 0; k < NM; k++) {
            stemp[k] = k < P.nmix ? exp(s[k] - mx) : 0.0;
            tot += stemp[k];
        }
        const double inv = 1.0 / tot;
#pragma unroll
        for (int k = 0; k < NM; k++) stemp[k] = stemp[k] * inv;
        return;
    }
#pragma unroll
    for (int kk = 0; kk < NM; kk++) {
        if (kk >= P.nmix) break;
        double skk = s[kk], sk = 0.0;
        bool overflow = false;
#pragma unroll
        for (int j = 0; j < NM; j++) {
            if (j >= P.nmix) break;
            if (j == kk) continue;
            double clike = s[j] - skk;
            if (clike < -700) continue;
            if (clike > 700) {
                overflow = true;
                break;
            }
            sk = sk + exp(clike);
        }
        stemp[kk] = overflow ? 0.0 : 1.0 / (1.0 + sk);
    }
}

// BayesB two-class probability of class 2 (models.f90:223-234)
template <class S>
__device__ __forceinline__ double b_prob(const S *sc, double rhs, double zz, double zz_vare,
                                         double vs)
{
    double vare = sc->vare;
    double s1 = sc->log_p[0];
    double logdetV = log(vs * zz_vare + 1.0);
    double uhat = rhs / (zz + vare / vs);
    double s2 = -0.5 * (logdetV - (rhs * uhat / vare)) + sc->log_p[1];
    double clike = s1 - s2;
    if (clike > 700) return 0.0;
    if (clike < -700) return 1.0;
    return 1.0 / (1.0 + exp(clike));
}

// New state of one SNP after its draw.
struct SnpDraw {
    double beta;   // new effect
    double vs;     // new per-SNP variance (BayesA/B)
    double delta;  // old minus new effect
    int cls;       // new class (not BayesA)
    bool chg;      // the residuals change (models.f90:416-422: gflag /= 1)
};

// Per-SNP draw of the chosen method from uniform source `src`, from the old
// effect gk, class cold (BayesA: 2) and variance vs (BayesA/B).  Mirrors
// models.f90:145-161 (A), 210-255 (B), 349-412 (C/Cpi/R/Rc).  S is Scal or
// any type with the fields read here (vare, scale_va, gp, vare_gp, log_p).
template <class Src, int NM = MAXMIX, class S = Scal>
__device__ inline SnpDraw draw_core(const Par &P, const S *sc, Src &src, double rhs0, double zz, double zz_vare,
                                   int g, double gk, int cold, double vs)
{
    SnpDraw d;
    double vare = sc->vare;
    if (P.method == SB_BAYES_A) {
        double rhs = rhs0 + zz * gk;
        double v1 = zz + vare / vs;
        double gn = draw_normal(src, rhs / v1, sqrt(vare / v1));
        double scale = (gn * gn + sc->scale_va * P.df_va) / (P.df_va + 1.0);
        d.beta = gn;
        d.vs = draw_sinvchi2(src, P.df_va + 1.0, scale);
        d.delta = gk - gn;
        d.cls = 2;
        d.chg = true;
        return d;
    }
    double rhs = rhs0;
    if (cold > 1) rhs = rhs + zz * gk;
    if (P.method == SB_BAYES_B) {
        double sk = b_prob(sc, rhs, zz, zz_vare, vs);
        double u = src.next();
        int flag = u < sk ? 2 : 1;
        double gv, gn;
        if (flag == 1) {
            gv = draw_normal(src, 0.0, sqrt(vs));
            gn = 0.0;
        } else {
            double v1 = zz + vare / vs;
            gv = draw_normal(src, rhs / v1, sqrt(vare / v1));
            gn = gv;
        }
        double scale = (gv * gv + sc->scale_va * P.df_va) / (P.df_va + 1.0);
        d.beta = gn;
        d.cls = flag;
        d.vs = draw_sinvchi2(src, P.df_va + 1.0, scale);
        d.delta = gk - gn;
        d.chg = !(flag == 1 && cold == 1);
        return d;
    }
    double stemp[NM];
    mix_probs<NM, std::is_same<Src, XoroSrc>::value>(P, sc, g, rhs, zz, zz_vare, stemp);
    double u = src.next(), ss = 0.0;
    int flag = 1;
#pragma unroll
    for (int k = 0; k < NM; k++) {
        if (k >= P.nmix) break;
        ss = ss + stemp[k];
        if (u < ss) {
            flag = k + 1;
            break;
        }
    }
    double gn = 0.0;
    if (flag > 1) {
        double v1 = zz + vare / sc->gp[flag - 1];
        gn = draw_normal(src, rhs / v1, sqrt(vare / v1));
    }
    d.beta = gn;
    d.vs = vs;
    d.cls = flag;
    d.delta = gk - gn;
    d.chg = !(flag == 1 && cold == 1);
    return d;
}

// draw_core's mixture branch (BayesC/Cpi/R/Rc) spread over the lanes of one
// wave, for a caller whose draws are serial (the single-site sweep): lane k
// forms the class term s_k and then class k's probability, each with the
// operations and in the order of mix_probs, so the doubles are draw_core's;
// lane 0 then takes the uniform and the effect from `src`.  Every lane of the
// wave calls it with the same arguments; the result is lane 0's.
template <class Src, int NM, class S>
__device__ SnpDraw draw_mix_wave(const Par &P, const S *sc, Src &src, double rhs0, double zz, double zz_vare, int g,
                                 double gk, int cold, int lane)
{
    constexpr bool EXACT = std::is_same<Src, XoroSrc>::value;
    const double vare = sc->vare;
    double rhs = rhs0;
    if (cold > 1) rhs = rhs + zz * gk;
    double sv = 0.0;
    if (lane == 0) {
        sv = sc->log_p[bin_of(P, g, 0)];
    } else if (lane < P.nmix) {
        double logdetV = log(sc->gp[lane] * zz_vare + 1.0);
        double uhat = rhs / (zz + sc->vare_gp[lane]);
        sv = -0.5 * (logdetV - (rhs * uhat / vare)) + sc->log_p[bin_of(P, g, lane)];
    }
    double s[NM];
#pragma unroll
    for (int k = 0; k < NM; k++) s[k] = __shfl(sv, k, WAVE);
    double st = 0.0;
    if (EXACT) {
        if (lane < P.nmix) {
            double sk = 0.0;
            bool overflow = false;
#pragma unroll
            for (int j = 0; j < NM; j++) {
                if (j >= P.nmix) break;
                if (j == lane) continue;
                double clike = s[j] - sv;
                if (clike < -700) continue;
                if (clike > 700) {
                    overflow = true;
                    break;
                }
                sk = sk + exp(clike);
            }
            st = overflow ? 0.0 : 1.0 / (1.0 + sk);
        }
    } else {
        double mx = s[0];
#pragma unroll
        for (int k = 1; k < NM; k++)
            if (k < P.nmix) mx = fmax(mx, s[k]);
        st = lane < P.nmix ? exp(sv - mx) : 0.0;
    }
    double stemp[NM];
#pragma unroll
    for (int k = 0; k < NM; k++) stemp[k] = __shfl(st, k, WAVE);
    SnpDraw d{};
    if (lane != 0) return d;
    if (!EXACT) {
        double tot = 0.0;
#pragma unroll
        for (int k = 0; k < NM; k++) tot += stemp[k];
        const double inv = 1.0 / tot;
#pragma unroll
        for (int k = 0; k < NM; k++) stemp[k] = stemp[k] * inv;
    }
    double u = src.next(), ss = 0.0;
    int flag = 1;
#pragma unroll
    for (int k = 0; k < NM; k++) {
        if (k >= P.nmix) break;
        ss = ss + stemp[k];
        if (u < ss) {
            flag = k + 1;
            break;
        }
    }
    double gn = 0.0;
    if (flag > 1) {
        double v1 = zz + vare / sc->gp[flag - 1];
        gn = draw_normal(src, rhs / v1, sqrt(vare / v1));
    }
    d.beta = gn;
    d.vs = 0.0;
    d.cls = flag;
    d.delta = gk - gn;
    d.chg = !(flag == 1 && cold == 1);
    return d;
}

// draw_core on the state arrays at position pos (loads, draw, stores).
template <class Src, int NM = MAXMIX>
__device__ inline void draw_snp(const Par &P, const Scal *sc, Src &src, double rhs0, double zz, double zz_vare,
                                int pos, int g, double *beta, int32_t *cls, double *vara_s, double *delta,
                                uint8_t *chg)
{
    const bool ab = P.method == SB_BAYES_A || P.method == SB_BAYES_B;
    const SnpDraw d = draw_core<Src, NM>(P, sc, src, rhs0, zz, zz_vare, g, beta[pos],
                                         P.method == SB_BAYES_A ? 2 : cls[pos], ab ? vara_s[pos] : 0.0);
    beta[pos] = d.beta;
    if (P.method != SB_BAYES_A) cls[pos] = d.cls;
    if (ab) vara_s[pos] = d.vs;
    delta[pos] = d.delta;
    chg[pos] = d.chg;
}

// Mixture conditionals of the counter-based sampler (BayesC/Cpi/R/Rc).  The
// terms that depend only on the chain, the class and the block are computed
// once per workgroup: with v1_k = zz + vare / gp_k,
//   s_k = log_p(g, k) + a_k + b_k rhs^2,  a_k = -log(gp_k zz / vare + 1) / 2,
//   b_k = 1 / (2 vare v1_k)                 (models.f90:363-371, uhat = rhs / v1_k)
// and the class-k effect ~ N(m_k rhs, sd_k^2), m_k = 1 / v1_k,
// sd_k = sqrt(vare / v1_k) (models.f90:402-405).  The same probabilities and
// effects as draw_core up to rounding; REPLAY keeps draw_core.
struct MixConst {
    double a[MAXMIX], b[MAXMIX], m[MAXMIX], sd[MAXMIX];
    double log_p[MAXBIN];   // copy of the chain's log mixture proportions
};

// Workgroup-wide: threads 1 .. nmix-1 fill class k of mc (LDS), every thread
// copies part of log_p, then sync.
__device__ inline void mix_const(const Par &P, const Scal *sc, double zz, MixConst &mc)
{
    for (int t = threadIdx.x; t < P.nbins; t += blockDim.x) mc.log_p[t] = sc->log_p[t];
    const int k = threadIdx.x;
    if (k >= 1 && k < P.nmix) {
        const double vare = sc->vare;
        const double v1 = zz + sc->vare_gp[k];
        mc.a[k] = -0.5 * log(sc->gp[k] * (zz / vare) + 1.0);
        mc.m[k] = 1.0 / v1;
        mc.b[k] = 0.5 * mc.m[k] / vare;
        mc.sd[k] = sqrt(vare * mc.m[k]);
    }
    __syncthreads();
}

template <int NM>
__device__ inline SnpDraw draw_mix_fast(const Par &P, const Scal *sc, PhiloxSrc &src, const MixConst &mc,
                                        double rhs0, double zz, int g, double gk, int cold)
{
    double rhs = rhs0;
    if (cold > 1) rhs = rhs + zz * gk;
    const double r2 = rhs * rhs;
    double s[NM];
    s[0] = mc.log_p[bin_of(P, g, 0)];
    double mx = s[0];
#pragma unroll
    for (int k = 1; k < NM; k++)
        if (k < P.nmix) {
            s[k] = mc.log_p[bin_of(P, g, k)] + mc.a[k] + mc.b[k] * r2;
            mx = fmax(mx, s[k]);
        }
    double tot = 0.0;
#pragma unroll
    for (int k = 0; k < NM; k++) {
        s[k] = k < P.nmix ? exp(s[k] - mx) : 0.0;
        tot += s[k];
    }
    const double u = src.next() * tot;   // u < cumulative sum / tot
    double ss = 0.0;
    int flag = 1;
#pragma unroll
    for (int k = 0; k < NM; k++) {
        if (k >= P.nmix) break;
        ss = ss + s[k];
        if (u < ss) {
            flag = k + 1;
            break;
        }
    }
    SnpDraw d;
    double gn = 0.0;
    if (flag > 1) {
        const double u1 = src.next(), u2 = src.next();
        gn = box_muller_pi(rhs * mc.m[flag - 1], mc.sd[flag - 1], u1, u2);
    }
    d.beta = gn;
    d.vs = 0.0;
    d.cls = flag;
    d.delta = gk - gn;
    d.chg = !(flag == 1 && cold == 1);
    return d;
}

// Host-side dispatch on a compile-time bound NM >= nmix of the class count
// (4, 8 or MAXMIX) for the draw kernels.
template <class F>
inline void with_nmix(int nmix, F &&f)
{
    if (nmix <= 4) f(std::integral_constant<int, 4>{});
    else if (nmix <= 8) f(std::integral_constant<int, 8>{});
    else f(std::integral_constant<int, MAXMIX>{});
}

}  // namespace sb

// The SODA Gibbs iteration on gfx950 (replaces soda_bayes, models.f90:9-605).
//
// Per iteration:  mu draw -> for each block: k_rhs (X_s'r + A_s'r_a, FP64),
// k_draw_* (mixture class + effect, per SNP), k_update (r += Z_s delta,
// r_a += A_s delta) -> k_stats + k_hyper (sigma2_g, sigma2_e, Dirichlet
// pi, fresh r_a) -> posterior accumulation -> resync every 1000 iterations.
//
// Genotypes stay 2-bit packed (16 per 32-bit word, block-position order).
// X_j'r is evaluated in centred form (sum_i x_ij r_i - 2 f_j sum_i r_i)/s_j;
// the residual updates use the reference's standardised values
// z = (x - 2f)/sqrt(2f(1-f)) (utilities.f90:531) from a 3-entry table.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "sb_device.h"
#include "sb_draw.h"
#include "sb_internal.h"

namespace sb {
namespace {

constexpr int NT = 256;



// ------------------------------------------------------------- mu draw --

// r = r + mu and partial sums of r (models.f90:130)
__global__ __launch_bounds__(NT) void k_mu_pre(Par P, double *r, const Scal *scal, double *red)
{
    __shared__ double lds[NT / WAVE];
    const int c = blockIdx.z;
    r += (int64_t)c * P.ldr;
    red += (int64_t)c * gridDim.x;
    double mu = scal[c].mu, s = 0.0;
    for (int i = blockIdx.x * NT + threadIdx.x; i < P.n; i += gridDim.x * NT) {
        double v = r[i] + mu;
        r[i] = v;
        s += v;
    }
    s = block_sum<NT>(s, lds);
    if (threadIdx.x == 0) red[blockIdx.x] = s;
}

// mu ~ N(sum(r)/n, vare/n) (models.f90:131) and the per-sweep constants
// log_p, vare_gp (models.f90:330-338)
// One workgroup per chain: the partial sums staged in LDS by coalesced loads,
// then added by thread 0 in the order of a serial loop.
constexpr int MU_LDS = 1024;   // doubles of staged partials
__global__ void k_mu_draw(Par P, Scal *scal, const double *red, int nred, const int *it)
{
    __shared__ double stg[MU_LDS];
    const int rep = it[0];
    const int c = blockIdx.x;
    red += (int64_t)c * nred;
    if (nred <= MU_LDS) {   // workgroup-uniform
        for (int i = threadIdx.x; i < nred; i += blockDim.x) stg[i] = red[i];
        __syncthreads();
        red = stg;
    }
    if (threadIdx.x != 0) return;
    scal += c;
    double s = 0.0;
    for (int i = 0; i < nred; i++) s += red[i];
    double nn = (double)P.n;
    double mean = s / nn, sd = sqrt(scal->vare / nn);
    if (P.rng_mode == SB_RNG_REPLAY) {
        XoroSrc src{{scal->rng[0], scal->rng[1]}};
        scal->mu = draw_normal(src, mean, sd);
        scal->rng[0] = src.st.s1;
        scal->rng[1] = src.st.s2;
    } else {
        PhiloxSrc src(chain_key(P, c), (uint32_t)rep, 0xFFFFFFF0u, TAG_MU);
        scal->mu = draw_normal(src, mean, sd);
    }
    if (P.method == SB_BAYES_B) {
        scal->log_p[0] = log(scal->P[0]);
        scal->log_p[1] = log(scal->P[1]);
    } else if (P.method != SB_BAYES_A) {
        for (int g = 0; g < P.ngroup; g++)
            for (int k = 0; k < P.nmix; k++) scal->log_p[bin_of(P, g, k)] = log(scal->P[bin_of(P, g, k)]);
        for (int k = 1; k < P.nmix; k++) scal->vare_gp[k] = scal->vare / scal->gp[k];
    }
}

__global__ void k_mu_post(Par P, double *r, const Scal *scal)
{
    const int c = blockIdx.z;
    r += (int64_t)c * P.ldr;
    double mu = scal[c].mu;
    for (int i = blockIdx.x * NT + threadIdx.x; i < P.n; i += gridDim.x * NT) r[i] = r[i] - mu;
}

// ----------------------------------------------------------- block rhs --

// Workgroups [0, nxg): X part.  WG -> (chunk c, SNP group); each wave holds
// r for 2048 individuals in registers (32 per lane) and streams `spw` SNP
// columns (a multiple of 8; one 8-byte load per lane per SNP, 512 B per
// wave, eight SNPs' loads in flight), reducing eight products at a time
// across the wave.  Workgroups [nxg, ...): A part, one wave per
// column of the block's packed triangle (Ud == nullptr) or of its dense
// n_b x n_b factor (blocks that took the pivoted-Cholesky fallback).
__global__ __launch_bounds__(NT) void k_rhs(Par P, const uint32_t *__restrict__ geno,
                                            const double *__restrict__ r,
                                            const double *__restrict__ ra,
                                            const double *__restrict__ U, int64_t tri0,
                                            const double *__restrict__ Ud, int p0, int nb,
                                            int spw, int ngs, int nxg, double *__restrict__ part,
                                            double *__restrict__ partA, double *__restrict__ partR)
{
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    // genotype workgroups first, then the factor's (interleaving the two
    // kinds in runs of 8 measured slower: 54 against 42 us per block)
    const int t = blockIdx.x;
    if (t < nxg) {
        const int c = t / ngs, sg = t % ngs;
        const int w0 = c * CHUNK_WORDS + 2 * lane;
        double rr[32];
        const double2 *r2 = reinterpret_cast<const double2 *>(r + (int64_t)w0 * 16);
#pragma unroll
        for (int e = 0; e < 16; e++) {
            double2 v = r2[e];
            rr[2 * e] = v.x;
            rr[2 * e + 1] = v.y;
        }
        if (sg == 0 && wave == 0) {
            double s = 0.0;
#pragma unroll
            for (int e = 0; e < 32; e++) s += rr[e];
            s = wave_sum(s);
            if (lane == 0) partR[c] = s;
        }
        const int j0 = (sg * 4 + wave) * spw;
        const int j1 = min(nb, j0 + spw);
        const uint32_t *gbase = geno + (int64_t)p0 * P.ldw + w0;
        // eight SNPs per step: their genotype loads in flight together
        // (clamped index), eight dot products, then one transposed
        // butterfly that leaves SNP (lane >> 3) & 7's sum in lanes 8k .. 8k+7
        for (int jb = j0; jb < j1; jb += 8) {
            uint64_t gv[8];
#pragma unroll
            for (int u = 0; u < 8; u++)
                gv[u] = *reinterpret_cast<const uint64_t *>(gbase + (int64_t)min(jb + u, j1 - 1) * P.ldw);
            double a[8];
#pragma unroll
            for (int u = 0; u < 8; u++) {
                a[u] = 0.0;
#pragma unroll
                for (int e = 0; e < 32; e++) a[u] = fma((double)(uint32_t)((gv[u] >> (2 * e)) & 3u), rr[e], a[u]);
            }
            const bool h5 = lane & 32, h4 = lane & 16, h3 = lane & 8;
            double b[4], d2[2];
#pragma unroll
            for (int k = 0; k < 4; k++)
                b[k] = (h5 ? a[k + 4] : a[k]) + __shfl_xor(h5 ? a[k] : a[k + 4], 32, WAVE);
#pragma unroll
            for (int k = 0; k < 2; k++)
                d2[k] = (h4 ? b[k + 2] : b[k]) + __shfl_xor(h4 ? b[k] : b[k + 2], 16, WAVE);
            double s = (h3 ? d2[1] : d2[0]) + __shfl_xor(h3 ? d2[0] : d2[1], 8, WAVE);
            s += __shfl_xor(s, 4, WAVE);
            s += __shfl_xor(s, 2, WAVE);
            s += __shfl_xor(s, 1, WAVE);
            const int u = (lane >> 3) & 7;
            if ((lane & 7) == 0 && jb + u < j1) part[(int64_t)c * P.blocksize + jb + u] = s;
        }
    } else {
        const int jk = (t - nxg) * 4 + wave;
        if (jk >= nb) return;
        // packed triangle, or (pivoted-Cholesky blocks) a dense n_b column;
        // 8 rows per lane in flight (a long column would otherwise wait on
        // each load in turn)
        const double *col = Ud ? Ud + (int64_t)jk * nb : U + tri0 + (int64_t)jk * (jk + 1) / 2;
        const int len = Ud ? nb : jk + 1;
        double acc = 0.0;
        for (int i0 = lane; i0 < len; i0 += 8 * WAVE) {
            double cv[8], rv[8];
#pragma unroll
            for (int u = 0; u < 8; u++) {
                const int i = min(i0 + u * WAVE, len - 1);
                cv[u] = col[i];
                rv[u] = ra[i];
            }
#pragma unroll
            for (int u = 0; u < 8; u++)
                if (i0 + u * WAVE < len) acc = fma(cv[u], rv[u], acc);
        }
        acc = wave_sum(acc);
        if (lane == 0) partA[jk] = acc;
    }
}

// rhs of block position jk (models.f90:358-361) from the partial products.
__device__ __forceinline__ double block_rhs(const Par &P, const double *part, const double *partA,
                                            const double *partR, const double *f2, const double *sdv,
                                            int p0, int jk)
{
    double sx = 0.0, sr = 0.0;
    // eight chunks' loads in flight at a time (clamped, masked), summed in
    // chunk order
    for (int c0 = 0; c0 < P.nchunk; c0 += 8) {
        double v[8], w[8];
#pragma unroll
        for (int u = 0; u < 8; u++) {
            const int c = min(c0 + u, P.nchunk - 1);
            v[u] = part[(int64_t)c * P.blocksize + jk];
            w[u] = partR[c];
        }
#pragma unroll
        for (int u = 0; u < 8; u++)
            if (c0 + u < P.nchunk) {
                sx += v[u];
                sr += w[u];
            }
    }
    int pos = p0 + jk;
    double s = sdv[pos];
    double rx = s > 0.0 ? (sx - f2[pos] * sr) / s : 0.0;
    return rx + partA[jk];
}

// COUNTER mode: one thread per SNP of the block, for chain `chain` (sc,
// beta, cls, vara_s, delta and chg point at that chain's state).
template <int NM>
__global__ __launch_bounds__(NT) void k_draw_counter(Par P, const Scal *sc, const double *part,
                                                     const double *partA, const double *partR,
                                                     const double *f2, const double *sdv,
                                                     const int32_t *grp, const int32_t *snp_of_pos,
                                                     int p0, int nb, double zz, const int *it, int chain, double *beta,
                                                     int32_t *cls, double *vara_s, double *delta,
                                                     uint8_t *chg, const uint8_t *skip)
{
    __shared__ MixConst mc;
    const bool mixm = P.method != SB_BAYES_A && P.method != SB_BAYES_B;
    if (mixm) mix_const(P, sc, zz, mc);   // workgroup-uniform
    int jk = blockIdx.x * NT + threadIdx.x;
    if (jk >= nb) return;
    int pos = p0 + jk;
    if (skip[pos]) {  // skip_j (models.f90:352): updated after the block loop
        chg[pos] = 0;
        return;
    }
    double rhs = block_rhs(P, part, partA, partR, f2, sdv, p0, jk);
    PhiloxSrc src(chain_key(P, chain), (uint32_t)it[0], (uint32_t)snp_of_pos[pos], TAG_SNP);
    if (mixm) {
        const SnpDraw d = draw_mix_fast<NM>(P, sc, src, mc, rhs, zz, grp[pos], beta[pos], cls[pos]);
        beta[pos] = d.beta;
        cls[pos] = d.cls;
        delta[pos] = d.delta;
        chg[pos] = d.chg;
    } else {
        draw_snp<PhiloxSrc, NM>(P, sc, src, rhs, zz, zz / sc->vare, pos, grp[pos], beta, cls, vara_s, delta, chg);
    }
}

// REPLAY mode: rhs in parallel, then every draw from stream 0 in block
// order on one thread (models.f90:350 with -nthreads 1).  BayesC/Cpi/R/Rc
// precompute the class probabilities in parallel; the serial walk only
// consumes uniforms; effects are finished in parallel.
__global__ __launch_bounds__(NT) void k_draw_replay(Par P, Scal *sc, const double *part,
                                                    const double *partA, const double *partR,
                                                    const double *f2, const double *sdv,
                                                    const int32_t *grp, int p0, int nb, double zz,
                                                    double *rhs_s, double *stemp_s, double *uu,
                                                    int32_t *flag_s, double *beta, int32_t *cls,
                                                    double *vara_s, double *delta, uint8_t *chg,
                                                    const uint8_t *skip)
{
    const double zz_vare = zz / sc->vare;
    const bool mix = P.method != SB_BAYES_A && P.method != SB_BAYES_B;
    for (int jk = threadIdx.x; jk < nb; jk += NT) {
        if (skip[p0 + jk]) {  // skip_j: no draw here, no stream-0 uniforms
            chg[p0 + jk] = 0;
            continue;
        }
        double rhs = block_rhs(P, part, partA, partR, f2, sdv, p0, jk);
        rhs_s[jk] = rhs;
        if (mix) {
            int pos = p0 + jk;
            double r2 = rhs;
            if (cls[pos] > 1) r2 = r2 + zz * beta[pos];
            rhs_s[jk] = r2;
            mix_probs<MAXMIX>(P, sc, grp[pos], r2, zz, zz_vare, stemp_s + (int64_t)jk * P.nmix);
        }
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        XoroSrc src{{sc->rng[0], sc->rng[1]}};
        if (mix) {
            for (int jk = 0; jk < nb; jk++) {
                if (skip[p0 + jk]) continue;
                const double *st = stemp_s + (int64_t)jk * P.nmix;
                double u = src.next(), ss = 0.0;
                int flag = 1;
                for (int k = 0; k < P.nmix; k++) {
                    ss = ss + st[k];
                    if (u < ss) {
                        flag = k + 1;
                        break;
                    }
                }
                flag_s[jk] = flag;
                if (flag > 1) {
                    uu[2 * jk] = src.next();
                    uu[2 * jk + 1] = src.next();
                }
            }
        } else {
            for (int jk = 0; jk < nb; jk++) {
                int pos = p0 + jk;
                if (skip[pos]) continue;
                draw_snp(P, sc, src, rhs_s[jk], zz, zz_vare, pos, 0, beta, cls, vara_s, delta, chg);
            }
        }
        sc->rng[0] = src.st.s1;
        sc->rng[1] = src.st.s2;
    }
    __syncthreads();
    if (!mix) return;
    const double vare = sc->vare;
    for (int jk = threadIdx.x; jk < nb; jk += NT) {
        if (skip[p0 + jk]) continue;
        int pos = p0 + jk, flag = flag_s[jk], cold = cls[pos];
        double gk = beta[pos], gn = 0.0;
        if (flag > 1) {
            double rhs = rhs_s[jk];
            double v1 = zz + vare / sc->gp[flag - 1];
            gn = box_muller(rhs / v1, sqrt(vare / v1), uu[2 * jk], uu[2 * jk + 1]);
        }
        beta[pos] = gn;
        cls[pos] = flag;
        delta[pos] = gk - gn;
        chg[pos] = !(flag == 1 && cold == 1);
    }
}

// ------------------------------------------------------- block update --

// Ordered list of the block's changed positions of chain blockIdx.z
// (models.f90:416-422: the SNPs with gflag /= 1, in block order), written
// once per block for the update kernel, with delta_j, c_j = delta_j / s_j,
// the centring constant sum_j 2 f_j c_j and, per 64-row tile of r_a, the
// first list entry whose column reaches it (U is upper triangular: row tile
// rt meets only the columns jk >= 64 rt, a suffix of the ordered list; a
// dense factor of the pivoted-Cholesky fallback meets them all).  One
// workgroup: round k covers positions 256 k .. 256 k + 255, one per thread,
// so that every load is coalesced; the loads of up to LR rounds are issued
// before the first is used, then each round is compacted in order by wave
// ballots and wave offsets.
constexpr int LR = 16;   // rounds of loads in flight
__global__ __launch_bounds__(NT) void k_list(Par P, const uint8_t *chg, const double *delta,
                                             const double *f2, const double *sdv, int p0, int nb, int *lst,
                                             double *dl, double *cl, int *count, double *kconst,
                                             int *tstart, int dense, double *ubytes)
{
    __shared__ int wsum[2][NT / WAVE];
    __shared__ double lds[NT / WAVE];
    const int64_t c = blockIdx.z;
    chg += c * P.p;
    delta += c * P.p;
    lst += c * P.blocksize;
    dl += c * P.blocksize;
    cl += c * P.blocksize;
    int *ts = tstart + c * (P.blocksize / WAVE + 2);
    const int lane = threadIdx.x & (WAVE - 1), wv = threadIdx.x / WAVE;
    const uint64_t below = lane ? (~0ull >> (WAVE - lane)) : 0ull;   // lanes < this one
    int base = 0;
    double kc = 0.0, ub = 0.0;
    for (int c0 = 0; c0 < nb; c0 += NT * LR) {
        bool fl[LR];
        double dv[LR], sv[LR], fv[LR];
#pragma unroll
        for (int u = 0; u < LR; u++) {
            const int j = c0 + u * NT + (int)threadIdx.x;
            const int pos = p0 + min(j, nb - 1);
            fl[u] = chg[pos] != 0 && j < nb;
            dv[u] = delta[pos];
            sv[u] = sdv[pos];
            fv[u] = f2[pos];
        }
#pragma unroll
        for (int u = 0; u < LR; u++) {
            const int j = c0 + u * NT + (int)threadIdx.x;
            if (c0 + u * NT >= nb) break;   // workgroup-uniform
            const uint64_t mask = __ballot(fl[u]);
            if (lane == 0) wsum[u & 1][wv] = __popcll(mask);
            __syncthreads();
            int woff = 0, tot = 0;
#pragma unroll
            for (int w = 0; w < NT / WAVE; w++) {
                const int v = wsum[u & 1][w];
                woff += w < wv ? v : 0;
                tot += v;
            }
            const int o = base + woff + __popcll(mask & below);   // entries before position j
            if (j < nb && (j & (WAVE - 1)) == 0) ts[j / WAVE] = dense ? 0 : o;
            if (fl[u]) {
                const double cj = sv[u] > 0.0 ? dv[u] / sv[u] : 0.0;
                lst[o] = j;
                dl[o] = dv[u];
                cl[o] = cj;
                kc = fma(fv[u], cj, kc);
                // kernel timing: the update's algorithmic bytes, i.e. the
                // SNP's two-bit genotype column and its column of the packed
                // upper factor (block position j holds j + 1 doubles)
                ub += (double)((P.n + 3) / 4) + 8.0 * (double)(j + 1);
            }
            base += tot;
        }
    }
    if (threadIdx.x == 0) count[c] = base;
    kc = block_sum<NT>(kc, lds);
    if (threadIdx.x == 0) kconst[c] = kc;
    if (ubytes) {
        ub = block_sum<NT>(ub, lds);
        if (threadIdx.x == 0) atomicAdd(ubytes, ub);
    }
}

// r += sum_j z_j delta_j and r_a += sum_j A_j delta_j over the changed list
// (models.f90:416-430), z = (x - 2f)/s.  The genotype part is evaluated in
// centred form: sum_j x_ij c_j - sum_j 2 f_j c_j with c_j = delta_j / s_j.
// Workgroups [0, nrw): 256 individuals each; lane l of every wave owns
// individuals 4l..4l+3 (one .bed-sized byte of every listed SNP); the UW
// waves take interleaved runs of the list (UNR loads in flight per wave) and
// their partial sums are added in wave order through LDS.  This half is
// bound by FP64 VALU work (bit extract, convert, FMA per genotype), not by
// its load instructions: a variant with one 32-bit word (16 individuals) per
// lane and four SNPs per wave load measured slower.  Workgroups after that:
// (64 rows of r_a) x (RCH list entries), each over the suffix of the list
// that reaches its row tile (U is upper triangular; tstart from k_list), so
// that the rows of the first tiles, which meet every changed column of the
// triangle, spread over many workgroups; each writes its partial sums (r_a
// is updated by k_ra_sum, in chunk order).  Their loads have clamped, valid
// addresses and are masked afterwards (no branch per load).
constexpr int UW = 16;   // waves per update workgroup
constexpr int UNR = 8;   // list entries in flight per lane
constexpr int RCH = UW * UNR * 2;   // list entries per r_a workgroup
__global__ __launch_bounds__(UW * WAVE) void k_update(Par P, const uint32_t *__restrict__ geno,
                                                      const double *__restrict__ U, int64_t tri0,
                                                      const double *__restrict__ Ud, int p0,
                                                      int nb, int nrw, const int *lst, const double *dl,
                                                      const double *cl, const int *count,
                                                      const double *kconst, const int *tstart, double *r,
                                                      double *rapart)
{
    __shared__ double red[UW][WAVE][4];
    const int64_t c = blockIdx.z;
    const int m = count[c];
    if (m == 0) return;
    lst += c * P.blocksize;
    dl += c * P.blocksize;
    cl += c * P.blocksize;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    // genotype workgroups first, then r_a's (interleaving the two kinds in
    // runs of 8 measured slower: 70 against 44 us per block)
    const int local = (int)blockIdx.x < nrw ? (int)blockIdx.x : (int)blockIdx.x - nrw;
    if ((int)blockIdx.x < nrw) {
        r += c * P.ldr;
        // neighbouring 256-individual ranges share 128-byte genotype lines:
        // keep them on one XCD (one L2)
        const int bx = xcd_contiguous(local, nrw);
        const int i0 = bx * 256 + 4 * lane;       // first of 4 individuals
        const int64_t byte = i0 >> 2;
        double acc[4] = {0.0, 0.0, 0.0, 0.0};
        // each wave walks its own run of the list; 64 entries at a time are
        // read as one vector load and taken out lane by lane (no load waits
        // on another inside the run; entries past the run carry c = 0)
        const int qa = (int)((int64_t)m * wave / UW), qz = (int)((int64_t)m * (wave + 1) / UW);
        for (int qb = qa; qb < qz; qb += WAVE) {
            const int ql = min(qb + lane, qz - 1);
            const int pos_l = p0 + lst[ql];
            const double cj_l = qb + lane < qz ? cl[ql] : 0.0;
            const int cnt = min(WAVE, qz - qb);
            for (int u0 = 0; u0 < cnt; u0 += UNR) {
                uint32_t gb[UNR];
                double cj[UNR];
#pragma unroll
                for (int u = 0; u < UNR; u++) {
                    const int src = min(u0 + u, WAVE - 1);
                    const int pos = __builtin_amdgcn_readlane(pos_l, src);
                    cj[u] = __hiloint2double(__builtin_amdgcn_readlane(__double2hiint(cj_l), src),
                                             __builtin_amdgcn_readlane(__double2loint(cj_l), src));
                    gb[u] = reinterpret_cast<const uint8_t *>(geno + (int64_t)pos * P.ldw)[byte];
                }
#pragma unroll
                for (int u = 0; u < UNR; u++)
#pragma unroll
                    for (int e = 0; e < 4; e++)
                        acc[e] = fma((double)((gb[u] >> (2 * e)) & 3u), cj[u], acc[e]);
            }
        }
#pragma unroll
        for (int e = 0; e < 4; e++) red[wave][lane][e] = acc[e];
        __syncthreads();
        if (wave == 0 && i0 < P.n) {
            const double kc = kconst[c];
#pragma unroll
            for (int e = 0; e < 4; e++) {
                double s = red[0][lane][e];
                for (int w = 1; w < UW; w++) s += red[w][lane][e];
                if (i0 + e < P.n) r[i0 + e] = r[i0 + e] + (s - kc);
            }
        }
    } else {
        const int nrt = (nb + WAVE - 1) / WAVE;
        const int rt = local % nrt, cg = local / nrt;
        // the list entries from the first column that reaches row tile rt
        const int qb = tstart[c * (P.blocksize / WAVE + 2) + rt] + cg * RCH, qe = min(m, qb + RCH);
        if (qb >= m) return;   // chunk past the list (not summed)
        rapart += c * P.blocksize * ((P.blocksize + RCH - 1) / RCH) + (int64_t)cg * P.blocksize;
        const int i = rt * WAVE + lane;
        const int ic = min(i, nb - 1);
        // column jk of the factor: packed triangle (rows <= jk) or dense n_b
        // column; the row index is clamped into the column and the value
        // masked, so that every load has a valid address and no branch
        const bool dense = Ud != nullptr;
        const double *Ub = dense ? Ud : U + tri0;
        double acc = 0.0;
        for (int q0 = qb + wave * UNR; q0 < qe; q0 += UW * UNR) {
            double uv[UNR], dv[UNR];
#pragma unroll
            for (int u = 0; u < UNR; u++) {
                const int q = q0 + u, qc = min(q, qe - 1);
                const int jk = lst[qc];
                const double d = dl[qc];
                const int64_t off = dense ? (int64_t)jk * nb : (int64_t)jk * (jk + 1) / 2;
                const double uval = Ub[off + (dense ? ic : min(ic, jk))];
                uv[u] = (q < qe && (dense || i <= jk)) ? uval : 0.0;
                dv[u] = q < qe ? d : 0.0;
            }
#pragma unroll
            for (int u = 0; u < UNR; u++) acc = fma(uv[u], dv[u], acc);
        }
        red[wave][lane][0] = acc;
        __syncthreads();
        if (wave == 0 && i < nb) {
            double s = red[0][lane][0];
            for (int w = 1; w < UW; w++) s += red[w][lane][0];
            rapart[i] = s;
        }
    }
}

// r_a += the list chunks' partial sums, in chunk order (chunks of the
// suffix of the list that reaches the row's tile).
__global__ __launch_bounds__(NT) void k_ra_sum(Par P, const double *rapart, const int *count, const int *tstart,
                                               int nb, double *ra)
{
    const int i = blockIdx.x * NT + threadIdx.x;
    const int64_t c = blockIdx.z;
    if (i >= nb) return;
    const int ncg = (count[c] - tstart[c * (P.blocksize / WAVE + 2) + i / WAVE] + RCH - 1) / RCH;
    rapart += c * P.blocksize * ((P.blocksize + RCH - 1) / RCH) + i;
    double s = 0.0;
    for (int g = 0; g < ncg; g++) s += rapart[(int64_t)g * P.blocksize];
    if (ncg > 0) ra[c * P.blocksize + i] += s;
}

// ------------------------------------------------ end-of-iteration stats --

// partial r'r over [0, n)
__global__ __launch_bounds__(NT) void k_rtr(Par P, const double *r, double *red)
{
    __shared__ double lds[NT / WAVE];
    const int c = blockIdx.z;
    r += (int64_t)c * P.ldr;
    red += (int64_t)c * gridDim.x;
    double s = 0.0;
    for (int i = blockIdx.x * NT + threadIdx.x; i < P.n; i += gridDim.x * NT) s += r[i] * r[i];
    s = block_sum<NT>(s, lds);
    if (threadIdx.x == 0) red[blockIdx.x] = s;
}

// per (group, class) bin: count and sum of g^2 over a chunk of positions
// (models.f90:496-503; BayesB 321-325; BayesA 194)
__global__ __launch_bounds__(NT) void k_bins(Par P, const double *beta, const int32_t *cls,
                                             const int32_t *grp, int chunk, double *cnt_out,
                                             double *var_out)
{
    __shared__ double lds[NT / WAVE];
    int bin = blockIdx.y, wg = blockIdx.x;
    const int ch = blockIdx.z;
    beta += (int64_t)ch * P.p;
    cls += (int64_t)ch * P.p;
    cnt_out += (int64_t)ch * P.nbins * gridDim.x;
    var_out += (int64_t)ch * P.nbins * gridDim.x;
    int g = bin % P.ngroup, k = bin / P.ngroup;
    int a = wg * chunk, e = min(P.p, a + chunk);
    double c = 0.0, v = 0.0;
    for (int pos = a + threadIdx.x; pos < e; pos += NT) {
        bool in = P.method == SB_BAYES_A ? true : (grp[pos] == g && cls[pos] == k + 1);
        if (in) {
            c += 1.0;
            v += beta[pos] * beta[pos];
        }
    }
    c = block_sum<NT>(c, lds);
    v = block_sum<NT>(v, lds);
    if (threadIdx.x == 0) {
        cnt_out[(int64_t)bin * gridDim.x + wg] = c;
        var_out[(int64_t)bin * gridDim.x + wg] = v;
    }
}

// Posterior sums of one (SNP, chain) sample, from its effect b and class k
// (models.f90:528-544); the chain's arrays.  A class-1 draw sets the effect
// to exactly 0.0 (sb_draw.h draw_core) and every sample adds 1.0 to exactly
// one class of each SNP, so class 1's count is the number of samples minus
// the other classes' counts: only classes above 1 are read, modified and
// written, the samples are counted in it[2], and k_class1 writes class 1's
// counts where the sums are read out (the same integers in FP64).  Adding a
// zero effect leaves the effect sum's bits unchanged (the sum is never
// -0.0), so zero effects skip it too.
__device__ __forceinline__ void accum_one(const Par &P, int pos, double b, int k, const double *vara_s,
                                          double *gstore, double *indiststore, double *vara_sstore)
{
    if (P.method == SB_BAYES_A || k > 1) {
        if (b != 0.0) gstore[pos] += b;
        if (P.method != SB_BAYES_A) indiststore[(int64_t)(k - 1) * P.p + pos] += 1.0;
    }
    if (P.method == SB_BAYES_A || P.method == SB_BAYES_B) vara_sstore[pos] += vara_s[pos];
}

// The same, every bin in one pass over the chunk (nbins <= NBM): each thread
// adds to its bins' sums in the order of the per-bin pass, so the sums are
// bit-identical, and beta / cls / grp are read once instead of once per bin
// (four positions' loads in flight).  When the iteration's sample goes into
// the posterior sums (it[0] past burn-in and on the thinning grid), the same
// pass adds it (accum_one): nothing between this launch and the end of the
// iteration changes an effect, class or per-SNP variance.
template <int NBM>
__global__ __launch_bounds__(NT) void k_bins_all(Par P, const double *beta, const int32_t *cls,
                                                 const int32_t *grp, int chunk, double *cnt_out, double *var_out,
                                                 int *it, const double *vara_s, double *gstore,
                                                 double *indiststore, double *vara_sstore)
{
    __shared__ double lds[NT / WAVE];
    const int wg = blockIdx.x, ch = blockIdx.z;
    const int rep = it[0];
    const bool acc = rep > P.burnin && rep % P.thin == 0;   // models.f90:526
    if (acc && wg == 0 && ch == 0 && threadIdx.x == 0) it[2] += 1;   // samples in the sums
    beta += (int64_t)ch * P.p;
    cls += (int64_t)ch * P.p;
    vara_s += (int64_t)ch * P.p;
    gstore += (int64_t)ch * P.p;
    vara_sstore += (int64_t)ch * P.p;
    indiststore += (int64_t)ch * P.p * P.nmix;
    cnt_out += (int64_t)ch * P.nbins * gridDim.x;
    var_out += (int64_t)ch * P.nbins * gridDim.x;
    const bool all = P.method == SB_BAYES_A;
    const int a = wg * chunk, e = min(P.p, a + chunk);
    double c[NBM], v[NBM];
#pragma unroll
    for (int b = 0; b < NBM; b++) c[b] = v[b] = 0.0;
    constexpr int U = 4;
    for (int p0 = a + threadIdx.x; p0 < e; p0 += U * NT) {
        int gq[U], kc[U];
        double bt[U];
#pragma unroll
        for (int u = 0; u < U; u++) {
            const int pos = min(p0 + u * NT, e - 1);
            gq[u] = grp[pos];
            kc[u] = cls[pos];
            bt[u] = beta[pos];
        }
#pragma unroll
        for (int u = 0; u < U; u++) {
            const int pos = p0 + u * NT;
            if (pos >= e) break;
            const int kq = kc[u] - 1;   // bin (class kq, group gq), as k_bins matches it
            const int bp = (gq[u] >= 0 && gq[u] < P.ngroup && kq >= 0 && kq < P.nmix) ? kq * P.ngroup + gq[u] : -1;
            const double x = bt[u] * bt[u];
#pragma unroll
            for (int b = 0; b < NBM; b++) {
                const bool in = all || b == bp;
                c[b] += in ? 1.0 : 0.0;
                v[b] += in ? x : 0.0;
            }
            if (acc) accum_one(P, pos, bt[u], kc[u], vara_s, gstore, indiststore, vara_sstore);
        }
    }
    for (int b = 0; b < P.nbins; b++) {   // workgroup-uniform
        double cb = 0.0, vb = 0.0;
#pragma unroll
        for (int u = 0; u < NBM; u++) {
            cb = u == b ? c[u] : cb;
            vb = u == b ? v[u] : vb;
        }
        cb = block_sum<NT>(cb, lds);
        vb = block_sum<NT>(vb, lds);
        if (threadIdx.x == 0) {
            cnt_out[(int64_t)b * gridDim.x + wg] = cb;
            var_out[(int64_t)b * gridDim.x + wg] = vb;
        }
    }
}

// Hyperparameter draws, stores and the trace record (models.f90:496-569).
// One workgroup per chain: its threads stage the chain's partial sums (bins,
// r'r) in LDS with coalesced loads, then thread 0 adds them in the order of
// a serial loop and draws.
constexpr int HYP_LDS = 4096;   // doubles of staged partials
__global__ void k_hyper(Par P, Scal *sc, const double *red_rtr, int nrtr, const double *cnt,
                        const double *var, int nwg, int *it, double *hyp)
{
    __shared__ double stg[HYP_LDS];
    const int ch = blockIdx.x;
    const int nb = P.nbins;
    const int64_t nbw = (int64_t)nb * nwg;
    red_rtr += (int64_t)ch * nrtr;
    cnt += (int64_t)ch * nbw;
    var += (int64_t)ch * nbw;
    if (2 * nbw + nrtr <= HYP_LDS) {   // workgroup-uniform
        for (int i = threadIdx.x; i < nbw; i += blockDim.x) {
            stg[i] = cnt[i];
            stg[nbw + i] = var[i];
        }
        for (int i = threadIdx.x; i < nrtr; i += blockDim.x) stg[2 * nbw + i] = red_rtr[i];
        __syncthreads();
        cnt = stg;
        var = stg + nbw;
        red_rtr = stg + 2 * nbw;
    }
    if (threadIdx.x != 0) return;
    // it[0]: this iteration; it[1]: its record index in hyp, read and advanced
    // by chain 0's workgroup only (models.f90:526, 549)
    const int rep = it[0];
    const bool store = rep > P.burnin && rep % P.thin == 0;
    const bool record = hyp != nullptr && rep % P.thin == 0;
    const int hyp_idx = record && ch == 0 ? it[1] : -1;
    sc += ch;
    if (ch != 0) hyp = nullptr;
    for (int b = 0; b < nb; b++) {
        double c = 0.0, v = 0.0;
        for (int w = 0; w < nwg; w++) {
            c += cnt[(int64_t)b * nwg + w];
            v += var[(int64_t)b * nwg + w];
        }
        sc->snpindist[b] = c;
        sc->varindist[b] = v;
    }
    XoroSrc xs{{sc->rng[0], sc->rng[1]}};
    PhiloxSrc ps(chain_key(P, ch), (uint32_t)rep, 0xFFFFFFF1u, TAG_HYP);
    const bool replay = P.rng_mode == SB_RNG_REPLAY;
    if (P.method == SB_BAYES_A) {
        sc->snpindist[0] = 0.0;  // BayesA keeps a single variance bin
    } else if (P.method == SB_BAYES_B) {
        double inc = 0.0, v2 = 0.0;
        inc = sc->snpindist[1];
        v2 = sc->varindist[1];
        sc->included = inc;
        sc->snpindist[0] = (double)P.p - inc;
        sc->snpindist[1] = inc;
        sc->varindist[0] = 0.0;
        sc->varindist[1] = v2;
    } else {
        double excl = 0.0;
        for (int g = 0; g < P.ngroup; g++) excl += sc->snpindist[bin_of(P, g, 0)];
        sc->included = (double)P.p - excl;
        double scale = 0.0;
        for (int g = 0; g < P.ngroup; g++)
            for (int k = 1; k < P.nmix; k++) scale = scale + sc->varindist[bin_of(P, g, k)] / P.mix[k];
        scale = (scale + sc->scale_va * P.df_va) / (P.df_va + sc->included);
        double dof = sc->included + P.df_va;
        sc->vara = replay ? draw_sinvchi2(xs, dof, scale) : draw_sinvchi2(ps, dof, scale);
        for (int k = 0; k < P.nmix; k++) sc->gp[k] = P.mix[k] * sc->vara;
    }
    double rtr = 0.0;
    for (int i = 0; i < nrtr; i++) rtr += red_rtr[i];
    double nn = (double)P.n;
    double scale = (rtr + P.scale_ve * P.df_ve) / (nn + P.df_ve);
    sc->vare = replay ? draw_sinvchi2(xs, nn + P.df_ve, scale) : draw_sinvchi2(ps, nn + P.df_ve, scale);
    if (P.method == SB_BAYES_CPI || P.method == SB_BAYES_R || P.method == SB_BAYES_RC) {
        for (int g = 0; g < P.ngroup; g++) {
            double x[MAXMIX], sx = 0.0;
            for (int k = 0; k < P.nmix; k++) {
                double a = sc->snpindist[bin_of(P, g, k)] + P.alpha[k];
                x[k] = replay ? draw_gamma(xs, a, 1.0) : draw_gamma(ps, a, 1.0);
            }
            for (int k = 0; k < P.nmix; k++) sx += x[k];
            for (int k = 0; k < P.nmix; k++) sc->P[bin_of(P, g, k)] = x[k] / sx;
        }
    }
    sc->rng[0] = xs.st.s1;
    sc->rng[1] = xs.st.s2;
    if (store) {
        sc->counter += 1.0;
        sc->mu_store += sc->mu;
        // the single-site sampler sums varindist for BayesA only in SODA runs
        // (models.f90:530 against 888-899)
        if (!(P.regular && P.method == SB_BAYES_A))
            for (int b = 0; b < nb; b++) sc->varstore[b] += sc->varindist[b];
        sc->vare_store += sc->vare;
        if (P.method != SB_BAYES_A) {
            sc->included_store += sc->included;
            for (int b = 0; b < nb; b++) sc->snpstore[b] += sc->snpindist[b];
            if (P.method != SB_BAYES_B && P.method != SB_BAYES_C)
                for (int b = 0; b < nb; b++) sc->pstore[b] += sc->P[b];
        }
        if (P.method != SB_BAYES_A && P.method != SB_BAYES_B) sc->vara_store += sc->vara;
    }
    if (hyp && hyp_idx >= 0) {
        double *h = hyp + (int64_t)hyp_idx * (5 + 2 * nb);
        h[0] = rep;
        h[1] = sc->included;
        h[2] = sc->mu;
        h[3] = sc->vara;
        h[4] = sc->vare;
        for (int b = 0; b < nb; b++) {
            h[5 + b] = sc->snpindist[b];
            h[5 + nb + b] = sc->varindist[b];
        }
    }
    if (ch == 0 && record) it[1] = hyp_idx + 1;
}

// fresh r_a ~ N(0, vare) (models.f90:523-525; init 118-120)
__global__ void k_ra_replay(Par P, Scal *sc, double *ra)
{
    if (threadIdx.x != 0) return;
    XoroSrc src{{sc->rng[0], sc->rng[1]}};
    double sd = sqrt(sc->vare);
    for (int i = 0; i < P.naind; i++) ra[i] = draw_normal(src, 0.0, sd);
    sc->rng[0] = src.st.s1;
    sc->rng[1] = src.st.s2;
}

__global__ void k_ra_counter(Par P, const Scal *sc, double *ra, const int *it, int rep, int tag)
{
    if (it) rep = it[0];   // the sweep's iteration (rep: the initial draws' fixed counter)
    int i = blockIdx.x * NT + threadIdx.x;
    const int c = blockIdx.z;
    if (i >= P.naind) return;
    PhiloxSrc src(chain_key(P, c), (uint32_t)rep, (uint32_t)i, tag);
    const double v = draw_normal(src, 0.0, sqrt(sc[c].vare));
    ra[(int64_t)c * P.blocksize + i] = v;
}

// posterior sums per SNP (models.f90:528-544), for the sweeps whose bin
// sums take one launch per bin (k_bins): accum_one per (SNP, chain)
__global__ void k_accum(Par P, int *it, const double *beta, const int32_t *cls, const double *vara_s,
                        double *gstore, double *indiststore, double *vara_sstore)
{
    const int rep = it[0];
    if (!(rep > P.burnin && rep % P.thin == 0)) return;   // models.f90:526
    int pos = blockIdx.x * NT + threadIdx.x;
    if (blockIdx.x == 0 && blockIdx.z == 0 && threadIdx.x == 0) it[2] += 1;   // samples in the sums
    if (pos >= P.p) return;
    const int64_t c = blockIdx.z;
    accum_one(P, pos, beta[c * P.p + pos], P.method == SB_BAYES_A ? 2 : cls[c * P.p + pos], vara_s + c * P.p,
              gstore + c * P.p, indiststore + c * P.p * P.nmix, vara_sstore + c * P.p);
}

// Class 1's posterior counts: the samples (it[2]) minus the other classes'
// counts, per (SNP, chain); not used by BayesA, which keeps no class counts.
__global__ void k_class1(Par P, const int *it, double *indiststore)
{
    const int pos = blockIdx.x * NT + threadIdx.x;
    if (pos >= P.p) return;
    double *d = indiststore + (int64_t)blockIdx.z * P.p * P.nmix;
    double o = 0.0;
    for (int k = 1; k < P.nmix; k++) o += d[(int64_t)k * P.p + pos];
    d[pos] = (double)it[2] - o;
}

// it[2] = the samples in the posterior sums of a loaded state: the sum of
// SNP 0's class counts of chain 0 (every sample counted one class of it)
__global__ void k_samples_from_counts(Par P, const double *indiststore, int *it)
{
    double s = 0.0;
    for (int k = 0; k < P.nmix; k++) s += indiststore[(int64_t)k * P.p];
    it[2] = P.method == SB_BAYES_A ? 0 : (int)s;
}

// ------------------------------------------------------------- resync --

// Positions with g != 0 in SNP-id order (one workgroup).
__global__ __launch_bounds__(NT) void k_nonzero(Par P, const int32_t *pos_by_id, const double *beta,
                                                int32_t *list, int32_t *count)
{
    __shared__ int cnt[NT + 1];
    const int64_t ch = blockIdx.z;
    beta += ch * P.p;
    list += ch * P.p;
    count += ch;
    const int R = (P.p + NT - 1) / NT;
    int a = threadIdx.x * R, e = min(P.p, a + R), c = 0;
    for (int id = a; id < e; id++) c += beta[pos_by_id[id]] != 0.0;
    cnt[threadIdx.x] = c;
    __syncthreads();
    if (threadIdx.x == 0) {
        int s = 0;
        for (int i = 0; i < NT; i++) {
            int v = cnt[i];
            cnt[i] = s;
            s += v;
        }
        *count = s;
    }
    __syncthreads();
    int o = cnt[threadIdx.x];
    for (int id = a; id < e; id++) {
        int pos = pos_by_id[id];
        if (beta[pos] != 0.0) list[o++] = pos;
    }
}

// compute_residuals (utilities.f90:609-614): r_i = y_i - sum_j z_ij g_j - mu,
// the sum over SNP ids in order (zero effects add nothing).
__global__ __launch_bounds__(NT) void k_resync(Par P, const uint32_t *geno, const double *zlut,
                                               const double *beta, const int32_t *list,
                                               const int32_t *count, const double *why,
                                               const Scal *sc, double *r)
{
    int w = blockIdx.x * NT + threadIdx.x;
    if (w >= P.ldw || w * 16 >= P.n) return;
    const int64_t ch = blockIdx.z;
    why += ch * P.ldr;
    beta += ch * P.p;
    list += ch * P.p;
    count += ch;
    sc += ch;
    r += ch * P.ldr;
    double acc[16];
#pragma unroll
    for (int e = 0; e < 16; e++) acc[e] = 0.0;
    int m = *count;
    for (int q = 0; q < m; q++) {
        int pos = list[q];
        uint32_t gw = geno[(int64_t)pos * P.ldw + w];
        const double *zl = zlut + 4 * (int64_t)pos;
        double z0 = zl[0], z1 = zl[1], z2 = zl[2], b = beta[pos];
#pragma unroll
        for (int e = 0; e < 16; e++) {
            uint32_t x = (gw >> (2 * e)) & 3u;
            double z = x == 0 ? z0 : (x == 1 ? z1 : z2);
            acc[e] = __dadd_rn(acc[e], __dmul_rn(z, b));
        }
    }
    double mu = sc->mu;
    int i0 = w * 16;
#pragma unroll
    for (int e = 0; e < 16; e++)
        if (i0 + e < P.n) r[i0 + e] = __dsub_rn(__dsub_rn(why[i0 + e], acc[e]), mu);
}

// ---------------------------------------------------------------- init --

// Draws of the initial state (models.f90:115-120): sigma2_e from y'y of the
// initial residuals, then r_a.
// (each chain's y'y of its initial residuals is in its sc->rtr)
__global__ void k_init_draws(Par P, Scal *sc)
{
    const int c = threadIdx.x;
    if (c >= P.nchains) return;
    sc += c;
    double nn = (double)P.n;
    double scale = (sc->rtr + P.scale_ve * P.df_ve) / (nn + P.df_ve);
    if (P.rng_mode == SB_RNG_REPLAY) {
        XoroSrc src{{sc->rng[0], sc->rng[1]}};
        sc->vare = draw_sinvchi2(src, nn + P.df_ve, scale);
        sc->rng[0] = src.st.s1;
        sc->rng[1] = src.st.s2;
    } else {
        PhiloxSrc src(chain_key(P, c), 0xFFFFFFFEu, 0xFFFFFFF2u, TAG_INIT);
        sc->vare = draw_sinvchi2(src, nn + P.df_ve, scale);
    }
}

__global__ void k_init_r(Par P, const double *why, const Scal *sc, double *r)
{
    int i = blockIdx.x * NT + threadIdx.x;
    const int64_t c = blockIdx.z;
    if (i < P.n) r[c * P.ldr + i] = (why[c * P.ldr + i] - 0.0) - sc[c].mu;
}

// Held-out rows (validation individuals of a fold) are missing phenotypes:
// given mu, g and vare their full conditional is y_i ~ N(mu + z_i'g, vare),
// so after vare's update every held-out y_i is redrawn, which makes its
// residual a fresh N(0, vare) draw.  The phenotype keeps the drawn value
// (y = y - r + e), so that the resync recomputes the same residual.
__global__ void k_impute(Par P, const Scal *sc, const uint8_t *heldout, double *why, double *r, const int *it)
{
    const int rep = it[0];
    const int i = blockIdx.x * NT + threadIdx.x;
    const int64_t c = blockIdx.z;
    if (i >= P.n) return;
    const int64_t o = c * P.ldr + i;
    if (!heldout[o]) return;
    PhiloxSrc src(chain_key(P, (int)c), (uint32_t)rep, (uint32_t)i, TAG_IMPUTE);
    const double e = draw_normal(src, 0.0, sqrt(sc[c].vare));
    why[o] = (why[o] - r[o]) + e;
    r[o] = e;
}

// -------------------------------------------------------- skipped SNPs --

// z'z of every skipped SNP (xpx = dot_product(X(:,j), X(:,j)), models.f90:
// 100-104), z from the position's standardisation table; one workgroup per
// skipped SNP.
__global__ __launch_bounds__(NT) void k_xpx(Par P, const uint32_t *__restrict__ geno,
                                            const double *__restrict__ zlut, const int32_t *list, double *xpx)
{
    __shared__ double lds[NT / WAVE];
    const int pos = list[blockIdx.x];
    const double *zl = zlut + 4 * (int64_t)pos;
    const uint32_t *gw = geno + (int64_t)pos * P.ldw;
    double s = 0.0;
    for (int i = threadIdx.x; i < P.n; i += NT) {
        const double z = zl[(gw[i >> 4] >> (2 * (i & 15))) & 3u];
        s = fma(z, z, s);
    }
    s = block_sum<NT>(s, lds);
    if (threadIdx.x == 0) xpx[blockIdx.x] = s;
}

// SNPs skipped by the pivoted-Cholesky fallback, after the block loop in
// i_skipped order (models.f90:174-190 BayesA, 276-320 BayesB, 433-495
// BayesC/Cpi/R/Rc): rhs = z'r with zz = z'z and no augmented part, the
// method's draw, then r += z (g_old - g_new) at once, so each SNP sees the
// previous one's update.  One workgroup: the products are block reductions
// over the individuals and the draw runs on thread 0 (stream 0 in REPLAY
// mode, the SNP's own Philox counter in COUNTER mode).  The reference reads
// log_p / vare_gp in the BayesB and mixture loops without setting them
// there; the values of the block sweep are used.
constexpr int SKT = 1024;
template <int NM>
__global__ __launch_bounds__(SKT) void k_skipped(Par P, Scal *sc, const uint32_t *__restrict__ geno,
                                                 const double *__restrict__ zlut, const int32_t *list, int nsk,
                                                 const double *xpx, const int32_t *grp,
                                                 const int32_t *snp_of_pos, int rep, double *r, double *beta,
                                                 int32_t *cls, double *vara_s, double *delta, uint8_t *chg)
{
    __shared__ double lds[SKT / WAVE];
    __shared__ double dsh;
    const bool replay = P.rng_mode == SB_RNG_REPLAY;
    const int c = blockIdx.x;   // one workgroup per chain
    sc += c;
    r += (int64_t)c * P.ldr;
    beta += (int64_t)c * P.p;
    cls += (int64_t)c * P.p;
    vara_s += (int64_t)c * P.p;
    delta += (int64_t)c * P.p;
    chg += (int64_t)c * P.p;
    XoroSrc xs{{sc->rng[0], sc->rng[1]}};
    for (int k = 0; k < nsk; k++) {
        const int pos = list[k];
        const double *zl = zlut + 4 * (int64_t)pos;
        const double z0 = zl[0], z1 = zl[1], z2 = zl[2];
        const uint32_t *gw = geno + (int64_t)pos * P.ldw;
        double s = 0.0;
        for (int w = threadIdx.x; w * 16 < P.n; w += SKT) {
            const uint32_t g = gw[w];
#pragma unroll
            for (int e = 0; e < 16; e++) {
                const int i = w * 16 + e;
                const uint32_t x = (g >> (2 * e)) & 3u;
                if (i < P.n) s = fma(x == 0 ? z0 : (x == 1 ? z1 : z2), r[i], s);
            }
        }
        s = block_sum<SKT>(s, lds);
        if (threadIdx.x == 0) {
            const double zz = xpx[k];
            if (replay) {
                draw_snp<XoroSrc, NM>(P, sc, xs, s, zz, zz / sc->vare, pos, grp[pos], beta, cls, vara_s, delta,
                                      chg);
            } else {
                PhiloxSrc src(chain_key(P, c), (uint32_t)rep, (uint32_t)snp_of_pos[pos], TAG_SNP);
                draw_snp<PhiloxSrc, NM>(P, sc, src, s, zz, zz / sc->vare, pos, grp[pos], beta, cls, vara_s,
                                        delta, chg);
            }
            dsh = delta[pos];
        }
        __syncthreads();
        const double d = dsh;
        for (int w = threadIdx.x; w * 16 < P.n; w += SKT) {
            const uint32_t g = gw[w];
#pragma unroll
            for (int e = 0; e < 16; e++) {
                const int i = w * 16 + e;
                const uint32_t x = (g >> (2 * e)) & 3u;
                if (i < P.n) r[i] = r[i] + (x == 0 ? z0 : (x == 1 ? z1 : z2)) * d;
            }
        }
        __syncthreads();
    }
    if (replay && threadIdx.x == 0) {
        sc->rng[0] = xs.st.s1;
        sc->rng[1] = xs.st.s2;
    }
}

}  // namespace

// X_s'r + U_s'r_a of block position jk for the FP64 path (diagnostics)
__global__ void k_rhs_out(Par P, const double *part, const double *partA, const double *partR,
                          const double *f2, const double *sdv, int p0, int nb, double *out)
{
    int jk = blockIdx.x * NT + threadIdx.x;
    if (jk < nb) out[jk] = block_rhs(P, part, partA, partR, f2, sdv, p0, jk);
}

// ------------------------------------------------------------- host side --

// dense factor of block b (pivoted-Cholesky fallback), or nullptr
// The sweep's device iteration state: it[0] = the iteration being run,
// it[1] = the index of its .hyp record (it[2]: the samples in the posterior
// sums, counted by the pass that adds them, cleared by mcmc_init).  Set at the start of each
// sb_mcmc_run call; every iteration begins by advancing it[0], so that one
// captured iteration replays for any iteration number.
__global__ void k_set_it(int *it, int rep, int rec)
{
    it[0] = rep;
    it[1] = rec;
}
__global__ void k_next_it(int *it) { it[0] += 1; }

// Diagnostics of the counter-based generator: for site k, `per` uniforms of
// the stream the sampler's kernels open as PhiloxSrc(chain_key(seed, chain),
// ctr[3k], ctr[3k+1], ctr[3k+2]) (iteration, site, tag), in draw order.
__global__ void k_counter_uniforms(uint64_t seed, int chain, const uint32_t *ctr, int n, int per, double *out)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    Par P{};
    P.seed = seed;
    PhiloxSrc src(chain_key(P, chain), ctr[3 * k], ctr[3 * k + 1], ctr[3 * k + 2]);
    for (int u = 0; u < per; u++) out[(int64_t)k * per + u] = src.next();
}

int counter_uniforms(int64_t seed, int chain, const uint32_t *ctr, int n, int per, double *out)
{
    TmpDev tmp;
    uint32_t *dc = nullptr;
    double *dout = nullptr;
    int rc;
    if ((rc = tmp.alloc(&dc, sizeof(uint32_t) * 3 * (size_t)n, "counters")) ||
        (rc = tmp.alloc(&dout, sizeof(double) * (size_t)n * per, "uniforms")))
        return rc;
    rc = hip_check(hipMemcpy(dc, ctr, sizeof(uint32_t) * 3 * (size_t)n, hipMemcpyHostToDevice), "upload counters");
    if (!rc) {
        k_counter_uniforms<<<(n + NT - 1) / NT, NT>>>((uint64_t)seed, chain, dc, n, per, dout);
        rc = hip_check(hipGetLastError(), "k_counter_uniforms");
    }
    if (!rc) rc = hip_check(hipDeviceSynchronize(), "k_counter_uniforms");
    if (!rc)
        rc = hip_check(hipMemcpy(out, dout, sizeof(double) * (size_t)n * per, hipMemcpyDeviceToHost),
                       "download uniforms");
    return rc;
}

static const double *dense_block(const sb_handle *h, int b)
{
    return (b < (int)h->dense_off.size() && h->dense_off[b] >= 0) ? h->Ud + h->dense_off[b] : nullptr;
}

int mcmc_init(sb_handle *h, const sb_priors *pr, const double *why, int64_t ldy, const uint8_t *heldout,
              int64_t ldh)
{
    if (!h->aug_built) return fail(SB_ERR_STATE, "sb_mcmc_init before sb_build_aug");
    Par &P = h->par;
    if (heldout && P.rng_mode != SB_RNG_COUNTER)
        return fail(SB_ERR_ARG, "held-out rows need SB_RNG_COUNTER (their draws have no reference order)");
    P.pizero = pr->pizero;
    P.scale_va0 = pr->scale_va;
    P.scale_ve = pr->scale_ve;
    P.df_va = pr->df_va;
    P.df_ve = pr->df_ve;
    P.vara0 = pr->vara0;
    for (int k = 0; k < MAXMIX; k++) {
        P.mix[k] = pr->mix[k];
        P.alpha[k] = pr->alpha[k];
    }
    if (P.method == SB_BAYES_C || P.method == SB_BAYES_CPI) {  // useroptions.f90:503-507
        P.mix[0] = 0.0;
        P.mix[1] = 1.0;
    }
    P.numit = pr->numit;
    P.burnin = pr->burnin;
    P.thin = pr->thin > 0 ? pr->thin : 1;
    P.resync = pr->resync > 0 ? pr->resync : 1000;   // models.f90:571
    if (P.rng_mode == SB_RNG_COUNTER) P.seed = (uint64_t)pr->seed;
    const int n = P.n, p = P.p, nmix = P.nmix, ngroup = P.ngroup;
    const int C = P.nchains;
    const size_t Cs = (size_t)C;

    // Host part of models.f90:56-114 (no draws), per chain: sums in the
    // reference's sequential order so the initial scalars are the
    // reference's doubles.  A chain's held-out rows enter at the mean of its
    // training rows (residual 0) and are redrawn every iteration.
    Scal base;
    std::memset(&base, 0, sizeof base);
    if (hip_check(hipMemcpy(&base.rng, &h->scal->rng, sizeof base.rng, hipMemcpyDeviceToHost), "rng state"))
        return SB_ERR_HIP;
    auto IDX = [&](int g, int k) { return k * ngroup + g; };
    if (P.method != SB_BAYES_A) {
        if (P.method == SB_BAYES_B || P.method == SB_BAYES_C) {
            for (int g = 0; g < ngroup; g++) {
                base.P[IDX(g, 0)] = P.pizero;
                base.P[IDX(g, 1)] = 1.0 - P.pizero;
            }
        } else {
            for (int g = 0; g < ngroup; g++) {
                base.P[IDX(g, 0)] = 0.5;
                double s = 0.0;
                for (int k = 1; k < nmix; k++) base.P[IDX(g, k)] = 1.0 / P.mix[k];
                for (int k = 1; k < nmix; k++) s += base.P[IDX(g, k)];
                for (int k = 1; k < nmix; k++) base.P[IDX(g, k)] = 0.5 * base.P[IDX(g, k)] / s;
            }
        }
    }
    std::vector<Scal> scs(Cs, base);
    std::vector<double> ys((size_t)P.ldr * Cs, 0.0);
    std::vector<uint8_t> held;
    if (heldout) held.assign((size_t)P.ldr * Cs, 0);
    const double nn = (double)n;
    for (int c = 0; c < C; c++) {
        const double *yc = why + (int64_t)c * ldy;
        double *y = ys.data() + (size_t)c * P.ldr;
        if (heldout) {
            const uint8_t *hc = heldout + (int64_t)c * ldh;
            double st = 0.0;
            int nt = 0;
            for (int i = 0; i < n; i++)
                if (!hc[i]) {
                    st += yc[i];
                    nt++;
                }
            if (nt == 0) return fail(SB_ERR_ARG, "a chain has no training rows");
            const double tmean = st / nt;
            for (int i = 0; i < n; i++) {
                held[(size_t)c * P.ldr + i] = hc[i] ? 1 : 0;
                y[i] = hc[i] ? tmean : yc[i];
            }
        } else {
            for (int i = 0; i < n; i++) y[i] = yc[i];
        }
        Scal &sc = scs[c];
        double sw = 0.0;
        for (int i = 0; i < n; i++) sw += y[i];
        sc.mu = sw / nn;
        double vara = P.vara0, scale_va = P.scale_va0;
        if (P.df_va <= 0.0) {
            double yhat = sw / nn, vary = 0.0;
            for (int i = 0; i < n; i++) vary += (y[i] - yhat) * (y[i] - yhat);
            vary = vary / (nn - 1.0);
            if (P.method == SB_BAYES_A) {
                vara = scale_va * vary / p;
            } else if (P.method == SB_BAYES_B) {
                vara = scale_va * vary / p / sc.P[IDX(0, 1)];
            } else {
                for (int g = 0; g < ngroup; g++) {
                    double d = 0.0;
                    for (int k = 0; k < nmix; k++) d += sc.P[IDX(g, k)] * P.mix[k];
                    vara = vara + d;
                }
                vara = scale_va * vary / p / vara;
            }
            scale_va = 0.0;
        } else {
            vara = scale_va * P.df_va / (P.df_va + 2.0);
        }
        sc.vara = vara;
        sc.scale_va = scale_va;
        for (int k = 0; k < nmix; k++) sc.gp[k] = P.mix[k] * vara;
        double yty = 0.0;
        for (int i = 0; i < n; i++) {
            double r = (y[i] - 0.0) - sc.mu;
            yty += r * r;
        }
        sc.rtr = yty;   // k_init_draws' input
    }
    if (hip_check(hipMemcpyAsync(h->scal, scs.data(), sizeof(Scal) * Cs, hipMemcpyHostToDevice, h->stream),
                  "init scal"))
        return SB_ERR_HIP;
    hipMemcpyAsync(h->why, ys.data(), sizeof(double) * ys.size(), hipMemcpyHostToDevice, h->stream);
    h->use_heldout = heldout != nullptr;
    if (heldout) {
        if (!h->heldout && dev_alloc(h, (void **)&h->heldout, held.size())) return SB_ERR_NOMEM;
        hipMemcpyAsync(h->heldout, held.data(), held.size(), hipMemcpyHostToDevice, h->stream);
    }
    hipMemsetAsync(h->r, 0, sizeof(double) * (size_t)P.ldr * Cs, h->stream);
    hipMemsetAsync(h->ra, 0, sizeof(double) * P.blocksize * Cs, h->stream);
    hipMemsetAsync(h->beta, 0, sizeof(double) * p * Cs, h->stream);
    hipMemsetAsync(h->gstore, 0, sizeof(double) * p * Cs, h->stream);
    hipMemsetAsync(h->vara_sstore, 0, sizeof(double) * p * Cs, h->stream);
    hipMemsetAsync(h->indiststore, 0, sizeof(double) * p * nmix * Cs, h->stream);
    hipMemsetAsync(h->d_it + 2, 0, sizeof(int), h->stream);   // samples in the posterior sums
    std::vector<int32_t> two(p * Cs, 2);
    hipMemcpyAsync(h->cls, two.data(), sizeof(int32_t) * p * Cs, hipMemcpyHostToDevice, h->stream);
    std::vector<double> vs(p * Cs);
    for (int c = 0; c < C; c++) std::fill(vs.begin() + (size_t)c * p, vs.begin() + (size_t)(c + 1) * p, scs[c].vara);
    hipMemcpyAsync(h->vara_s, vs.data(), sizeof(double) * p * Cs, hipMemcpyHostToDevice, h->stream);
    k_init_r<<<dim3((n + NT - 1) / NT, 1, C), NT, 0, h->stream>>>(P, h->why, h->scal, h->r);
    if (int rc = init_xpx(h)) return rc;
    if (int rc = batch_clear_carry(h)) return rc;   // r is exact: nothing carried
    h->rstat_nwg = 0;
    k_init_draws<<<1, 64, 0, h->stream>>>(P, h->scal);
    if (P.naind > 0) {   // no augmented pseudo-data with -SODAOff
        if (P.rng_mode == SB_RNG_REPLAY)
            k_ra_replay<<<1, 1, 0, h->stream>>>(P, h->scal, h->ra);
        else
            k_ra_counter<<<dim3((P.naind + NT - 1) / NT, 1, C), NT, 0, h->stream>>>(P, h->scal, h->ra, nullptr,
                                                                                   0x7FFFFFFF, TAG_INIT);
    }
    if (hip_check(hipGetLastError(), "init kernels")) return SB_ERR_HIP;
    if (hip_check(hipStreamSynchronize(h->stream), "init")) return SB_ERR_HIP;
    h->initialised = true;
    h->graph_ok = false;   // a captured sweep may name another state (held-out rows)
    h->iterations = 0;
    return SB_OK;
}

// z'z of the SNPs the pivoted-Cholesky fallback skipped (models.f90:82-85)
int init_xpx(sb_handle *h)
{
    if (h->nskipped > 0 && !h->par.regular)   // the single-site sweep has its windows' Grams
        k_xpx<<<h->nskipped, NT, 0, h->stream>>>(h->par, h->geno, h->zlut, h->d_skip_list, h->xpx);
    return hip_check(hipGetLastError(), "xpx");
}

static void fp64_products(sb_handle *h, int b, int c);

int block_rhs(sb_handle *h, int b, double *out_host)
{
    if (!h->initialised) return fail(SB_ERR_STATE, "sb_block_rhs before sb_mcmc_init");
    if (b < 0 || b >= h->par.nblocks) return fail(SB_ERR_ARG, "block out of range");
    const Par &P = h->par;
    hipStream_t st = h->stream;
    const int p0 = h->ls[b], nb = h->ls[b + 1] - p0;
    double *dout = nullptr;
    if (hip_check(hipMalloc(&dout, sizeof(double) * (size_t)nb * P.nchains), "hipMalloc(rhs)")) return SB_ERR_NOMEM;
    int rc = SB_OK;
    if (P.precision == SB_PREC_MIXED && !dense_block(h, b)) {
        h->rstat_nwg = 0;
        rc = batch_block(h, b, 0.0, 0);
        if (!rc) rc = batch_draw(h, b, 0.0, h->sc.delta, reinterpret_cast<uint8_t *>(h->sc.flag), dout);
    } else {
        // FP64 products, chain by chain (the FP64 path, and the mixed path's
        // pivoted-Cholesky blocks)
        for (int c = 0; c < P.nchains; c++) {
            fp64_products(h, b, c);
            k_rhs_out<<<(nb + NT - 1) / NT, NT, 0, st>>>(P, h->sc.part, h->sc.partA, h->sc.partR, h->f2, h->sdv,
                                                         p0, nb, dout + (int64_t)c * nb);
        }
    }
    if (!rc) rc = hip_check(hipMemcpyAsync(out_host, dout, sizeof(double) * (size_t)nb * P.nchains,
                                           hipMemcpyDeviceToHost, st),
                            "download rhs");
    if (!rc) rc = hip_check(hipStreamSynchronize(st), "block rhs");
    hipFree(dout);
    return rc;
}

// The FP64 block products X_s'r + U_s'r_a of chain c into the per-block
// scratch (part, partA, partR).
static void fp64_products(sb_handle *h, int b, int c)
{
    const Par &P = h->par;
    KClock &kc = h->kc;
    hipStream_t st = h->stream;
    const int p0 = h->ls[b], nb = h->ls[b + 1] - p0;
    int spw = 8 * (int)std::max<int64_t>(1, std::min<int64_t>(4, (int64_t)nb * P.nchunk / 16384));
    int ngs = (nb + 4 * spw - 1) / (4 * spw);
    int nxg = ngs * P.nchunk;
    int nag = (nb + 3) / 4;
    const double gbytes = (double)nb * (((double)P.n + 3.0) / 4.0);   // 2-bit genotypes of the block
    const double ubytes = 8.0 * (double)nb * (nb + 1) / 2.0;           // packed upper factor
    kc.begin(st);
    kc.last(KS_RHS, gbytes + ubytes + 8.0 * ((double)P.n + 2.0 * nb));
    launch(kc, k_rhs, dim3(nxg + nag), dim3(NT), 0, st, P, h->geno, h->r + (int64_t)c * P.ldr,
           h->ra + (int64_t)c * P.blocksize, h->U, h->tri_off[b], dense_block(h, b), p0, nb, spw, ngs, nxg,
           h->sc.part, h->sc.partA, h->sc.partR);
}

// One block on the FP64 kernels for every chain: products and draws chain by
// chain, then the ordered changed lists, r and r_a updates of all chains.
// The FP64 path's only sweep; in a mixed run, the blocks that took the
// pivoted-Cholesky fallback (dense factor, skipped SNPs), after which the
// next block recomputes its r limbs (its X_s'Q launch also requantises r_a).
static int fp64_block(sb_handle *h, int b, const std::vector<double> &xxd)
{
    const Par &P = h->par;
    KClock &kc = h->kc;
    hipStream_t st = h->stream;
    const int C = P.nchains;
    const int p0 = h->ls[b], nb = h->ls[b + 1] - p0;
    const int64_t tri0 = h->tri_off[b];
    uint8_t *chg = reinterpret_cast<uint8_t *>(h->sc.flag);
    double *delta = h->sc.delta;
    for (int c = 0; c < C; c++) {
        fp64_products(h, b, c);
        const int64_t o = (int64_t)c * P.p;
        kc.begin(st);
        kc.last(KS_DRAW, 8.0 * nb * (P.nchunk + 8.0));
        if (P.rng_mode == SB_RNG_REPLAY)
            launch(kc, k_draw_replay, dim3(1), dim3(NT), 0, st, P, h->scal, h->sc.part, h->sc.partA, h->sc.partR,
                   h->f2, h->sdv, h->grp, p0, nb, xxd[b], h->sc.rhs, h->sc.stemp, h->sc.uu, h->sc.chg, h->beta,
                   h->cls, h->vara_s, delta, chg, h->skip);
        else
            with_nmix(P.nmix, [&](auto nm) {
                launch(kc, k_draw_counter<decltype(nm)::value>, dim3((nb + NT - 1) / NT), dim3(NT), 0, st, P,
                       h->scal + c, h->sc.part, h->sc.partA, h->sc.partR, h->f2, h->sdv, h->grp,
                       h->d_pos_by_id + P.p, p0, nb, xxd[b], h->d_it, c, h->beta + o, h->cls + o, h->vara_s + o,
                       delta + o, chg + o, h->skip);
            });
    }
    const int nrw = (P.n + 255) / 256;
    const int nra = ((nb + WAVE - 1) / WAVE) * ((nb + RCH - 1) / RCH);
    kc.begin(st);
    kc.last(KS_LIST, 25.0 * nb);
    launch(kc, k_list, dim3(1, 1, C), dim3(NT), 0, st, P, chg, delta, h->f2, h->sdv, p0, nb, h->u_lst, h->u_dl,
           h->u_cl, h->u_cnt, h->u_kc, h->u_tstart, dense_block(h, b) != nullptr ? 1 : 0, kc.extra(KS_UPDATE));
    kc.begin(st);
    launch(kc, k_update, dim3(nrw + nra, 1, C), dim3(UW * WAVE), 0, st, P, h->geno, h->U, tri0, dense_block(h, b),
           p0, nb, nrw, h->u_lst, h->u_dl, h->u_cl, h->u_cnt, h->u_kc, h->u_tstart, h->r, h->u_rapart);
    kc.last(KS_UPDATE, 16.0 * P.n + 16.0 * nb);
    launch(kc, k_ra_sum, dim3((nb + NT - 1) / NT, 1, C), dim3(NT), 0, st, P, h->u_rapart, h->u_cnt, h->u_tstart,
           nb, h->ra);
    if (P.precision == SB_PREC_MIXED) {
        h->rstat_nwg = 0;   // r moved outside the batched update: fresh statistics and limbs
        h->q_block = -1;
    }
    return hip_check(hipGetLastError(), "FP64 block");
}

// One sweep iteration (models.f90:129-570, without the resync): mu, the
// block loop, the skipped SNPs or the single-site sweep, the
// hyperparameters, fresh r_a, held-out imputation and the posterior sums.
// Iteration-dependent values come from the device state h->d_it, which the
// first kernel advances, so the same launches serve any iteration (and can
// be captured as a graph).  `rep` is used only by the paths that are never
// captured (the single-site sweep and the skipped SNPs).
static int sweep_iteration(sb_handle *h, int rep, const std::vector<double> &xxd)
{
    const Par &P = h->par;
    hipStream_t st = h->stream;
    const int nred = std::min(1024, (P.n + NT - 1) / NT);
    const int C = P.nchains;
    const bool mixed = P.precision == SB_PREC_MIXED;
    const int bin_chunk = 8192;
    const int nwg_bins = (P.p + bin_chunk - 1) / bin_chunk;
    double *bcnt = h->sc.red + (int64_t)1024 * C, *bvar = bcnt + (int64_t)C * P.nbins * nwg_bins;
    uint8_t *chg = reinterpret_cast<uint8_t *>(h->sc.flag);
    double *delta = h->sc.delta;
    KClock &kc = h->kc;
    int rc = SB_OK;
    k_next_it<<<1, 1, 0, st>>>(h->d_it);
    k_mu_pre<<<dim3(nred, 1, C), NT, 0, st>>>(P, h->r, h->scal, h->sc.red);
    k_mu_draw<<<C, 64, 0, st>>>(P, h->scal, h->sc.red, nred, h->d_it);
    k_mu_post<<<dim3(nred, 1, C), NT, 0, st>>>(P, h->r, h->scal);
    h->rstat_nwg = 0;   // r changed: the batched path recomputes its statistics
    for (int b = 0; b < P.nblocks && !rc; b++) {
        if (mixed && !dense_block(h, b)) {
            rc = batch_block(h, b, xxd[b], rep);
            if (!rc) rc = batch_draw(h, b, xxd[b], delta, chg);
            if (!rc) rc = batch_update(h, b);
        } else {
            rc = fp64_block(h, b, xxd);
        }
    }
    if (rc) return rc;
    if (P.regular) {
        // -SODAOff: the single-site sweep over every SNP
        rc = single_sweep(h, rep);
        if (rc) return rc;
    } else if (h->nskipped > 0) {
        // the SNPs the pivoted-Cholesky fallback skipped, one by one, every
        // chain (models.f90:433-495)
        with_nmix(P.nmix, [&](auto nm) {
            k_skipped<decltype(nm)::value><<<C, SKT, 0, st>>>(P, h->scal, h->geno, h->zlut, h->d_skip_list,
                                                              h->nskipped, h->xpx, h->grp, h->d_pos_by_id + P.p,
                                                              rep, h->r, h->beta, h->cls, h->vara_s, delta, chg);
        });
    }
    k_rtr<<<dim3(nred, 1, C), NT, 0, st>>>(P, h->r, h->sc.red);
    // one-pass bins also add the posterior sums (k_accum otherwise, below)
    const bool bins1 = P.nbins <= 16;
    if (P.nbins <= 4)
        k_bins_all<4><<<dim3(nwg_bins, 1, C), NT, 0, st>>>(P, h->beta, h->cls, h->grp, bin_chunk, bcnt, bvar,
                                                           h->d_it, h->vara_s, h->gstore, h->indiststore,
                                                           h->vara_sstore);
    else if (P.nbins <= 16)
        k_bins_all<16><<<dim3(nwg_bins, 1, C), NT, 0, st>>>(P, h->beta, h->cls, h->grp, bin_chunk, bcnt, bvar,
                                                            h->d_it, h->vara_s, h->gstore, h->indiststore,
                                                            h->vara_sstore);
    else
        k_bins<<<dim3(nwg_bins, P.nbins, C), NT, 0, st>>>(P, h->beta, h->cls, h->grp, bin_chunk, bcnt, bvar);
    k_hyper<<<C, 64, 0, st>>>(P, h->scal, h->sc.red, nred, bcnt, bvar, nwg_bins, h->d_it, h->hyp);
    if (P.naind > 0) {
        if (P.rng_mode == SB_RNG_REPLAY)
            k_ra_replay<<<1, 1, 0, st>>>(P, h->scal, h->ra);
        else
            k_ra_counter<<<dim3((P.naind + NT - 1) / NT, 1, C), NT, 0, st>>>(P, h->scal, h->ra, h->d_it, 0, TAG_RA);
    }
    if (h->use_heldout) {
        k_impute<<<dim3((P.n + NT - 1) / NT, 1, C), NT, 0, st>>>(P, h->scal, h->heldout, h->why, h->r, h->d_it);
        h->rstat_nwg = 0;
    }
    if (!bins1)
        k_accum<<<dim3((P.p + NT - 1) / NT, 1, C), NT, 0, st>>>(P, h->d_it, h->beta, h->cls, h->vara_s,
                                                               h->gstore, h->indiststore, h->vara_sstore);
    return hip_check(hipGetLastError(), "sweep launch");
}

// compute_residuals every P.resync iterations (models.f90:571-573)
static int sweep_resync(sb_handle *h)
{
    const Par &P = h->par;
    hipStream_t st = h->stream;
    const int C = P.nchains;
    k_nonzero<<<dim3(1, 1, C), NT, 0, st>>>(P, h->d_pos_by_id, h->beta, h->resync_list, h->d_count);
    k_resync<<<dim3((P.ldw + NT - 1) / NT, 1, C), NT, 0, st>>>(P, h->geno, h->zlut, h->beta, h->resync_list,
                                                              h->d_count, h->why, h->scal, h->r);
    h->rstat_nwg = 0;
    int rc = batch_clear_carry(h);   // r is exact again
    if (!rc) rc = hip_check(hipGetLastError(), "resync launch");
    return rc;
}

// HIP graphs for the sweep, opt-in with SODABAYES_GRAPH=1: replaying one
// captured iteration measured 11.18 against 11.20 ms per iteration (mixed,
// 16 chains) and 13.00 against 13.08 ms (FP64) at 50,000 x 500,000 with
// subsets of 4,096 -- the host already enqueues far ahead of the device, and
// a graph's kernel nodes are still dispatched one by one -- so eager
// launches stay the default.  Never for the per-kernel timing, the
// single-site sampler or the skipped-SNP updates (their launches depend on
// the iteration on the host).
static bool graph_wanted(const sb_handle *h)
{
    const char *e = std::getenv("SODABAYES_GRAPH");
    if (!e || e[0] != '1') return false;
    return !h->kc.on && !h->par.regular && h->nskipped == 0 && h->par.nblocks > 0;
}

// class 1's posterior counts into indiststore (before it is read out)
int class1_counts(sb_handle *h)
{
    const Par &P = h->par;
    if (P.method == SB_BAYES_A) return SB_OK;
    k_class1<<<dim3((P.p + NT - 1) / NT, 1, P.nchains), NT, 0, h->stream>>>(P, h->d_it, h->indiststore);
    return hip_check(hipGetLastError(), "class-1 counts");
}

// the sample count of the posterior sums after a state load
int samples_from_counts(sb_handle *h)
{
    k_samples_from_counts<<<1, 1, 0, h->stream>>>(h->par, h->indiststore, h->d_it);
    return hip_check(hipGetLastError(), "sample count");
}

int mcmc_run(sb_handle *h, int first, int count, double *hyp_host)
{
    if (!h->initialised) return fail(SB_ERR_STATE, "sb_mcmc_run before sb_mcmc_init");
    if (count < 0 || first < 1) return fail(SB_ERR_ARG, "bad iteration range");
    const Par &P = h->par;
    hipStream_t st = h->stream;
    const int hw = 5 + 2 * P.nbins;
    int nrec = 0;
    for (int rep = first; rep < first + count; rep++) nrec += (rep % P.thin) == 0;
    const bool mixed = P.precision == SB_PREC_MIXED;
    std::vector<double> xxd(P.nblocks);
    if (P.nblocks > 0) hipMemcpy(xxd.data(), h->xxd, sizeof(double) * P.nblocks, hipMemcpyDeviceToHost);
    int rc = SB_OK;
    if (mixed && !h->tuned) rc = batch_autotune(h);
    if (rc) return rc;
    // the call's .hyp records go to a buffer of the handle (a larger call
    // replaces it, and with it any captured graph that names it)
    if (nrec > h->hyp_cap) {
        if (h->gexec) {
            hipGraphExecDestroy(h->gexec);
            h->gexec = nullptr;
        }
        h->graph_ok = false;
        if (h->hyp) hipFree(h->hyp);
        h->hyp = nullptr;
        h->hyp_cap = 0;
        const int cap = std::max(nrec, 1024);
        if (hip_check(hipMalloc(&h->hyp, sizeof(double) * (size_t)cap * hw), "hipMalloc(hyp records)"))
            return SB_ERR_NOMEM;
        h->hyp_cap = cap;
    }
    // optional per-kernel timing on this stream
    const char *kt = std::getenv("SODABAYES_KERNEL_TIMING");
    h->kc.on = kt && kt[0] == '1';
    KClock &kc = h->kc;
    const bool use_graph = graph_wanted(h);
    auto t0 = std::chrono::steady_clock::now();
    k_set_it<<<1, 1, 0, st>>>(h->d_it, first - 1, 0);
    // every failure below leaves the loop with rc set and goes through the
    // one exit at the end (stream drained, clock reset)
    for (int rep = first; rep < first + count && !rc; rep++) {
        if (use_graph && !h->graph_ok) {
            // capture one iteration; its launches run when the graph does
            if (h->gexec) {
                hipGraphExecDestroy(h->gexec);
                h->gexec = nullptr;
            }
            hipGraph_t g = nullptr;
            rc = hip_check(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal), "begin capture");
            if (rc) break;
            const int rb = sweep_iteration(h, rep, xxd);
            const int re = hip_check(hipStreamEndCapture(st, &g), "end capture");
            rc = rb ? rb : re;
            if (!rc) rc = hip_check(hipGraphInstantiate(&h->gexec, g, nullptr, nullptr, 0), "graph instantiate");
            if (g) hipGraphDestroy(g);
            if (rc) break;
            h->graph_ok = true;
        }
        if (use_graph)
            rc = hip_check(hipGraphLaunch(h->gexec, st), "graph launch");
        else
            rc = sweep_iteration(h, rep, xxd);
        if (!rc && rep % P.resync == 0) rc = sweep_resync(h);
    }
    const int rs = hip_check(hipStreamSynchronize(st), "sweep");
    if (!rc) rc = rs;
    auto t1 = std::chrono::steady_clock::now();
    if (kc.on) kc.collect();   // also clears the records of a failed run
    kc.on = false;
    if (count > 0) h->last_sweep_ms = std::chrono::duration<double, std::milli>(t1 - t0).count() / count;
    h->iterations += count;
    if (!rc && hyp_host && nrec > 0)
        rc = hip_check(hipMemcpy(hyp_host, h->hyp, sizeof(double) * (size_t)nrec * hw, hipMemcpyDeviceToHost),
                       "download hyp");
    return rc;
}

}  // namespace sb

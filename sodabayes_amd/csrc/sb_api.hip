// C ABI of the SODA sampler (include/sodabayes.h): handle lifetime, host-side
// reference RNG helpers, argument checks and state download.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "sb_internal.h"

namespace {
thread_local std::string g_err;

constexpr double DEPS = 2.220446049250313080847e-16;  // epsilon(1.0d0)

// xoroshiro128+ of parallelrng.f90:100-149, host copy for the caller's
// setup draws (seeding, block permutation).
struct HostXoro {
    uint64_t s1, s2;
    static uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
    uint64_t next()
    {
        uint64_t a = s1, b = s2, r = a + b;
        b ^= a;
        s1 = rotl(a, 55) ^ b ^ (b << 14);
        s2 = rotl(b, 36);
        return r;
    }
    double uniform()
    {
        uint64_t b = 0x3FF0000000000000ULL | (next() >> 12);
        double d;
        std::memcpy(&d, &b, sizeof d);
        return d - 1.0;
    }
    void jump()
    {
        static const uint64_t J[2] = {0xbeac0467eba5facbULL, 0xd86b048b86aa9922ULL};
        uint64_t a = 0, b = 0;
        for (int i = 0; i < 2; i++)
            for (int bit = 0; bit < 64; bit++) {
                if (J[i] & (1ULL << bit)) {
                    a ^= s1;
                    b ^= s2;
                }
                next();
            }
        s1 = a;
        s2 = b;
    }
};
}  // namespace

namespace sb {

int fail(int code, const std::string &msg)
{
    g_err = msg;
    return code;
}

int hip_check(hipError_t e, const char *what)
{
    if (e == hipSuccess) return SB_OK;
    g_err = std::string(what) + ": " + hipGetErrorString(e);
    return SB_ERR_HIP;
}

int dev_alloc(sb_handle *h, void **p, size_t bytes)
{
    if (bytes == 0) bytes = 8;
    hipError_t e = hipMalloc(p, bytes);
    if (e != hipSuccess) {
        g_err = "hipMalloc(" + std::to_string(bytes) + " B): " + hipGetErrorString(e);
        *p = nullptr;
        return SB_ERR_NOMEM;
    }
    h->allocs.push_back(*p);
    h->bytes += bytes;
    return SB_OK;
}

// ---- device-state checkpoints ----
struct StateHeader {
    char magic[8];             // "SBSTATE1"
    int64_t bytes;             // whole checkpoint
    sb_dims dims;
    int32_t naind, nskipped, use_heldout, nsec;
    int64_t iterations;
    int64_t sec[16];           // bytes of each device section, in state_sections order
};
constexpr char STATE_MAGIC[8] = {'S', 'B', 'S', 'T', 'A', 'T', 'E', '1'};

struct Section {
    void *dev;
    size_t bytes;
};

// Device sections of a handle's chain state, in checkpoint order.
std::vector<Section> state_sections(sb_handle *h, bool heldout)
{
    const Par &P = h->par;
    const size_t C = (size_t)P.nchains, p = (size_t)P.p, ldr = (size_t)P.ldr, bs = (size_t)P.blocksize;
    std::vector<Section> v = {
        {h->scal, sizeof(Scal) * C},     {h->r, sizeof(double) * ldr * C},
        {h->ra, sizeof(double) * bs * C},    {h->why, sizeof(double) * ldr * C},
        {h->beta, sizeof(double) * p * C},   {h->cls, sizeof(int32_t) * p * C},
        {h->vara_s, sizeof(double) * p * C}, {h->gstore, sizeof(double) * p * C},
        {h->vara_sstore, sizeof(double) * p * C},
        {h->indiststore, sizeof(double) * p * (size_t)P.nmix * C},
    };
    if (heldout) v.push_back({h->heldout, ldr * C});
    // the mixed path's carried quantisation remainders: r is exact only with them
    if (h->b_carry) v.push_back({h->b_carry, sizeof(double) * p * C});
    return v;
}

int64_t state_bytes(sb_handle *h, bool heldout)
{
    int64_t b = sizeof(StateHeader) + sizeof(Par) + sizeof(double) * h->par.nblocks;
    for (const Section &s : state_sections(h, heldout)) b += (int64_t)s.bytes;
    return b;
}

}  // namespace sb

using namespace sb;

extern "C" {

const char *sb_last_error(void) { return g_err.c_str(); }

int sb_rng_seed(int64_t seed, int32_t stream, uint64_t state[2])
{
    if (!state || stream < 0) return fail(SB_ERR_ARG, "sb_rng_seed: bad argument");
    uint64_t a = (uint64_t)(seed < 0 ? -seed : seed);
    HostXoro s{a, a + 1};
    for (int k = 0; k <= stream; k++) s.jump();
    state[0] = s.s1;
    state[1] = s.s2;
    return SB_OK;
}

int sb_permutate(int32_t n, int32_t s, uint64_t state[2], int32_t *p)
{
    if (!state || !p || n <= 0 || s <= 0) return fail(SB_ERR_ARG, "sb_permutate: bad argument");
    HostXoro x{state[0], state[1]};
    for (int i = 0; i < n; i++) p[i] = i;
    std::vector<double> u(s);
    for (int i = 1; i <= n; i += s) {
        int m = (n - i + 1) < s ? (n - i + 1) : s;
        for (int j = 0; j < s; j++) u[j] = x.uniform();
        for (int j = 1; j <= m; j++) {
            int ipj = i + j - 1;
            int k = (int)(u[j - 1] * (n - ipj + 1)) + ipj;
            int t = p[ipj - 1];
            p[ipj - 1] = p[k - 1];
            p[k - 1] = t;
        }
    }
    state[0] = x.s1;
    state[1] = x.s2;
    return SB_OK;
}

int sb_create(const sb_dims *dims, const uint8_t *bed, const double *freq, const int32_t *permvec,
              const int32_t *snpgrp, sb_handle **out)
{
    if (!dims || !bed || !freq || !permvec || !out) return fail(SB_ERR_ARG, "sb_create: null argument");
    *out = nullptr;
    const sb_dims &d = *dims;
    if (d.method < SB_BAYES_A || d.method > SB_BAYES_RC) return fail(SB_ERR_ARG, "unknown method");
    if (d.nind <= 0 || d.nloci <= 0) return fail(SB_ERR_ARG, "nind and nloci must be positive");
    int want_nmix = d.method == SB_BAYES_A ? 1 : (d.method <= SB_BAYES_CPI ? 2 : d.nmix);
    if (d.nmix != want_nmix || d.nmix < 1 || d.nmix > MAXMIX) return fail(SB_ERR_ARG, "bad nmix for method");
    int want_ngroup = d.method == SB_BAYES_RC ? d.ngroup : 1;
    if (d.ngroup != want_ngroup || d.ngroup < 1 || d.ngroup > MAXGRP) return fail(SB_ERR_ARG, "bad ngroup");
    if (d.method == SB_BAYES_RC && !snpgrp) return fail(SB_ERR_ARG, "BayesRc needs snpgrp");
    // nblocks = blocksize = 0: -SODAOff, the reference's single-site sampler
    const bool regular = d.nblocks == 0 && d.blocksize == 0;
    if (!regular) {
        if (d.blocksize < 2 || d.blocksize > 4096) return fail(SB_ERR_ARG, "blocksize must be in [2, 4096]");
        if (d.nblocks < 1 || d.blockt < 0 || d.blockt > d.nblocks) return fail(SB_ERR_ARG, "bad block partition");
    }
    if (d.rng_mode != SB_RNG_REPLAY && d.rng_mode != SB_RNG_COUNTER) return fail(SB_ERR_ARG, "bad rng_mode");
    if (d.precision != SB_PREC_FP64 && d.precision != SB_PREC_MIXED) return fail(SB_ERR_ARG, "bad precision");
    if (d.nchains < 1 || d.nchains > 64) return fail(SB_ERR_ARG, "nchains must be in [1, 64]");
    if (regular && d.precision != SB_PREC_FP64)
        return fail(SB_ERR_ARG, "the single-site sampler (-SODAOff) runs in FP64 (no block products)");
    if (!regular && d.nchains > 1 && d.precision != SB_PREC_MIXED)
        return fail(SB_ERR_ARG, "several SODA chains per handle need the mixed-precision batched path");
    if (d.rng_mode == SB_RNG_REPLAY && (d.nchains != 1 || d.precision != SB_PREC_FP64))
        return fail(SB_ERR_ARG, "replay runs one FP64 chain (the reference's -nthreads 1 stream)");

    // block bounds, soda.f90:463-469
    std::vector<int> ls(d.nblocks + 1);
    int pos = 0;
    for (int b = 0; b < d.nblocks; b++) {
        int s0, sz;
        if (b < d.blockt) {
            s0 = b * d.blocksize;
            sz = std::min(d.blocksize, d.nloci - s0);
        } else {
            s0 = d.blocksize * d.blockt + (b - d.blockt) * (d.blocksize - 1);
            sz = std::min(d.blocksize - 1, d.nloci - s0);
        }
        if (s0 != pos || sz < 1) return fail(SB_ERR_ARG, "block partition does not tile the SNPs");
        ls[b] = s0;
        pos = s0 + sz;
    }
    ls[d.nblocks] = pos;
    if (!regular && pos != d.nloci) return fail(SB_ERR_ARG, "block partition does not cover nloci");

    sb_handle *h = new sb_handle();
    h->d = d;
    hipGetDevice(&h->device);
    if (hip_check(hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking), "stream")) {
        delete h;
        return SB_ERR_HIP;
    }
    Par &P = h->par;
    std::memset(&P, 0, sizeof P);
    P.method = d.method;
    P.n = d.nind;
    P.p = d.nloci;
    P.nmix = d.nmix;
    P.ngroup = d.ngroup;
    P.nbins = d.nmix * d.ngroup;
    P.nblocks = d.nblocks;
    P.blocksize = d.blocksize;
    P.blockt = d.blockt;
    P.naind = d.blocksize;  // naind - naless with every block positive definite
    P.regular = regular ? 1 : 0;
    int words = (d.nind + 15) / 16;
    P.ldw = ((words + CHUNK_WORDS - 1) / CHUNK_WORDS) * CHUNK_WORDS;
    P.nchunk = P.ldw / CHUNK_WORDS;
    P.ldr = (int64_t)P.ldw * 16;
    P.rng_mode = d.rng_mode;
    P.nchains = d.nchains;
    P.precision = d.precision;
    h->ls = ls;
    h->tri_off.assign(d.nblocks + 1, 0);
    for (int b = 0; b < d.nblocks; b++) {
        int64_t nb = ls[b + 1] - ls[b];
        h->tri_off[b + 1] = h->tri_off[b] + nb * (nb + 1) / 2;
    }
    h->permvec.assign(permvec, permvec + d.nloci);
    h->pos_of.assign(d.nloci, -1);
    for (int k = 0; k < d.nloci; k++) {
        int s = permvec[k];
        if (s < 0 || s >= d.nloci || h->pos_of[s] >= 0) {
            delete h;
            return fail(SB_ERR_ARG, "permvec is not a permutation of 0..nloci-1");
        }
        h->pos_of[s] = k;
    }

    const int64_t p = d.nloci, bs = d.blocksize, ldr = (int64_t)P.ldw * 16, C = d.nchains;
    const int bin_chunk = 8192, nwg_bins = (int)((p + bin_chunk - 1) / bin_chunk);
    int rc = 0;
    auto A = [&](void **ptr, size_t bytes) {
        if (!rc) rc = dev_alloc(h, ptr, bytes);
    };
    A((void **)&h->geno, sizeof(uint32_t) * (size_t)(p * P.ldw));
    A((void **)&h->f2, sizeof(double) * 2 * p);
    A((void **)&h->sdv, sizeof(double) * p);
    A((void **)&h->zlut, sizeof(double) * 4 * p);
    A((void **)&h->grp, sizeof(int32_t) * p);
    A((void **)&h->d_pos_by_id, sizeof(int32_t) * 2 * p);
    A((void **)&h->xxd, sizeof(double) * d.nblocks);
    A((void **)&h->U, sizeof(double) * (size_t)h->tri_off[d.nblocks]);
    A((void **)&h->skip, (size_t)p + 64);
    h->dense_off.assign(d.nblocks, -1);
    A((void **)&h->d_tri_off, sizeof(int64_t) * (d.nblocks + 1));
    A((void **)&h->d_ls, sizeof(int32_t) * (d.nblocks + 1));
    A((void **)&h->r, sizeof(double) * ldr * C);
    A((void **)&h->ra, sizeof(double) * bs * C);
    A((void **)&h->why, sizeof(double) * ldr * C);
    A((void **)&h->beta, sizeof(double) * p * C);
    A((void **)&h->cls, sizeof(int32_t) * p * C);
    A((void **)&h->vara_s, sizeof(double) * p * C);
    A((void **)&h->gstore, sizeof(double) * p * C);
    A((void **)&h->vara_sstore, sizeof(double) * p * C);
    A((void **)&h->indiststore, sizeof(double) * p * d.nmix * C);
    A((void **)&h->scal, sizeof(Scal) * C);
    A((void **)&h->sc.part, sizeof(double) * P.nchunk * bs);
    A((void **)&h->sc.partA, sizeof(double) * bs);
    A((void **)&h->sc.partR, sizeof(double) * P.nchunk);
    A((void **)&h->sc.rhs, sizeof(double) * bs);
    A((void **)&h->sc.stemp, sizeof(double) * bs * d.nmix);
    A((void **)&h->sc.uu, sizeof(double) * 2 * bs);
    A((void **)&h->sc.chg, sizeof(int32_t) * bs);
    A((void **)&h->sc.flag, sizeof(uint8_t) * p * C + 64);
    A((void **)&h->sc.delta, sizeof(double) * p * C);
    A((void **)&h->sc.red, sizeof(double) * C * (1024 + 2 * (size_t)P.nbins * nwg_bins) + 64);
    A((void **)&h->resync_list, sizeof(int32_t) * p * C);
    A((void **)&h->u_lst, sizeof(int) * bs * C);
    A((void **)&h->u_dl, sizeof(double) * bs * C);
    A((void **)&h->u_cl, sizeof(double) * bs * C);
    A((void **)&h->u_kc, sizeof(double) * (C + 4));
    A((void **)&h->d_it, sizeof(int) * 4);
    A((void **)&h->u_cnt, sizeof(int) * (C + 4));
    A((void **)&h->u_rapart, sizeof(double) * bs * C * ((bs + 255) / 256 + 1));
    A((void **)&h->u_tstart, sizeof(int) * C * (bs / 64 + 2));
    A((void **)&h->d_count, sizeof(int32_t) * (C + 4));
    if (d.precision == SB_PREC_MIXED) {
        h->nwg_r = (int)std::min<int64_t>(64, (d.nind + 255) / 256);
        const int ntc = (int)((C + 15) / 16);
        h->b_ntile = h->limbs * ntc;   // limb-major rows: limb k of chain c at row k * 16 ntc + c
        {
            int dev = 0, ncu = 0;
            hipGetDevice(&dev);
            if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || ncu < 1)
                ncu = 1;
            h->ncu = ncu;
        }
        h->b_ldq = ldr;
        const int nS = P.ldw / 16;
        h->b_rw = h->b_ntile <= 6 ? 4 : 2;
        h->b_nch = sb::prod_chunks(h, h->b_rw, 4);
        if (const char *e = std::getenv("SODABAYES_ROWS_PER_WAVE")) {
            h->b_rw = (std::atoi(e) == 4 && h->b_ntile <= 6) ? 4 : 2;
            h->b_nch = sb::prod_chunks(h, h->b_rw, 4);
            h->tuned = true;
        }
        if (const char *e = std::getenv("SODABAYES_KSS")) {
            const int k = std::max(1, std::atoi(e));
            h->b_nch = std::max(1, (nS + k - 1) / k);
            h->tuned = true;
        }
        h->kss = (nS + h->b_nch - 1) / h->b_nch;
        // the chunk partials are sized for the most chunks the tuner tries
        h->uq_group = sb::batch_uq_group(C);   // before the chunk candidates, which count the U pairs
        h->rows_per_group = 64 * h->uq_group;
        const int nch = sb::batch_max_chunks(h);
        const int nwg_d = (int)((bs + 255) / 256);
        int nrg = (int)((bs + h->rows_per_group - 1) / h->rows_per_group);
        const int64_t nwg_stat = std::max<int64_t>(h->nwg_r, sb::RSLOT);  // k_rstat partials or update slots
        h->b_nq = (int)(P.ldw / 16);   // quantisation workgroups per chain tile: one per super-step
        A((void **)&h->b_rmax, sizeof(uint64_t) * C * nwg_stat);
        A((void **)&h->b_qsum, sizeof(long long) * C * h->b_nq);
        A((void **)&h->b_Q, (size_t)16 * h->b_ntile * h->b_ldq);
        A((void **)&h->b_expo, sizeof(int) * C);
        A((void **)&h->b_T, sizeof(long long) * (size_t)nch * bs * C);
        A((void **)&h->b_rapart, sizeof(double) * (size_t)nrg * bs * C);
        A((void **)&h->b_uapart, sizeof(double) * (size_t)nrg * bs * C);
        h->b_ldd = (int64_t)16 * sb::batch_words((int)bs);
        A((void **)&h->b_Dq, (size_t)16 * h->b_ntile * h->b_ldd);
        A((void **)&h->b_expo2, sizeof(int) * C);
        A((void **)&h->b_dmax, sizeof(uint64_t) * C * nwg_d);
        A((void **)&h->b_kpart, sizeof(double) * C * nwg_d);
        A((void **)&h->b_cj, sizeof(double) * C * bs);
        A((void **)&h->b_carry, sizeof(double) * C * (size_t)d.nloci);
        A((void **)&h->b_ramax, sizeof(unsigned long long) * 2 * C);
        h->gt_off.assign(d.nblocks + 1, 0);
        h->gs_off.assign(d.nblocks + 1, 0);
        h->uq0_off.assign(d.nblocks + 1, 0);
        h->uq1_off.assign(d.nblocks + 1, 0);
        for (int b = 0; b < d.nblocks; b++) {
            const int64_t nbb = ls[b + 1] - ls[b], nbt = (nbb + 15) / 16;
            h->gt_off[b + 1] = h->gt_off[b] + (ldr / 16) * (sb::batch_words((int)nbb) / 16) * 64;
            h->gs_off[b + 1] = h->gs_off[b] + nbt * nS * 64;
            h->uq0_off[b + 1] = h->uq0_off[b] + sb::batch_uq_elems((int)nbb, 0);
            h->uq1_off[b + 1] = h->uq1_off[b] + sb::batch_uq_elems((int)nbb, 1);
        }
        A((void **)&h->geno_t, (size_t)16 * h->gt_off[d.nblocks]);
        A((void **)&h->geno_sw, (size_t)16 * h->gs_off[d.nblocks]);
        A((void **)&h->Uq0, (size_t)16 * h->uq0_off[d.nblocks]);
        A((void **)&h->Uq1, (size_t)16 * h->uq1_off[d.nblocks]);
        A((void **)&h->ue0, sizeof(int) * (size_t)d.nblocks * bs);
        A((void **)&h->ue1, sizeof(int) * (size_t)d.nblocks * bs);
    }
    if (rc) {
        sb_destroy(h);
        return rc;
    }
    // per-position constants, computed with the reference's expressions
    // (xscale, utilities.f90:526-533)
    std::vector<double> f2(2 * p), sdv(p), zl(4 * p);
    std::vector<int32_t> g(p, 0), pb(2 * p);
    for (int64_t k = 0; k < p; k++) {
        int s = permvec[k];
        double f = freq[s];
        f2[k] = 2.0 * f;
        f2[p + k] = f;
        if (f < DEPS || (1.0 - f) < DEPS) {
            sdv[k] = 0.0;
            zl[4 * k] = zl[4 * k + 1] = zl[4 * k + 2] = 0.0;
        } else {
            double sd = std::sqrt(2.0 * f * (1.0 - f));
            sdv[k] = sd;
            for (int x = 0; x < 3; x++) zl[4 * k + x] = ((double)x - 2.0 * f) / sd;
        }
        zl[4 * k + 3] = 0.0;
        if (snpgrp) g[k] = snpgrp[s];
        pb[p + k] = s;
    }
    for (int64_t s = 0; s < p; s++) pb[s] = h->pos_of[s];
    for (int64_t k = 0; k < p; k++)
        if (g[k] < 0 || g[k] >= d.ngroup) {
            sb_destroy(h);
            return fail(SB_ERR_ARG, "snpgrp out of range");
        }
    hipStream_t st = h->stream;
    hipMemcpyAsync(h->f2, f2.data(), sizeof(double) * 2 * p, hipMemcpyHostToDevice, st);
    hipMemcpyAsync(h->sdv, sdv.data(), sizeof(double) * p, hipMemcpyHostToDevice, st);
    hipMemcpyAsync(h->zlut, zl.data(), sizeof(double) * 4 * p, hipMemcpyHostToDevice, st);
    hipMemcpyAsync(h->grp, g.data(), sizeof(int32_t) * p, hipMemcpyHostToDevice, st);
    hipMemcpyAsync(h->d_pos_by_id, pb.data(), sizeof(int32_t) * 2 * p, hipMemcpyHostToDevice, st);
    hipMemcpyAsync(h->d_tri_off, h->tri_off.data(), sizeof(int64_t) * (d.nblocks + 1), hipMemcpyHostToDevice, st);
    hipMemcpyAsync(h->d_ls, ls.data(), sizeof(int32_t) * (d.nblocks + 1), hipMemcpyHostToDevice, st);
    hipMemsetAsync(h->scal, 0, sizeof(Scal) * C, st);
    if (h->b_Q) {
        hipMemsetAsync(h->b_Q, 0, (size_t)16 * h->b_ntile * h->b_ldq, st);
        hipMemsetAsync(h->b_Dq, 0, (size_t)16 * h->b_ntile * h->b_ldd, st);
        hipMemsetAsync(h->b_ramax, 0, sizeof(unsigned long long) * 2 * C, st);
    }
    hipMemsetAsync(h->U, 0, sizeof(double) * (size_t)h->tri_off[d.nblocks], st);
    hipMemsetAsync(h->skip, 0, (size_t)p + 64, st);
    rc = setup_pack(h, bed, d.bed_on_device != 0);
    if (!rc && h->geno_t) rc = batch_transpose(h);
    if (rc) {
        sb_destroy(h);
        return rc;
    }
    *out = h;
    return SB_OK;
}

int sb_destroy(sb_handle *h)
{
    if (!h) return SB_OK;
    if (h->stream) hipStreamSynchronize(h->stream);
    if (h->gexec) hipGraphExecDestroy(h->gexec);
    if (h->hyp) hipFree(h->hyp);
    for (void *p : h->allocs) hipFree(p);
    if (h->stream) hipStreamDestroy(h->stream);
    delete h;
    return SB_OK;
}

int sb_rng_set(sb_handle *h, const uint64_t state[2])
{
    if (!h || !state) return fail(SB_ERR_ARG, "sb_rng_set: null argument");
    if (hip_check(hipMemcpyAsync(&h->scal->rng, state, 16, hipMemcpyHostToDevice, h->stream), "rng set"))
        return SB_ERR_HIP;
    return hip_check(hipStreamSynchronize(h->stream), "rng set");
}

int sb_rng_get(sb_handle *h, uint64_t state[2])
{
    if (!h || !state) return fail(SB_ERR_ARG, "sb_rng_get: null argument");
    if (hip_check(hipMemcpyAsync(state, &h->scal->rng, 16, hipMemcpyDeviceToHost, h->stream), "rng get"))
        return SB_ERR_HIP;
    return hip_check(hipStreamSynchronize(h->stream), "rng get");
}

int sb_build_aug(sb_handle *h, double *xxd, int32_t *fail_block, int32_t *fail_col)
{
    if (!h) return fail(SB_ERR_ARG, "sb_build_aug: null handle");
    return setup_build_aug(h, xxd, fail_block, fail_col);
}

int sb_counter_uniforms(int64_t seed, int32_t chain, const uint32_t *ctr, int32_t n, int32_t per, double *out)
{
    if (!ctr || !out || n < 0 || per < 0 || chain < 0) return fail(SB_ERR_ARG, "sb_counter_uniforms: bad argument");
    if (n == 0 || per == 0) return SB_OK;
    return counter_uniforms(seed, chain, ctr, n, per, out);
}

int sb_layout_stream(sb_handle *h, double *out)
{
    if (!h || !out) return fail(SB_ERR_ARG, "sb_layout_stream: null argument");
    if (!h->aug_built) return fail(SB_ERR_STATE, "sb_layout_stream before sb_build_aug");
    return layout_stream(h, out);
}

int sb_test_shift_d(sb_handle *h, int32_t block, double amount)
{
    if (!h) return fail(SB_ERR_ARG, "sb_test_shift_d: null handle");
    if (block >= h->par.nblocks || !(amount >= 0.0))
        return fail(SB_ERR_ARG, "sb_test_shift_d: block out of range or negative amount");
    h->test_dshift_block = block < 0 ? -1 : block;
    h->test_dshift = block < 0 ? 0.0 : amount;
    return SB_OK;
}

int sb_mcmc_init(sb_handle *h, const sb_priors *pr, const double *why)
{
    if (!h || !pr || !why) return fail(SB_ERR_ARG, "sb_mcmc_init: null argument");
    return mcmc_init(h, pr, why);
}

int sb_mcmc_init_chains(sb_handle *h, const sb_priors *pr, const double *why, int64_t ldy,
                        const uint8_t *heldout, int64_t ldh)
{
    if (!h || !pr || !why) return fail(SB_ERR_ARG, "sb_mcmc_init_chains: null argument");
    if (ldy < 0 || (ldy > 0 && ldy < h->par.n) || (heldout && ldh < 0) || (heldout && ldh > 0 && ldh < h->par.n))
        return fail(SB_ERR_ARG, "sb_mcmc_init_chains: column stride below nind");
    return mcmc_init(h, pr, why, ldy, heldout, ldh);
}

int sb_mcmc_run(sb_handle *h, int32_t first, int32_t count, double *hyp)
{
    if (!h) return fail(SB_ERR_ARG, "sb_mcmc_run: null handle");
    return mcmc_run(h, first, count, hyp);
}

int32_t sb_hyp_width(const sb_handle *h) { return h ? 5 + 2 * h->par.nbins : 0; }

int sb_get(sb_handle *h, int32_t what, int32_t chain, double *out, int64_t count)
{
    if (!h || !out) return fail(SB_ERR_ARG, "sb_get: null argument");
    const Par &P = h->par;
    const int64_t p = P.p;
    const bool by_block = what == SB_Q_GRAM || what == SB_Q_XA_BLOCK;
    if (!by_block && (chain < 0 || chain >= P.nchains)) return fail(SB_ERR_ARG, "sb_get: bad chain");
    const int64_t cp = by_block ? 0 : (int64_t)chain * p;
    hipStream_t st = h->stream;
    auto need = [&](int64_t n) { return count >= n ? SB_OK : fail(SB_ERR_ARG, "sb_get: output too small"); };
    auto by_id = [&](const void *dsrc, bool is_int, int64_t ncol) -> int {
        int rc = need(p * ncol);
        if (rc) return rc;
        std::vector<double> tmp(p * ncol);
        if (is_int) {
            std::vector<int32_t> ti(p);
            if (hip_check(hipMemcpy(ti.data(), dsrc, sizeof(int32_t) * p, hipMemcpyDeviceToHost), "get"))
                return SB_ERR_HIP;
            for (int64_t k = 0; k < p; k++) tmp[k] = ti[k];
        } else if (hip_check(hipMemcpy(tmp.data(), dsrc, sizeof(double) * p * ncol, hipMemcpyDeviceToHost), "get")) {
            return SB_ERR_HIP;
        }
        for (int64_t c = 0; c < ncol; c++)
            for (int64_t k = 0; k < p; k++) out[c * p + h->permvec[k]] = tmp[c * p + k];
        return SB_OK;
    };
    if (what == SB_Q_INDISTSTORE) {
        const int rc = class1_counts(h);
        if (rc) return rc;
    }
    hipStreamSynchronize(st);
    Scal sc;
    bool need_scal = what == SB_Q_P || what == SB_Q_SCALARS || what == SB_Q_STORES || what == SB_Q_SNPSTORE ||
                     what == SB_Q_VARSTORE || what == SB_Q_PSTORE || what == SB_Q_SNPINDIST ||
                     what == SB_Q_VARINDIST;
    if (need_scal && hip_check(hipMemcpy(&sc, h->scal + chain, sizeof sc, hipMemcpyDeviceToHost), "get scal"))
        return SB_ERR_HIP;
    switch (what) {
    case SB_Q_G: return by_id(h->beta + cp, false, 1);
    case SB_Q_CLASS: return by_id(h->cls + cp, true, 1);
    case SB_Q_VARA_S: return by_id(h->vara_s + cp, false, 1);
    case SB_Q_GSTORE: return by_id(h->gstore + cp, false, 1);
    case SB_Q_VARA_SSTORE: return by_id(h->vara_sstore + cp, false, 1);
    case SB_Q_INDISTSTORE: return by_id(h->indiststore + cp * P.nmix, false, P.nmix);
    case SB_Q_YADJ:
        if (need(P.n)) return SB_ERR_ARG;
        return hip_check(hipMemcpy(out, h->r + (int64_t)chain * P.ldr, sizeof(double) * P.n, hipMemcpyDeviceToHost),
                         "get r");
    case SB_Q_WHY:
        if (need(P.n)) return SB_ERR_ARG;
        return hip_check(hipMemcpy(out, h->why + (int64_t)chain * P.ldr, sizeof(double) * P.n,
                                   hipMemcpyDeviceToHost),
                         "get why");
    case SB_Q_YAADJ:
        if (need(P.blocksize)) return SB_ERR_ARG;
        return hip_check(hipMemcpy(out, h->ra + (int64_t)chain * P.blocksize, sizeof(double) * P.blocksize,
                                   hipMemcpyDeviceToHost),
                         "get ra");
    case SB_Q_XXD:
        if (need(P.nblocks)) return SB_ERR_ARG;
        return hip_check(hipMemcpy(out, h->xxd, sizeof(double) * P.nblocks, hipMemcpyDeviceToHost), "get xxd");
    case SB_Q_P:
        if (need(P.nbins)) return SB_ERR_ARG;
        std::memcpy(out, sc.P, sizeof(double) * P.nbins);
        return SB_OK;
    case SB_Q_SNPSTORE:
    case SB_Q_VARSTORE:
    case SB_Q_PSTORE:
    case SB_Q_SNPINDIST:
    case SB_Q_VARINDIST: {
        if (need(P.nbins)) return SB_ERR_ARG;
        const double *src = what == SB_Q_SNPSTORE   ? sc.snpstore
                            : what == SB_Q_VARSTORE ? sc.varstore
                            : what == SB_Q_PSTORE   ? sc.pstore
                            : what == SB_Q_SNPINDIST ? sc.snpindist
                                                     : sc.varindist;
        std::memcpy(out, src, sizeof(double) * P.nbins);
        return SB_OK;
    }
    case SB_Q_SCALARS:
        if (need(4)) return SB_ERR_ARG;
        out[0] = sc.mu;
        out[1] = sc.vare;
        out[2] = sc.vara;
        out[3] = sc.included;
        return SB_OK;
    case SB_Q_STORES:
        if (need(5)) return SB_ERR_ARG;
        out[0] = sc.counter;
        out[1] = sc.mu_store;
        out[2] = sc.vare_store;
        out[3] = sc.vara_store;
        out[4] = sc.included_store;
        return SB_OK;
    case SB_Q_XA: {
        if (need((int64_t)P.blocksize * p)) return SB_ERR_ARG;
        std::vector<double> U(h->tri_off[P.nblocks]);
        if (hip_check(hipMemcpy(U.data(), h->U, sizeof(double) * U.size(), hipMemcpyDeviceToHost), "get U"))
            return SB_ERR_HIP;
        std::memset(out, 0, sizeof(double) * (size_t)P.blocksize * p);
        for (int b = 0; b < P.nblocks; b++) {
            int p0 = h->ls[b], nb = h->ls[b + 1] - p0;
            if (b < (int)h->dense_off.size() && h->dense_off[b] >= 0) {  // pivoted-Cholesky block
                std::vector<double> D((size_t)nb * nb);
                if (hip_check(hipMemcpy(D.data(), h->Ud + h->dense_off[b], sizeof(double) * D.size(),
                                        hipMemcpyDeviceToHost),
                              "get dense XA"))
                    return SB_ERR_HIP;
                for (int j = 0; j < nb; j++)
                    for (int i = 0; i < nb; i++) out[(int64_t)(p0 + j) * P.blocksize + i] = D[(size_t)j * nb + i];
                continue;
            }
            for (int j = 0; j < nb; j++)
                for (int i = 0; i <= j; i++)
                    out[(int64_t)(p0 + j) * P.blocksize + i] = U[h->tri_off[b] + (int64_t)j * (j + 1) / 2 + i];
        }
        return SB_OK;
    }
    case SB_Q_XA_BLOCK: {
        // one block's augmenting matrix, n_b x n_b column-major (block positions)
        if (chain < 0 || chain >= P.nblocks) return fail(SB_ERR_ARG, "block out of range");
        const int b = chain, p0 = h->ls[b], nb = h->ls[b + 1] - p0;
        if (need((int64_t)nb * nb)) return SB_ERR_ARG;
        std::memset(out, 0, sizeof(double) * (size_t)nb * nb);
        if (b < (int)h->dense_off.size() && h->dense_off[b] >= 0)
            return hip_check(hipMemcpy(out, h->Ud + h->dense_off[b], sizeof(double) * (size_t)nb * nb,
                                       hipMemcpyDeviceToHost),
                             "get dense XA block");
        std::vector<double> U(h->tri_off[b + 1] - h->tri_off[b]);
        if (hip_check(hipMemcpy(U.data(), h->U + h->tri_off[b], sizeof(double) * U.size(), hipMemcpyDeviceToHost),
                      "get U block"))
            return SB_ERR_HIP;
        for (int j = 0; j < nb; j++)
            for (int i = 0; i <= j; i++) out[(int64_t)j * nb + i] = U[(int64_t)j * (j + 1) / 2 + i];
        return SB_OK;
    }
    case SB_Q_GRAM: {
        if (chain < 0 || chain >= P.nblocks) return fail(SB_ERR_ARG, "block out of range");
        int64_t nb = h->ls[chain + 1] - h->ls[chain];
        if (need(nb * nb)) return SB_ERR_ARG;
        return setup_gram_block(h, chain, out);
    }
    case SB_Q_KTIMES: {
        static_assert((int)SB_KT_N == (int)sb::KS_N, "kernel timing slots");
        if (need(3 * SB_KT_N)) return SB_ERR_ARG;
        for (int k = 0; k < SB_KT_N; k++) {
            out[3 * k] = h->kc.ms[k];
            out[3 * k + 1] = h->kc.launches[k];
            out[3 * k + 2] = h->kc.tbytes[k];
        }
        return SB_OK;
    }
    case SB_Q_SKIPPED: {
        if (need(h->nskipped)) return SB_ERR_ARG;
        for (int k = 0; k < h->nskipped; k++) out[k] = h->permvec[h->skip_list[k]] + 1;
        return SB_OK;
    }
    case SB_Q_NAIND:
        if (need(2)) return SB_ERR_ARG;
        out[0] = P.naind;
        out[1] = h->nskipped;
        return SB_OK;
    default: return fail(SB_ERR_ARG, "sb_get: unknown quantity");
    }
}

int sb_block_rhs(sb_handle *h, int32_t block, double *out)
{
    if (!h || !out) return fail(SB_ERR_ARG, "sb_block_rhs: null argument");
    return block_rhs(h, block, out);
}


int sb_state_size(sb_handle *h, int64_t *bytes)
{
    if (!h || !bytes) return fail(SB_ERR_ARG, "sb_state_size: null argument");
    if (!h->initialised) return fail(SB_ERR_STATE, "sb_state_size before sb_mcmc_init");
    *bytes = sb::state_bytes(h, h->use_heldout);
    return SB_OK;
}

int sb_state_save(sb_handle *h, void *buf, int64_t size)
{
    if (!h || !buf) return fail(SB_ERR_ARG, "sb_state_save: null argument");
    if (!h->initialised) return fail(SB_ERR_STATE, "sb_state_save before sb_mcmc_init");
    const int64_t need = sb::state_bytes(h, h->use_heldout);
    if (size < need) return fail(SB_ERR_ARG, "sb_state_save: buffer smaller than sb_state_size");
    if (int rc = sb::class1_counts(h)) return rc;   // the checkpoint holds every class's counts
    if (hip_check(hipStreamSynchronize(h->stream), "state save")) return SB_ERR_HIP;
    const std::vector<sb::Section> secs = sb::state_sections(h, h->use_heldout);
    sb::StateHeader hd;
    std::memset(&hd, 0, sizeof hd);
    std::memcpy(hd.magic, sb::STATE_MAGIC, 8);
    hd.bytes = need;
    hd.dims = h->d;
    hd.naind = h->par.naind;
    hd.nskipped = h->nskipped;
    hd.use_heldout = h->use_heldout ? 1 : 0;
    hd.nsec = (int32_t)secs.size();
    hd.iterations = h->iterations;
    for (size_t k = 0; k < secs.size(); k++) hd.sec[k] = (int64_t)secs[k].bytes;
    char *o = static_cast<char *>(buf);
    std::memcpy(o, &hd, sizeof hd);
    o += sizeof hd;
    std::memcpy(o, &h->par, sizeof(sb::Par));
    o += sizeof(sb::Par);
    if (hip_check(hipMemcpy(o, h->xxd, sizeof(double) * h->par.nblocks, hipMemcpyDeviceToHost), "state save"))
        return SB_ERR_HIP;
    o += sizeof(double) * h->par.nblocks;
    for (const sb::Section &s : secs) {
        if (hip_check(hipMemcpy(o, s.dev, s.bytes, hipMemcpyDeviceToHost), "state save")) return SB_ERR_HIP;
        o += s.bytes;
    }
    return SB_OK;
}

int sb_state_load(sb_handle *h, const void *buf, int64_t size)
{
    if (!h || !buf) return fail(SB_ERR_ARG, "sb_state_load: null argument");
    if (!h->aug_built) return fail(SB_ERR_STATE, "sb_state_load before sb_build_aug");
    sb::StateHeader hd;
    if (size < (int64_t)sizeof hd) return fail(SB_ERR_ARG, "sb_state_load: not a checkpoint");
    std::memcpy(&hd, buf, sizeof hd);
    if (std::memcmp(hd.magic, sb::STATE_MAGIC, 8) != 0) return fail(SB_ERR_ARG, "sb_state_load: not a checkpoint");
    if (hd.bytes != size || std::memcmp(&hd.dims, &h->d, sizeof(sb_dims)) != 0 || hd.naind != h->par.naind ||
        hd.nskipped != h->nskipped)
        return fail(SB_ERR_ARG, "sb_state_load: checkpoint of another job shape");
    const bool heldout = hd.use_heldout != 0;
    if (heldout && !h->heldout &&
        sb::dev_alloc(h, (void **)&h->heldout, (size_t)h->par.ldr * h->par.nchains))
        return SB_ERR_NOMEM;
    const std::vector<sb::Section> secs = sb::state_sections(h, heldout);
    if (hd.nsec != (int32_t)secs.size() || sb::state_bytes(h, heldout) != size)
        return fail(SB_ERR_ARG, "sb_state_load: checkpoint layout differs");
    for (size_t k = 0; k < secs.size(); k++)
        if (hd.sec[k] != (int64_t)secs[k].bytes) return fail(SB_ERR_ARG, "sb_state_load: checkpoint layout differs");
    const char *o = static_cast<const char *>(buf) + sizeof hd;
    sb::Par par;
    std::memcpy(&par, o, sizeof par);
    o += sizeof par;
    std::vector<double> xxd(h->par.nblocks);
    if (hip_check(hipMemcpy(xxd.data(), h->xxd, sizeof(double) * xxd.size(), hipMemcpyDeviceToHost), "state load"))
        return SB_ERR_HIP;
    if (std::memcmp(xxd.data(), o, sizeof(double) * xxd.size()) != 0)
        return fail(SB_ERR_ARG, "sb_state_load: the augmenting matrices differ (other genotypes or setup)");
    o += sizeof(double) * xxd.size();
    for (const sb::Section &s : secs) {
        if (hip_check(hipMemcpy(s.dev, o, s.bytes, hipMemcpyHostToDevice), "state load")) return SB_ERR_HIP;
        o += s.bytes;
    }
    h->par = par;   // the structural fields are equal (dims checked); the priors come from the checkpoint
    h->use_heldout = heldout;
    h->iterations = hd.iterations;
    h->rstat_nwg = 0;
    h->q_block = -1;
    h->ra_block = -1;
    int rc = sb::init_xpx(h);
    if (rc) return rc;
    if ((rc = sb::samples_from_counts(h))) return rc;
    if (hip_check(hipStreamSynchronize(h->stream), "state load")) return SB_ERR_HIP;
    h->initialised = true;
    h->graph_ok = false;   // a captured sweep may name another state (held-out rows)
    return SB_OK;
}

int sb_get_info(sb_handle *h, double *out, int32_t n)
{
    if (!h || !out) return fail(SB_ERR_ARG, "sb_get_info: null argument");
    // grouped kernel times: block products, draws, updates
    const sb::KClock &k = h->kc;
    const double prod_ms = k.ms[sb::KS_RHS] + k.ms[sb::KS_RQUANT] + k.ms[sb::KS_XQ] + k.ms[sb::KS_UT0];
    const double prod_n = k.launches[sb::KS_RHS] + k.launches[sb::KS_XQ];
    const double prod_b = k.tbytes[sb::KS_RHS] + k.tbytes[sb::KS_RQUANT] + k.tbytes[sb::KS_XQ] + k.tbytes[sb::KS_UT0];
    const double draw_ms = k.ms[sb::KS_DRAW] + k.ms[sb::KS_DRAWB];
    const double draw_n = k.launches[sb::KS_DRAW] + k.launches[sb::KS_DRAWB];
    const double upd_ms = k.ms[sb::KS_LIST] + k.ms[sb::KS_UPDATE] + k.ms[sb::KS_DQUANT] + k.ms[sb::KS_XUPD] +
                          k.ms[sb::KS_UT1] + k.ms[sb::KS_RARED];
    const double upd_n = k.launches[sb::KS_UPDATE] + k.launches[sb::KS_XUPD];
    double v[15] = {h->setup_seconds, (double)h->iterations, h->last_sweep_ms, (double)h->bytes,
                    prod_ms, prod_n, prod_b, draw_ms, draw_n, upd_ms, upd_n, h->pm_iters_max,
                    (double)h->kss, (double)h->b_rw, h->tune_us};
    for (int i = 0; i < n && i < 15; i++) out[i] = v[i];
    return SB_OK;
}

}  // extern "C"

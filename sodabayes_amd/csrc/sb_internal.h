// Internal state of a SODA sampler handle (device buffers and launchers).
#pragma once
#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>
#include <vector>

#include "sodabayes.h"

namespace sb {

constexpr int MAXMIX = 20;          // globals.f90:8 maxdist
constexpr int MAXGRP = 100;         // globals.f90:34 maxgrp
constexpr int MAXBIN = MAXMIX * MAXGRP;
constexpr int CHUNK_WORDS = 128;    // 2048 individuals per genotype chunk (16 per word)
constexpr int CHUNK_IND = CHUNK_WORDS * 16;
// slots per chain into which the residual update folds its workgroups' max |r|
// (atomic max into slot workgroup mod RSLOT)
constexpr int RSLOT = 16;

// Read-only run parameters, passed by value to kernels.
struct Par {
    int method, n, p, nmix, ngroup, nbins;
    int nblocks, blocksize, blockt, naind;
    int ldw;           // 32-bit genotype words per SNP column (multiple of CHUNK_WORDS)
    int nchunk;        // ldw / CHUNK_WORDS
    int64_t ldr;       // residual stride per chain (ldw * 16)
    int rng_mode, nchains, precision;
    int regular;       // -SODAOff: the single-site sampler (nblocks = blocksize = 0)
    int numit, burnin, thin;
    int resync;        // iterations between residual resyncs (1000 in the reference)
    double pizero, scale_va0, scale_ve, df_va, df_ve, vara0;
    double mix[MAXMIX], alpha[MAXMIX];
    uint64_t seed;
};

// Mutable per-chain scalars, kept on the device (one struct per chain).
struct Scal {
    double mu, vare, vara, scale_va, included;
    double sum_r, rtr;                         // reduction results
    double counter, mu_store, vare_store, vara_store, included_store;
    double gp[MAXMIX], vare_gp[MAXMIX];
    double P[MAXBIN], log_p[MAXBIN];           // (group, class), group fastest
    double snpindist[MAXBIN], varindist[MAXBIN];
    double snpstore[MAXBIN], varstore[MAXBIN], pstore[MAXBIN];
    uint64_t rng[2];                           // replay stream-0 state
    int32_t nchg;                              // changed SNPs of the current block
    int32_t pad;
};

// Per-block replay / draw scratch (one block at a time).
struct Scratch {
    double *part;      // nchunk * blocksize partial X's products
    double *partA;     // blocksize augmented products
    double *partR;     // nchunk partial sums of r
    double *rhs;       // blocksize
    double *stemp;     // blocksize * nmix class probabilities
    double *uu;        // blocksize * 2 recorded uniforms (replay)
    int32_t *flag;     // blocksize drawn class
    int32_t *chg;      // blocksize changed positions (ordered)
    double *delta;     // blocksize beta_old - beta_new of changed positions
    double *red;       // generic reduction partials
};

// Per-kernel timing on a handle's stream (enabled for a run of sb_mcmc_run by
// the environment variable SODABAYES_KERNEL_TIMING=1).  Each slot sums the
// elapsed ms, the launches and the algorithmic bytes of its timed launches.
// The events are taken by the kernels themselves (hipExtLaunchKernelGGL's
// start / stop events), so a slot's time is kernel execution from the start
// of its first kernel to the end of its last, without the dispatch.
enum KSlot {
    KS_RHS, KS_DRAW, KS_LIST, KS_UPDATE,                          // FP64 path
    KS_RQUANT, KS_XQ, KS_UT0, KS_DRAWB, KS_DQUANT, KS_XUPD, KS_UT1, KS_RARED,  // batched path
    KS_N
};

struct KClock {
    bool on = false;
    std::vector<hipEvent_t> ev;   // event pool, reused across runs
    size_t used = 0;
    struct Rec {
        hipEvent_t a, b;
        int slot;
        double bytes;
    };
    std::vector<Rec> recs;        // timed (start, stop) pairs of this run
    double ms[KS_N] = {}, launches[KS_N] = {}, tbytes[KS_N] = {};
    bool want_start = false, want_stop = false;
    hipEvent_t cur_start = nullptr;
    int stop_slot = 0;
    double stop_bytes = 0.0;
    // algorithmic bytes that only the device knows (the changed-SNP columns
    // of the FP64 update), added by kernels per slot while timing is on
    double *dextra = nullptr;
    double *extra(int s)
    {
        if (!on) return nullptr;
        if (!dextra) {
            if (hipMalloc(&dextra, sizeof(double) * KS_N) != hipSuccess) {
                dextra = nullptr;
                return nullptr;
            }
            hipMemset(dextra, 0, sizeof(double) * KS_N);
        }
        return dextra + s;
    }
    hipEvent_t get()
    {
        if (used == ev.size()) {
            hipEvent_t e;
            hipEventCreate(&e);
            ev.push_back(e);
        }
        return ev[used++];
    }
    // the next launch starts a timed span
    void begin(hipStream_t) { want_start = on; }
    // the next launch ends the span, which counts for `slot` with `b` bytes
    void last(int slot, double b)
    {
        want_stop = on;
        stop_slot = slot;
        stop_bytes = b;
    }
    hipEvent_t take_start()
    {
        if (!want_start) return nullptr;
        want_start = false;
        cur_start = get();
        return cur_start;
    }
    hipEvent_t take_stop()
    {
        if (!want_stop) return nullptr;
        want_stop = false;
        hipEvent_t e = get();
        recs.push_back({cur_start, e, stop_slot, stop_bytes});
        return e;
    }
    // after the stream has drained
    void collect()
    {
        for (const Rec &r : recs) {
            float t = 0.f;
            hipEventElapsedTime(&t, r.a, r.b);
            ms[r.slot] += t;
            launches[r.slot] += 1.0;
            tbytes[r.slot] += r.bytes;
        }
        recs.clear();
        used = 0;
        if (dextra) {
            double x[KS_N];
            if (hipMemcpy(x, dextra, sizeof x, hipMemcpyDeviceToHost) == hipSuccess)
                for (int s = 0; s < KS_N; s++) tbytes[s] += x[s];
            hipMemset(dextra, 0, sizeof(double) * KS_N);
        }
    }
    ~KClock()
    {
        for (auto e : ev) hipEventDestroy(e);
        if (dextra) hipFree(dextra);
    }
};

// Launches kernel k; while the clock is on, through hipExtLaunchKernelGGL
// with the clock's pending start / stop events.  Every argument is converted
// to the kernel's own parameter type first.
template <typename... P, typename... A>
void launch(KClock &kc, void (*k)(P...), dim3 grid, dim3 block, uint32_t shm, hipStream_t st, A &&...a)
{
    static_assert(sizeof...(P) == sizeof...(A), "kernel argument count");
    if (kc.on) {
        hipEvent_t e0 = kc.take_start();
        hipEvent_t e1 = kc.take_stop();
        hipExtLaunchKernelGGL(k, grid, block, shm, st, e0, e1, 0u, static_cast<P>(a)...);
    } else {
        k<<<grid, block, shm, st>>>(static_cast<P>(a)...);
    }
}

}  // namespace sb

struct sb_handle {
    sb_dims d;
    sb::Par par;
    int device = 0;
    hipStream_t stream = nullptr;
    std::vector<int> ls;          // block start position (0-based), nblocks+1
    std::vector<int64_t> tri_off; // packed-triangle offset per block, nblocks+1
    std::vector<int32_t> permvec; // SNP id per position
    std::vector<int32_t> pos_of;  // position per SNP id
    size_t bytes = 0;
    bool aug_built = false, initialised = false;
    double setup_seconds = 0.0, last_sweep_ms = 0.0;
    int64_t iterations = 0;
    sb::KClock kc;                // per-kernel timing (SODABAYES_KERNEL_TIMING=1)
    double pm_iters_max = 0.0;        // largest power-method iteration count of the setup
    int test_dshift_block = -1;       // sb_test_shift_d: block whose d is lowered (-1: none)
    double test_dshift = 0.0;         // and by how much

    // genotype data
    uint32_t *geno = nullptr;     // nloci x ldw, position order, 2-bit counts
    double *f2 = nullptr;         // 2 f per position
    double *sdv = nullptr;        // sqrt(2 f (1-f)) per position (0: monomorphic)
    double *zlut = nullptr;       // 4 per position: z(0), z(1), z(2), 0
    int32_t *grp = nullptr;       // 0-based group per position
    int32_t *d_pos_by_id = nullptr; // position of SNP id (resync order)
    // augmenting matrices
    double *xxd = nullptr;        // nblocks
    double *U = nullptr;          // packed upper triangles
    int64_t *d_tri_off = nullptr;
    // pivoted-Cholesky fallback (blocks whose dI - G failed the factorisation)
    double *Ud = nullptr;          // dense n_b x n_b XA blocks of the fallback, column = block position
    std::vector<int64_t> dense_off; // per block: offset in Ud, or -1 (packed triangle in U)
    uint8_t *skip = nullptr;       // per position: 1 = skipped in the block sweep (skip_j)
    std::vector<int32_t> skip_list; // skipped positions in the reference's i_skipped order
    int32_t *d_skip_list = nullptr;
    double *xpx = nullptr;         // per skipped SNP: z'z
    int nskipped = 0;
    // single-site sampler (-SODAOff): windows of 16 SNPs in update order
    uint32_t *win = nullptr;       // per window: ldr words, SNP j's code in bits 2j..2j+1
    double *wz = nullptr;          // per window: z of codes 0..2 of its 16 SNPs (4 per SNP)
    double *wgram = nullptr;       // per window: packed lower triangle of z_j'z_l
    int nwin = 0;
    int32_t *d_ls = nullptr;
    // chain state (nchains copies)
    double *r = nullptr;          // ldw*16 per chain
    double *ra = nullptr;         // blocksize per chain
    double *why = nullptr;        // ldr per chain: phenotype (held-out rows: current draw)
    uint8_t *heldout = nullptr;   // ldr per chain, nonzero = validation row (imputed), or null
    bool use_heldout = false;     // the current chains have held-out rows
    double *beta = nullptr;       // nloci per chain (position order)
    int32_t *cls = nullptr;       // nloci per chain
    double *vara_s = nullptr;     // nloci per chain
    double *gstore = nullptr, *vara_sstore = nullptr, *indiststore = nullptr;
    sb::Scal *scal = nullptr;     // per chain
    double *hyp = nullptr;        // device .hyp records of one sb_mcmc_run call (hyp_cap records)
    int hyp_cap = 0;
    int *d_it = nullptr;          // device iteration state: [iteration, next .hyp record]
    // one sweep iteration captured as a HIP graph (kept across sb_mcmc_run
    // calls while its launch parameters stay valid; freed by sb_destroy)
    hipGraphExec_t gexec = nullptr;
    bool graph_ok = false;        // gexec matches the handle's current state
    sb::Scratch sc{};
    int *u_lst = nullptr;         // changed positions of the current block, per chain (blocksize each)
    double *u_dl = nullptr;       // their beta_old - beta_new
    double *u_cl = nullptr;       // delta / s per listed position
    double *u_kc = nullptr;       // centring constant sum 2 f c per chain
    int *u_cnt = nullptr;         // list length per chain
    double *u_rapart = nullptr;   // r_a partial sums per list chunk (FP64 update)
    int *u_tstart = nullptr;      // per chain, per 64-row tile of r_a: first list entry reaching it
    int32_t *resync_list = nullptr; // positions with beta != 0 in SNP-id order
    int32_t *d_count = nullptr;

    // mixed-precision batched products (precision SB_PREC_MIXED)
    int limbs = 3;                // int8 limbs per residual (8L-1 fixed-point bits)
    int nwg_r = 0;                // workgroups of the residual statistics (k_rstat)
    int ncu = 1;                  // compute units of the device (grid sizes of the int8 products)
    int rstat_nwg = 0;            // valid partials of max |r| in b_rmax (0: stale)
    int q_block = -1;             // block whose r limbs the last update already wrote to b_Q (-1: none)
    uint64_t *b_rmax = nullptr;   // nchains x partials of max |r| (bit patterns)
    int8_t *b_Q = nullptr;        // (16 * b_ntile) x ldq int8 limbs, limb-major rows k * 16 ntc + c
    int64_t b_ldq = 0;
    int *b_expo = nullptr;        // fixed-point exponent per chain
    long long *b_qsum = nullptr;  // nchains x b_nq exact sums of the quantised residuals
    int b_nq = 1;                 // super-steps: quantisation workgroups per chain tile
    int kss = 16;                 // super-steps (16 genotype words) per X_s'Q chunk, about nS / b_nch
    int b_nch = 1;                // X_s'Q chunks (of the individuals) per block
    int b_rw = 4;                 // row tiles per wave of the int8 products (2 or 4)
    bool tuned = false;           // tile shape chosen (batch_autotune, or fixed by the environment)
    double tune_us = 0.0;         // block-0 products + update time of the chosen shape
    double *tune_r = nullptr;     // scratch residuals of the tuning updates
    int b_ntile = 3;              // 16-column MFMA tiles: limbs x ceil(nchains / 16)
    int *b_T = nullptr;           // chunks x nchains x blocksize int64 partials of X_s'Q
    int uq_group = 8;             // K steps per group of the int8 U products (batch_uq_group)
    int rows_per_group = 64 * 8;  // row split of the augmented product: 64 uq_group
    double *b_rapart = nullptr;   // (blocksize / rows_per_group) x blocksize x nchains
    double *b_uapart = nullptr;   // same shape: partial U_s delta of the update
    int8_t *b_Dq = nullptr;       // (16 * b_ntile) x b_ldd int8 limbs of delta / s
    int64_t b_ldd = 0;
    int *b_expo2 = nullptr;       // fixed-point exponent of delta / s per chain
    uint64_t *b_dmax = nullptr;   // per chain, per draw workgroup: max |c_j| (bit patterns)
    double *b_kpart = nullptr;    // per chain, per quantisation workgroup: sum 2 f c~_j (centring constant)
    double *b_cj = nullptr;       // blocksize per chain: c_j = delta_j / s_j + carry_j of the current block
    double *b_carry = nullptr;    // nloci per chain (position order): remainder of the last quantised update
    uint32_t *geno_t = nullptr;   // individual-major genotypes per block, MFMA operand tiles (update GEMM)
    std::vector<int64_t> gt_off;  // 16-byte element offset of each block in geno_t
    uint32_t *geno_sw = nullptr;  // SNP-major genotypes per block, MFMA operand tiles (product GEMM)
    std::vector<int64_t> gs_off;  // 16-byte element offset of each block in geno_sw
    int8_t *Uq0 = nullptr;        // U as 4 int8 limbs per entry, A-operand tiles of U' r_a (mode 0)
    int8_t *Uq1 = nullptr;        // U diag(s) as 4 int8 limbs per entry, A-operand tiles of U c~ (mode 1)
    std::vector<int64_t> uq0_off; // uint4 offset of each block in Uq0
    std::vector<int64_t> uq1_off; // uint4 offset of each block in Uq1
    int *ue0 = nullptr;           // nblocks x blocksize limb exponents of U's columns (mode 0)
    int *ue1 = nullptr;           // nblocks x blocksize limb exponents of U diag(s)'s rows (mode 1)
    unsigned long long *b_ramax = nullptr;   // 2 x nchains: max |r_a| bits per chain, block parity
    int ra_block = -1;            // block whose r_a maxima the last update already collected (-1: none)

    std::vector<void *> allocs;
};

namespace sb {
// error helpers (sb_api.hip)
int fail(int code, const std::string &msg);
int hip_check(hipError_t e, const char *what);
int dev_alloc(sb_handle *h, void **p, size_t bytes);

// Device buffers that live for one call: every allocation is checked
// (SB_ERR_NOMEM with the size and the buffer's name in sb_last_error) and
// everything is freed when the object goes out of scope, on every return
// path.  The caller synchronises its stream before returning.
struct TmpDev {
    std::vector<void *> ptrs;
    template <class T>
    int alloc(T **p, size_t bytes, const char *what)
    {
        void *q = nullptr;
        hipError_t e = hipMalloc(&q, bytes ? bytes : 8);
        if (e != hipSuccess) {
            *p = nullptr;
            return fail(SB_ERR_NOMEM, std::string("hipMalloc(") + what + ", " + std::to_string(bytes) +
                                          " B): " + hipGetErrorString(e));
        }
        ptrs.push_back(q);
        *p = static_cast<T *>(q);
        return SB_OK;
    }
    ~TmpDev()
    {
        for (void *q : ptrs) hipFree(q);
    }
};

// setup (sb_setup.hip)
int setup_pack(sb_handle *h, const uint8_t *bed, bool on_device);
int setup_build_aug(sb_handle *h, double *xxd_out, int32_t *fail_block, int32_t *fail_col);
int setup_gram_block(sb_handle *h, int block, double *G_host);

// MCMC (sb_sweep.hip)
int mcmc_init(sb_handle *h, const sb_priors *pr, const double *why, int64_t ldy = 0,
              const uint8_t *heldout = nullptr, int64_t ldh = 0);
int mcmc_run(sb_handle *h, int first, int count, double *hyp);
int class1_counts(sb_handle *h);
int samples_from_counts(sb_handle *h);
int counter_uniforms(int64_t seed, int chain, const uint32_t *ctr, int n, int per, double *out);

// chain-batched mixed-precision block products and draws (sb_batch.hip)
int batch_block(sb_handle *h, int b, double zz, int rep);
int batch_draw(sb_handle *h, int b, double zz, double *delta, uint8_t *chg, double *rhs_out = nullptr);
int batch_update(sb_handle *h, int b);
int batch_clear_carry(sb_handle *h);
int layout_stream(sb_handle *h, double *out);
int batch_transpose(sb_handle *h);
int batch_words(int nb);
int64_t batch_uq_elems(int nb, int mode);
int prod_chunks(const sb_handle *h, int rw, int q);
int batch_uq_group(int nchains);
int batch_max_chunks(const sb_handle *h);
int batch_autotune(sb_handle *h);
int init_xpx(sb_handle *h);
// single-site sampler (sb_single.hip)
int single_setup(sb_handle *h);
int single_sweep(sb_handle *h, int rep);
int batch_tiles(sb_handle *h);
int block_rhs(sb_handle *h, int b, double *out_host);
}  // namespace sb

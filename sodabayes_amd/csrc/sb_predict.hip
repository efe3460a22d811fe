// Genomic values of the prediction run (-predict, SODABayes.f90:145-148):
// xscale's prediction branch standardises the genotypes with the training
// frequencies (utilities.f90:540-553) and compute_dgv adds the mean to the
// row dot products with the effects (utilities.f90:616-622):
//   pred_i = mu + sum_j z_ij g_j,   z_ij = (x_ij - 2 f_j) / sqrt(2 f_j (1 - f_j))
// (z = 0 for monomorphic SNPs).  Evaluated in centred form over the packed
// two-bit counts, never forming z:
//   pred_i = (mu - sum_j 2 f_j c_j) + sum_j x_ij c_j,   c_j = g_j / s_j.
// The GEMV streams the packed genotypes once: a thread owns one 32-bit word
// (16 individuals) of one SNP chunk, workgroups tile (words x chunks), and
// the chunk partial sums are added in chunk order (deterministic).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <vector>

#include "sb_device.h"
#include "sb_internal.h"
#include "sodabayes.h"

namespace sb {
namespace {

constexpr int PT = 256;   // threads per workgroup
constexpr int PU = 8;     // SNPs in flight per thread
constexpr double DEPS = 2.220446049250313080847e-16;   // epsilon(1.0d0)

// .bed rows (SNP-major, ceil(n/4) bytes each) -> 16 counts per 32-bit word,
// 2 bits each, code c -> c - (c >> 1 & 1) (utilities.f90:446-451)
// (grid-stride over the nloci x ldw words: nloci can exceed a grid dimension)
__global__ void k_pred_pack(const uint8_t *__restrict__ bed, int64_t nbytes, int p, int ldw,
                            uint32_t *__restrict__ geno)
{
    const int64_t total = (int64_t)p * ldw;
    for (int64_t k = (int64_t)blockIdx.x * PT + threadIdx.x; k < total; k += (int64_t)gridDim.x * PT) {
        const int64_t j = k / ldw;
        const int64_t b0 = (k - j * ldw) * 4;
        const uint8_t *row = bed + j * nbytes;
        uint32_t v = 0;
#pragma unroll
        for (int b = 0; b < 4; b++)
            if (b0 + b < nbytes) v |= (uint32_t)row[b0 + b] << (8 * b);
        geno[k] = v - ((v >> 1) & 0x55555555u);
    }
}

// partial[ch][16 w + e] = sum_{j in chunk ch} x_(16w+e, j) c_j
__global__ __launch_bounds__(PT) void k_pred_gemv(const uint32_t *__restrict__ geno, int ldw, int p, int clen,
                                                  const double *__restrict__ c, double *__restrict__ partial,
                                                  int64_t ldp)
{
    const int w = blockIdx.x * PT + threadIdx.x;
    const int ch = blockIdx.y;
    const int j0 = ch * clen, j1 = min(p, j0 + clen);
    if (w >= ldw || j0 >= j1) return;
    double acc[16];
#pragma unroll
    for (int e = 0; e < 16; e++) acc[e] = 0.0;
    for (int j = j0; j < j1; j += PU) {
        uint32_t g[PU];
        double cv[PU];
        // unconditional loads (clamped), masked afterwards
#pragma unroll
        for (int u = 0; u < PU; u++) {
            const int jj = min(j + u, j1 - 1);
            g[u] = geno[(int64_t)jj * ldw + w];
            cv[u] = c[jj];
        }
#pragma unroll
        for (int u = 0; u < PU; u++) {
            const double cu = j + u < j1 ? cv[u] : 0.0;
#pragma unroll
            for (int e = 0; e < 16; e++) acc[e] = fma((double)((g[u] >> (2 * e)) & 3u), cu, acc[e]);
        }
    }
    double *out = partial + (int64_t)ch * ldp + 16 * (int64_t)w;
#pragma unroll
    for (int e = 0; e < 16; e++) out[e] = acc[e];
}

__global__ void k_pred_sum(const double *__restrict__ partial, int nch, int64_t ldp, int n, double base,
                           double *__restrict__ pred)
{
    const int i = blockIdx.x * PT + threadIdx.x;
    if (i >= n) return;
    double s = 0.0;
    for (int ch = 0; ch < nch; ch++) s += partial[(int64_t)ch * ldp + i];
    pred[i] = base + s;
}

// ---- the variant that allocates nothing (sb_predict_noalloc) ----
//
// Its partial sums live in `scratch`, a fixed device array of the library
// (each device holds its own copy until the process ends); the inputs stay
// in the caller's .bed layout, read byte-wise, so no packed copy is needed.
// Calls on one device share the array: calls that overlap on one device race.
constexpr int64_t NOALLOC_MAX = SB_PREDICT_NOALLOC_MAX_PARTIALS;
__device__ __align__(256) double scratch[NOALLOC_MAX];

// partial[ch][i] = sum over the SNPs j of chunk ch of (x_ij - 2 f_j) c_j,
// c_j = g_j / sqrt(2 f_j (1 - f_j)) (0 for monomorphic SNPs), for the 16
// individuals 16 w .. 16 w + 15 of thread w (four .bed bytes)
__global__ __launch_bounds__(PT) void k_pred_bytes(const uint8_t *__restrict__ bed, int64_t nbytes, int n, int p,
                                                   int clen, const double *__restrict__ freq,
                                                   const double *__restrict__ g, double *__restrict__ partial)
{
    const int w = blockIdx.x * PT + threadIdx.x;
    const int ch = blockIdx.y;
    const int j0 = ch * clen, j1 = min(p, j0 + clen);
    if (16 * w >= n || j0 >= j1) return;
    double acc[16];
#pragma unroll
    for (int e = 0; e < 16; e++) acc[e] = 0.0;
    const int64_t b0 = 4 * (int64_t)w;
    for (int j = j0; j < j1; j++) {
        const double f = freq[j];
        const bool mono = f < DEPS || (1.0 - f) < DEPS;
        const double c = mono ? 0.0 : g[j] / sqrt(2.0 * f * (1.0 - f));
        const double f2 = 2.0 * f;
        const uint8_t *row = bed + (int64_t)j * nbytes;
        uint32_t v = 0;
#pragma unroll
        for (int b = 0; b < 4; b++) v |= (uint32_t)(b0 + b < nbytes ? row[b0 + b] : 0) << (8 * b);
        v = v - ((v >> 1) & 0x55555555u);   // codes -> counts, missing -> 1 (utilities.f90:446-451)
#pragma unroll
        for (int e = 0; e < 16; e++) acc[e] = fma((double)((v >> (2 * e)) & 3u) - f2, c, acc[e]);
    }
    double *out = partial + (int64_t)ch * n + 16 * (int64_t)w;
#pragma unroll
    for (int e = 0; e < 16; e++)
        if (16 * w + e < n) out[e] = acc[e];
}

__global__ void k_pred_sum_n(const double *__restrict__ partial, int nch, int n, double mu,
                             double *__restrict__ pred)
{
    const int i = blockIdx.x * PT + threadIdx.x;
    if (i >= n) return;
    double s = 0.0;
    for (int ch = 0; ch < nch; ch++) s += partial[(int64_t)ch * n + i];
    pred[i] = mu + s;
}

}  // namespace

// Device address of `scratch` on each device, looked up once per device
// (hipGetSymbolAddress with that device current); host code uses only
// these addresses, never the symbol itself.
namespace {
constexpr int MAX_DEVICES = 64;
double *g_scratch[MAX_DEVICES] = {};
std::mutex g_scratch_mu;
}  // namespace

static int scratch_address(int dev, double **out)
{
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (dev < 0 || dev >= MAX_DEVICES) return fail(SB_ERR_ARG, "sb_predict_noalloc: device index out of range");
    if (!g_scratch[dev]) {
        void *ptr = nullptr;
        const int rc = hip_check(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)");
        if (rc) return rc;
        g_scratch[dev] = static_cast<double *>(ptr);
    }
    *out = g_scratch[dev];
    return SB_OK;
}

int predict_noalloc(int32_t nind, int32_t nloci, const uint8_t *bed, const double *freq, const double *effects,
                    double mu, double *pred, void *stream)
{
    if (nind < 1 || nloci < 1 || !bed || !freq || !effects || !pred)
        return fail(SB_ERR_ARG, "sb_predict_noalloc: bad sizes or null argument");
    if ((int64_t)nind > NOALLOC_MAX)
        return fail(SB_ERR_ARG, "sb_predict_noalloc: nind " + std::to_string(nind) + " exceeds the scratch limit of " +
                                    std::to_string(NOALLOC_MAX) + " individuals");
    int dev = 0;
    int rc = hip_check(hipGetDevice(&dev), "hipGetDevice");
    double *part = nullptr;
    if (!rc) rc = scratch_address(dev, &part);
    if (rc) return rc;
    hipStream_t st = static_cast<hipStream_t>(stream);
    const int64_t nbytes = ((int64_t)nind + 3) / 4;
    const int nw = (nind + 15) / 16, wgx = (nw + PT - 1) / PT;
    // SNP chunks: enough workgroups to fill the device, within the scratch
    const int64_t room = NOALLOC_MAX / nind;
    const int want = (int)std::min<int64_t>(room, std::max(1, 2048 / wgx));
    const int clen = std::max(1, (nloci + want - 1) / want);
    const int nch = (nloci + clen - 1) / clen;   // nch <= want <= room: nch * nind fits
    k_pred_bytes<<<dim3(wgx, nch), PT, 0, st>>>(bed, nbytes, nind, nloci, clen, freq, effects, part);
    k_pred_sum_n<<<(nind + PT - 1) / PT, PT, 0, st>>>(part, nch, nind, mu, pred);
    return hip_check(hipGetLastError(), "sb_predict_noalloc kernels");
}

int predict(int32_t nind, int32_t nloci, const uint8_t *bed, int32_t bed_on_device, const double *freq,
            const double *effects, double mu, double *pred)
{
    if (nind < 1 || nloci < 1 || !bed || !freq || !effects || !pred)
        return fail(SB_ERR_ARG, "sb_predict: bad sizes or null argument");
    const int64_t nbytes = ((int64_t)nind + 3) / 4;
    const int ldw = (nind + 15) / 16;
    const int64_t ldp = 16 * (int64_t)ldw;
    // c_j = g_j / s_j and the centring constant, in SNP order on the host
    std::vector<double> c((size_t)nloci);
    double K = 0.0;
    for (int j = 0; j < nloci; j++) {
        const double f = freq[j];
        const bool mono = f < DEPS || (1.0 - f) < DEPS;
        c[j] = mono ? 0.0 : effects[j] / std::sqrt(2.0 * f * (1.0 - f));
        K += 2.0 * f * c[j];
    }
    // chunks: enough (word group x chunk) workgroups to fill the device
    const int wgx = (ldw + PT - 1) / PT;
    const int want = std::max(1, 2048 / wgx);
    const int clen = std::max(PU, (int)((nloci + want - 1) / want));
    const int nch = (nloci + clen - 1) / clen;
    hipStream_t st = nullptr;
    uint8_t *dbed = nullptr;
    uint32_t *geno = nullptr;
    double *dc = nullptr, *part = nullptr, *dpred = nullptr;
    int rc = hip_check(hipStreamCreateWithFlags(&st, hipStreamNonBlocking), "sb_predict stream");
    auto alloc = [&](void **ptr, size_t bytes) {
        if (!rc && hip_check(hipMalloc(ptr, bytes), "sb_predict hipMalloc")) rc = SB_ERR_NOMEM;
    };
    if (!bed_on_device) alloc((void **)&dbed, (size_t)(nbytes * nloci));
    alloc((void **)&geno, sizeof(uint32_t) * (size_t)ldw * nloci);
    alloc((void **)&dc, sizeof(double) * (size_t)nloci);
    alloc((void **)&part, sizeof(double) * (size_t)(ldp * nch));
    alloc((void **)&dpred, sizeof(double) * (size_t)nind);
    const uint8_t *src = bed_on_device ? bed : dbed;
    if (!rc && !bed_on_device)
        rc = hip_check(hipMemcpyAsync(dbed, bed, (size_t)(nbytes * nloci), hipMemcpyHostToDevice, st), "upload bed");
    if (!rc)
        rc = hip_check(hipMemcpyAsync(dc, c.data(), sizeof(double) * nloci, hipMemcpyHostToDevice, st),
                       "upload effects");
    if (!rc) {
        const int64_t words = (int64_t)nloci * ldw;
        const int pwg = (int)std::min<int64_t>((words + PT - 1) / PT, 16384);
        k_pred_pack<<<pwg, PT, 0, st>>>(src, nbytes, nloci, ldw, geno);
        k_pred_gemv<<<dim3(wgx, nch), PT, 0, st>>>(geno, ldw, nloci, clen, dc, part, ldp);
        k_pred_sum<<<(nind + PT - 1) / PT, PT, 0, st>>>(part, nch, ldp, nind, mu - K, dpred);
        rc = hip_check(hipGetLastError(), "sb_predict kernels");
    }
    if (!rc)
        rc = hip_check(hipMemcpyAsync(pred, dpred, sizeof(double) * nind, hipMemcpyDeviceToHost, st), "download");
    if (st) {
        const int rs = hip_check(hipStreamSynchronize(st), "sb_predict");
        if (!rc) rc = rs;
    }
    hipFree(dbed);
    hipFree(geno);
    hipFree(dc);
    hipFree(part);
    hipFree(dpred);
    if (st) hipStreamDestroy(st);
    return rc;
}

}  // namespace sb

extern "C" int sb_predict(int32_t nind, int32_t nloci, const uint8_t *bed, int32_t bed_on_device, const double *freq,
                          const double *effects, double mu, double *pred)
{
    return sb::predict(nind, nloci, bed, bed_on_device, freq, effects, mu, pred);
}

extern "C" int sb_predict_noalloc(int32_t nind, int32_t nloci, const uint8_t *bed, const double *freq,
                                  const double *effects, double mu, double *pred, void *stream)
{
    return sb::predict_noalloc(nind, nloci, bed, freq, effects, mu, pred, stream);
}

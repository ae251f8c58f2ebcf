// rjk_j.hip — Coulomb (J) build on packed B, two streaming passes.
//
// Pass 1 (J-γ):  γ_P = Σ_p B_{p,P} w_p, with w the packed density weights
//   (w_p = 2 D_ij for i<j, D_ii on the diagonal: the "2·full − diag" form of
//   vj_upper_with_rimatr_v_sync, src/scf_io/mod.rs:2516-2529).  One work-group
//   reads G auxiliary columns over one chunk of pairs, so w is fetched once per
//   G columns; partial γ per chunk go to gpart and are summed by a small
//   kernel (deterministic, no atomics).
// Pass 2 (J-acc): J_p = Σ_P B_{p,P} γ_P (src/scf_io/mod.rs:2531-2541).  Each
//   thread owns 16 bytes of pairs and walks P; γ_P is wave-uniform and is
//   read through scalar loads.  Partial sums per P-split go to jpart, then
//   J = scale·Σ_split jpart (scale last, as in REST).
//
// Both passes stream B once: the J build moves 2·|B| bytes of HBM traffic,
// its roofline bound.  Spins share the B stream (both γ, both J at once).
//
// The allocation-free J entry point at the end of this file runs the same
// kernels on a file-scope device array instead of the context's workspace.
#include <mutex>

#include "rjk_internal.hpp"

namespace rjk {
namespace {

typedef double d2 __attribute__((ext_vector_type(2)));
typedef float f4 __attribute__((ext_vector_type(4)));

template <typename BT> struct V16;
template <> struct V16<double> { using T = d2; static constexpr int N = 2; };
template <> struct V16<float> { using T = f4; static constexpr int N = 4; };

constexpr int kJBlock = 256;
constexpr int kGammaCols = 8;

__device__ inline double wave_sum(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// packed index p -> (i, j), i <= j, p = j(j+1)/2 + i
__device__ inline void unpair(size_t p, size_t* i, size_t* j) {
    size_t jj = (size_t)((sqrt(8.0 * (double)p + 1.0) - 1.0) * 0.5);
    while (jj * (jj + 1) / 2 > p) --jj;
    while ((jj + 1) * (jj + 2) / 2 <= p) ++jj;
    *j = jj;
    *i = p - jj * (jj + 1) / 2;
}

__global__ __launch_bounds__(256) void k_dm_weights(const double* __restrict__ dm, size_t nbf, size_t npair,
                                                    size_t ldb, int nspin, int full_dm, double* __restrict__ w) {
    size_t p = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= ldb) return;
    if (p >= npair) {
        for (int s = 0; s < nspin; ++s) w[s * ldb + p] = 0.0;
        return;
    }
    size_t i, j;
    unpair(p, &i, &j);
    for (int s = 0; s < nspin; ++s) {
        const double* d = dm + (size_t)s * nbf * nbf;
        double v;
        if (i == j) v = d[i + j * nbf];
        else if (full_dm) v = d[i + j * nbf] + d[j + i * nbf];   // Σ_{kl} over both triangles
        else v = 2.0 * d[i + j * nbf];                         // upper triangle only
        w[s * ldb + p] = v;
    }
}

template <typename BT, int NS, int G>
__global__ __launch_bounds__(kJBlock) void k_jgamma(const BT* __restrict__ b, size_t ldb, int naux,
                                                    const double* __restrict__ w, size_t vec_per_chunk,
                                                    double* __restrict__ gpart) {
    using VT = typename V16<BT>::T;
    constexpr int N = V16<BT>::N;
    const size_t nvec = ldb / N;
    const size_t v0 = (size_t)blockIdx.x * vec_per_chunk;
    const size_t v1 = min(nvec, v0 + vec_per_chunk);
    const int P0 = blockIdx.y * G;
    const int gcnt = min(G, naux - P0);
    const VT* bv = reinterpret_cast<const VT*>(b) + (size_t)P0 * nvec;
    const d2* wv2 = reinterpret_cast<const d2*>(w);

    double acc[NS][G];
#pragma unroll
    for (int s = 0; s < NS; ++s)
#pragma unroll
        for (int g = 0; g < G; ++g) acc[s][g] = 0.0;

    if (gcnt == G) {
        for (size_t v = v0 + threadIdx.x; v < v1; v += kJBlock) {
            double wl[NS][N];
#pragma unroll
            for (int s = 0; s < NS; ++s)
#pragma unroll
                for (int h = 0; h < N / 2; ++h) {
                    d2 t = wv2[(s * ldb) / 2 + v * (N / 2) + h];
                    wl[s][2 * h] = t.x;
                    wl[s][2 * h + 1] = t.y;
                }
            VT x[G];
#pragma unroll
            for (int g = 0; g < G; ++g) x[g] = __builtin_nontemporal_load(bv + (size_t)g * nvec + v);
#pragma unroll
            for (int g = 0; g < G; ++g)
#pragma unroll
                for (int s = 0; s < NS; ++s)
#pragma unroll
                    for (int e = 0; e < N; ++e) acc[s][g] = fma((double)x[g][e], wl[s][e], acc[s][g]);
        }
    } else {
        for (size_t v = v0 + threadIdx.x; v < v1; v += kJBlock) {
            double wl[NS][N];
#pragma unroll
            for (int s = 0; s < NS; ++s)
#pragma unroll
                for (int h = 0; h < N / 2; ++h) {
                    d2 t = wv2[(s * ldb) / 2 + v * (N / 2) + h];
                    wl[s][2 * h] = t.x;
                    wl[s][2 * h + 1] = t.y;
                }
#pragma unroll
            for (int g = 0; g < G; ++g) {
                if (g < gcnt) {
                    VT x = bv[(size_t)g * nvec + v];
#pragma unroll
                    for (int s = 0; s < NS; ++s)
#pragma unroll
                        for (int e = 0; e < N; ++e) acc[s][g] = fma((double)x[e], wl[s][e], acc[s][g]);
                }
            }
        }
    }

    __shared__ double red[kJBlock / 64][NS * G];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int s = 0; s < NS; ++s)
#pragma unroll
        for (int g = 0; g < G; ++g) {
            double t = wave_sum(acc[s][g]);
            if (lane == 0) red[wave][s * G + g] = t;
        }
    __syncthreads();
    if (threadIdx.x < NS * G) {
        int s = threadIdx.x / G, g = threadIdx.x % G;
        if (g < gcnt) {
            double t = 0;
#pragma unroll
            for (int q = 0; q < kJBlock / 64; ++q) t += red[q][threadIdx.x];
            gpart[((size_t)blockIdx.x * NS + s) * naux + P0 + g] = t;
        }
    }
}

__global__ void k_gamma_reduce(const double* __restrict__ gpart, int nchunk, int nspin, int naux,
                               double* __restrict__ gamma) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nspin * naux) return;
    int s = t / naux, P = t % naux;
    double acc = 0;
    for (int c = 0; c < nchunk; ++c) acc += gpart[((size_t)c * nspin + s) * naux + P];
    gamma[(size_t)s * naux + P] = acc;
}

template <typename BT, int NS>
__global__ __launch_bounds__(kJBlock) void k_jacc(const BT* __restrict__ b, size_t ldb, int naux,
                                                  const double* __restrict__ gamma, int p_per_split,
                                                  double* __restrict__ jpart) {
    using VT = typename V16<BT>::T;
    constexpr int N = V16<BT>::N;
    const size_t nvec = ldb / N;
    const size_t v = (size_t)blockIdx.x * kJBlock + threadIdx.x;
    const int Pb = blockIdx.y * p_per_split;
    const int Pe = min(naux, Pb + p_per_split);
    if (v >= nvec) return;
    double acc[NS][N];
#pragma unroll
    for (int s = 0; s < NS; ++s)
#pragma unroll
        for (int e = 0; e < N; ++e) acc[s][e] = 0.0;
    const VT* bp = reinterpret_cast<const VT*>(b) + (size_t)Pb * nvec + v;
    int P = Pb;
    constexpr int U = 8;
    for (; P + U <= Pe; P += U, bp += U * nvec) {
        VT x[U];
#pragma unroll
        for (int u = 0; u < U; ++u) x[u] = __builtin_nontemporal_load(bp + (size_t)u * nvec);
#pragma unroll
        for (int u = 0; u < U; ++u)
#pragma unroll
            for (int s = 0; s < NS; ++s) {
                const double g = gamma[(size_t)s * naux + P + u];
#pragma unroll
                for (int e = 0; e < N; ++e) acc[s][e] = fma((double)x[u][e], g, acc[s][e]);
            }
    }
    for (; P < Pe; ++P, bp += nvec) {
        VT x = *bp;
#pragma unroll
        for (int s = 0; s < NS; ++s) {
            const double g = gamma[(size_t)s * naux + P];
#pragma unroll
            for (int e = 0; e < N; ++e) acc[s][e] = fma((double)x[e], g, acc[s][e]);
        }
    }
#pragma unroll
    for (int s = 0; s < NS; ++s) {
        double* o = jpart + ((size_t)blockIdx.y * NS + s) * ldb + v * N;
#pragma unroll
        for (int h = 0; h < N / 2; ++h) {
            d2 t;
            t.x = acc[s][2 * h];
            t.y = acc[s][2 * h + 1];
            *reinterpret_cast<d2*>(o + 2 * h) = t;
        }
    }
}

__global__ void k_jreduce(const double* __restrict__ jpart, int nsplit, int nspin, size_t ldb, size_t npair,
                          double scale, double* __restrict__ vj) {
    size_t p = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= npair) return;
    for (int s = 0; s < nspin; ++s) {
        double acc = 0;
        for (int q = 0; q < nsplit; ++q) acc += jpart[((size_t)q * nspin + s) * ldb + p];
        vj[(size_t)s * npair + p] = acc * scale;
    }
}

template <typename BT>
hipError_t jgamma_t(hipStream_t st, const BT* b, size_t ldb, size_t naux, int nspin, const double* w,
                    size_t nchunk, double* gpart) {
    constexpr int N = V16<BT>::N;
    const size_t nvec = ldb / N;
    size_t vpc = round_up((nvec + nchunk - 1) / nchunk, kJBlock);
    dim3 grid((unsigned)nchunk, (unsigned)((naux + kGammaCols - 1) / kGammaCols));
    if (nspin == 1)
        hipLaunchKernelGGL((k_jgamma<BT, 1, kGammaCols>), grid, dim3(kJBlock), 0, st, b, ldb, (int)naux, w, vpc,
                           gpart);
    else
        hipLaunchKernelGGL((k_jgamma<BT, 2, kGammaCols>), grid, dim3(kJBlock), 0, st, b, ldb, (int)naux, w, vpc,
                           gpart);
    return hipGetLastError();
}

template <typename BT>
hipError_t jacc_t(hipStream_t st, const BT* b, size_t ldb, size_t naux, int nspin, const double* gamma,
                  size_t nsplit, double* jpart) {
    constexpr int N = V16<BT>::N;
    const size_t nvec = ldb / N;
    int pps = (int)((naux + nsplit - 1) / nsplit);
    dim3 grid((unsigned)((nvec + kJBlock - 1) / kJBlock), (unsigned)nsplit);
    if (nspin == 1)
        hipLaunchKernelGGL((k_jacc<BT, 1>), grid, dim3(kJBlock), 0, st, b, ldb, (int)naux, gamma, pps, jpart);
    else
        hipLaunchKernelGGL((k_jacc<BT, 2>), grid, dim3(kJBlock), 0, st, b, ldb, (int)naux, gamma, pps, jpart);
    return hipGetLastError();
}

}  // namespace

void j_plan(size_t ldb, size_t naux, int prec, size_t* nchunk, size_t* nsplit) {
    const size_t N = (prec == RJK_PREC_FP64) ? 2 : 4;
    const size_t nvec = ldb / N;
    const size_t target = 4096;  // work-groups per pass: ≫ 256 CUs
    size_t groups = (naux + kGammaCols - 1) / kGammaCols;
    size_t c = (target + groups - 1) / groups;
    size_t cmax = std::max<size_t>(1, nvec / (4 * kJBlock));
    *nchunk = std::min(std::max<size_t>(c, 1), cmax);
    size_t bx = (nvec + kJBlock - 1) / kJBlock;
    size_t sp = (target + bx - 1) / bx;
    size_t smax = std::max<size_t>(1, naux / 16);
    *nsplit = std::min(std::max<size_t>(sp, 1), smax);
}

hipError_t launch_dm_weights(hipStream_t st, const double* dm, size_t nbf, size_t npair, size_t ldb, int nspin,
                             int full_dm, double* w) {
    hipLaunchKernelGGL(k_dm_weights, dim3((unsigned)((ldb + 255) / 256)), dim3(256), 0, st, dm, nbf, npair, ldb,
                       nspin, full_dm, w);
    return hipGetLastError();
}

hipError_t launch_jgamma(hipStream_t st, int prec, const void* b, size_t ldb, size_t naux, int nspin,
                         const double* w, size_t nchunk, double* gpart) {
    if (prec == RJK_PREC_FP64)
        return jgamma_t<double>(st, (const double*)b, ldb, naux, nspin, w, nchunk, gpart);
    return jgamma_t<float>(st, (const float*)b, ldb, naux, nspin, w, nchunk, gpart);
}

hipError_t launch_gamma_reduce(hipStream_t st, const double* gpart, size_t nchunk, int nspin, size_t naux,
                               double* gamma) {
    size_t n = (size_t)nspin * naux;
    hipLaunchKernelGGL(k_gamma_reduce, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, gpart, (int)nchunk,
                       nspin, (int)naux, gamma);
    return hipGetLastError();
}

hipError_t launch_jacc(hipStream_t st, int prec, const void* b, size_t ldb, size_t naux, int nspin,
                       const double* gamma, size_t nsplit, double* jpart) {
    if (prec == RJK_PREC_FP64) return jacc_t<double>(st, (const double*)b, ldb, naux, nspin, gamma, nsplit, jpart);
    return jacc_t<float>(st, (const float*)b, ldb, naux, nspin, gamma, nsplit, jpart);
}

hipError_t launch_jreduce(hipStream_t st, const double* jpart, size_t nsplit, int nspin, size_t ldb, size_t npair,
                          double scale, double* vj) {
    hipLaunchKernelGGL(k_jreduce, dim3((unsigned)((npair + 255) / 256)), dim3(256), 0, st, jpart, (int)nsplit, nspin,
                       ldb, npair, scale, vj);
    return hipGetLastError();
}

}  // namespace rjk

// ---------------------------------------------------------------------------
// Allocation-free J.  Its workspace is `scratch`, one fixed device array per GPU
// (held until the process ends).  Host code never passes `scratch` itself to a
// kernel: the array's name is a host-side stand-in there, and the device address
// comes from hipGetSymbolAddress, looked up once per device by scratch_init()
// (called from rjk_create with that device current) and kept per device.
// Calls on one GPU share the array: calls that overlap on one GPU race.
// ---------------------------------------------------------------------------
namespace rjk {
constexpr size_t kScratchBytes = size_t(128) << 20;   // (H2O)64's J fits with two spins
constexpr int kScratchMaxDevices = 64;
}  // namespace rjk

static __device__ __align__(256) unsigned char scratch[rjk::kScratchBytes];

namespace rjk {
namespace {
std::mutex g_scratch_mu;
void* g_scratch_dev[kScratchMaxDevices] = {};
}  // namespace

size_t scratch_bytes() { return kScratchBytes; }

int scratch_init(rjk_ctx* ctx) {
    if (ctx->device < 0 || ctx->device >= kScratchMaxDevices)
        return fail(ctx, RJK_ERR_ARG, "no-alloc scratch: device index beyond the per-device table");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_dev[ctx->device]) {
        void* ptr = nullptr;
        hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !ptr) {
            (void)hipGetLastError();
            return fail(ctx, RJK_ERR_HIP, std::string("no-alloc scratch: hipGetSymbolAddress failed: ") +
                                              hipGetErrorString(e));
        }
        g_scratch_dev[ctx->device] = ptr;
    }
    ctx->scratch = g_scratch_dev[ctx->device];
    return RJK_OK;
}

}  // namespace rjk

extern "C" int rjk_vj_upper_dev_noalloc(rjk_ctx* ctx, const double* dm_dev, size_t spin_channel,
                                        double scaling_factor, int prec, int full_dm, double* vj_dev) {
    using namespace rjk;
    if (!ctx) return RJK_ERR_ARG;
    if (!dm_dev || !vj_dev) return fail(ctx, RJK_ERR_ARG, "vj (no-alloc): NULL device pointer");
    if (spin_channel != 1 && spin_channel != 2)
        return fail(ctx, RJK_ERR_ARG, "spin_channel must be 1 or 2, got " + std::to_string(spin_channel));
    // the size limit does not depend on which B copy the call ends up reading
    const int ns = (int)spin_channel;
    const size_t need = std::max(j_ws_bytes(ctx, ns, RJK_PREC_FP64), j_ws_bytes(ctx, ns, RJK_PREC_FP32));
    if (need > kScratchBytes)
        return fail(ctx, RJK_ERR_ARG, "vj (no-alloc): the J workspace of this shape, " + std::to_string(need) +
                                          " bytes, exceeds the fixed scratch of " + std::to_string(kScratchBytes) +
                                          " bytes; use rjk_vj_upper_dev");
    const void* b = nullptr;
    int jprec = prec;
    int rc = j_source(ctx, spin_channel, prec, &b, &jprec);
    if (rc) return rc;
    if (!ctx->scratch) return fail(ctx, RJK_ERR_HIP, "vj (no-alloc): scratch address unavailable on this device");
    RJK_HIP(ctx, hipSetDevice(ctx->device));
    return j_run(ctx, b, jprec, dm_dev, ns, scaling_factor, full_dm, vj_dev, ctx->scratch, false);
}

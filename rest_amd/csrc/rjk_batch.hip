// rjk_batch.hip — J and K of many independent RI problems (water clusters of
// uneven size) in grouped, persistent launches replayed from a hipGraph
// (include/rest_batch.h).
//
// Every heavy kernel runs over one flattened work list that spans all
// problems: a work item's problem is found by binary search in a prefix table,
// and a fixed grid of work-groups (a multiple of 8, a few per CU) walks the
// list with stride gridDim.x.  Each problem's item count is a multiple of 8,
// so item i always runs on XCD i % 8 and the XCD-aware item order of the
// single-problem kernels carries over: all tiles of one auxiliary column (K
// half-transform) or of one SYRK split share one XCD's L2.
//
// Per problem (RHF, occupation weights occ_c):
//   w_p   = (2 − δ_ij) D_ij               (upper triangle of D, packed)
//   γ_P   = Σ_p B_pP w_p,  J_p = s_J Σ_P B_pP γ_P      (vj_upper_with_rimatr_v_sync)
//   M'_P  = B_P C',  C' = C diag(sqrt occ),  K = s_K Σ_P M'_P M'_Pᵀ (vk_upper_with_ri_v_sync)
// with the K tiles of rjk_ktiles.hpp.  B is stored in the batch's precision
// (FP64, or FP32 with FP64 J accumulation and FP32 MFMA K as in rjk_k.hip).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <string>
#include <vector>

#include "../../include/rest_batch.h"
#include "rjk_ktiles.hpp"

namespace rjk {
namespace {

using namespace kt;

struct BProb {
    const void* b;        // packed B [naux][ldb], batch precision
    size_t ldb, npair;
    int nbf, naux, nbf_pad, nk, nocc, ncb, nsplit, pps;
    const double* dm;     // [nbf][nbf] column-major
    const double* c;      // [nbf][nocc] column-major
    const double* occ;    // [nocc]
    double* w;            // [ldb], padding zero
    double* gamma;        // [naux]
    double* vj;           // [npair]
    double* vk;           // [npair]
    void* cocc;           // [nbf_pad][nk]
    void* sign;           // [nk]
    void* m;              // [naux][nk][nbf_pad]
    double* kpart;        // [nsplit][nbf_pad][nbf_pad]
};

// Prefix tables of the work lists (nprob + 1 entries each, device).
struct Tables {
    const size_t* pairs;   // Σ npair
    const size_t* cols;    // Σ naux
    const size_t* jseg;    // Σ J-accumulation segments
    const size_t* prep;    // Σ nbf_pad·nk
    const size_t* syrk;    // Σ nsplit·ntile2
};

__device__ inline int find_prob(const size_t* __restrict__ off, int n, size_t x) {
    int lo = 0, hi = n - 1;   // largest i with off[i] <= x
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (off[mid] <= x) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

__device__ inline void pair_ij(size_t p, size_t* i, size_t* j) {
    size_t jj = (size_t)((sqrt(8.0 * (double)p + 1.0) - 1.0) * 0.5);
    while (tri(jj) > p) --jj;
    while (tri(jj + 1) <= p) ++jj;
    *j = jj;
    *i = p - tri(jj);
}

__device__ inline double block_sum256(double v, double* sh) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
    __syncthreads();
    const double s = sh[0] + sh[1] + sh[2] + sh[3];
    __syncthreads();
    return s;
}

// ---- J -----------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_b_weights(const BProb* __restrict__ pr, int np, Tables tb) {
    const size_t total = tb.pairs[np];
    for (size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x; t < total; t += (size_t)gridDim.x * blockDim.x) {
        const int q = find_prob(tb.pairs, np, t);
        const BProb& P = pr[q];
        const size_t p = t - tb.pairs[q];
        size_t i, j;
        pair_ij(p, &i, &j);
        const double d = P.dm[i + j * (size_t)P.nbf];   // upper triangle (row <= column)
        P.w[p] = i == j ? d : 2.0 * d;
    }
}

template <typename T>
struct Vec16;
template <> struct Vec16<double> { typedef double v __attribute__((ext_vector_type(2))); static constexpr int n = 2; };
template <> struct Vec16<float> { typedef float v __attribute__((ext_vector_type(4))); static constexpr int n = 4; };

// γ_P for one (problem, auxiliary column) per work item
template <typename T>
__global__ __launch_bounds__(256) void k_b_gamma(const BProb* __restrict__ pr, int np, Tables tb) {
    using V = typename Vec16<T>::v;
    constexpr int NV = Vec16<T>::n;
    __shared__ double sh[4];
    const size_t total = tb.cols[np];
    for (size_t it = blockIdx.x; it < total; it += gridDim.x) {
        const int q = find_prob(tb.cols, np, it);
        const BProb& P = pr[q];
        const size_t col = it - tb.cols[q];
        const T* __restrict__ bc = (const T*)P.b + col * P.ldb;
        double acc = 0.0;
        for (size_t p = (size_t)threadIdx.x * NV; p < P.ldb; p += 256 * NV) {
            const V bv = *(const V*)(bc + p);
#pragma unroll
            for (int e = 0; e < NV; ++e) acc += (double)bv[e] * P.w[p + e];
        }
        const double s = block_sum256(acc, sh);
        if (threadIdx.x == 0) P.gamma[col] = s;
    }
}

// J_p = s Σ_P B_pP γ_P for one (problem, segment of 256·NV pairs) per work item
template <typename T>
__global__ __launch_bounds__(256) void k_b_jacc(const BProb* __restrict__ pr, int np, Tables tb, double scale) {
    using V = typename Vec16<T>::v;
    constexpr int NV = Vec16<T>::n;
    const size_t total = tb.jseg[np];
    for (size_t it = blockIdx.x; it < total; it += gridDim.x) {
        const int q = find_prob(tb.jseg, np, it);
        const BProb& P = pr[q];
        const size_t p0 = (it - tb.jseg[q]) * (256 * NV) + (size_t)threadIdx.x * NV;
        if (p0 >= P.ldb) continue;
        double acc[NV];
#pragma unroll
        for (int e = 0; e < NV; ++e) acc[e] = 0.0;
        const T* __restrict__ bp = (const T*)P.b + p0;
        int col = 0;
        for (; col + 4 <= P.naux; col += 4) {
            V v[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) v[u] = *(const V*)(bp + (size_t)(col + u) * P.ldb);
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const double g = P.gamma[col + u];
#pragma unroll
                for (int e = 0; e < NV; ++e) acc[e] += (double)v[u][e] * g;
            }
        }
        for (; col < P.naux; ++col) {
            const V v = *(const V*)(bp + (size_t)col * P.ldb);
            const double g = P.gamma[col];
#pragma unroll
            for (int e = 0; e < NV; ++e) acc[e] += (double)v[e] * g;
        }
#pragma unroll
        for (int e = 0; e < NV; ++e)
            if (p0 + e < P.npair) P.vj[p0 + e] = scale * acc[e];
    }
}

// ---- K -----------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(256) void k_b_prep(const BProb* __restrict__ pr, int np, Tables tb) {
    const size_t total = tb.prep[np];
    for (size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x; t < total; t += (size_t)gridDim.x * blockDim.x) {
        const int q = find_prob(tb.prep, np, t);
        const BProb& P = pr[q];
        const size_t e = t - tb.prep[q];
        const int c = (int)(e % P.nk), nu = (int)(e / P.nk);
        double v = 0.0;
        if (c < P.nocc && nu < P.nbf) v = P.c[(size_t)c * P.nbf + nu] * sqrt(fabs(P.occ[c]));
        ((T*)P.cocc)[e] = (T)v;
        if (nu == 0) {
            const double o = c < P.nocc ? P.occ[c] : 0.0;
            ((T*)P.sign)[c] = (T)(o > 0 ? 1.0 : (o < 0 ? -1.0 : 0.0));
        }
    }
}

// half-transform over the problems of one tile variant: items of problem j
// (subset order) start at off[j]; round_up(naux, 8) · ntm · ntc items each
template <typename T, int FM, int FN>
__global__ __launch_bounds__(kBlock) void k_b_khalf(const BProb* __restrict__ pr, const int* __restrict__ sub,
                                                    const size_t* __restrict__ off, int nsub) {
    constexpr int TM = 2 * FM * 16, TN = 2 * FN * 16;
    const size_t total = off[nsub];
    for (size_t it = blockIdx.x; it < total; it += gridDim.x) {
        const int j = find_prob(off, nsub, it);
        const BProb& P = pr[sub[j]];
        const size_t local = it - off[j];
        const int ntm = P.nbf_pad / TM, ntc = P.nk / TN, per = ntm * ntc;
        const size_t xcd = local % kXcd, kq = local / kXcd;
        const int Pl = (int)((kq / per) * kXcd + xcd);
        if (Pl >= P.naux) continue;   // padding items (uniform per work-group)
        const int rem = (int)(kq % per), mt = rem / ntc, ct = rem % ntc;
        khalf_tile<T, FM, FN>((const T*)P.b + (size_t)Pl * P.ldb, P.nbf, mt * TM, ct * TN, (const T*)P.cocc, P.nk,
                              P.nbf_pad, (T*)P.m + (size_t)Pl * P.nk * P.nbf_pad);
    }
}

template <typename T, int F>
__global__ __launch_bounds__(kBlock) void k_b_ksyrk(const BProb* __restrict__ pr, int np, Tables tb) {
    constexpr int Tt = 2 * F * 16;
    const size_t total = tb.syrk[np];
    for (size_t it = blockIdx.x; it < total; it += gridDim.x) {
        const int q = find_prob(tb.syrk, np, it);
        const BProb& P = pr[q];
        const size_t local = it - tb.syrk[q];
        const size_t nt = (size_t)P.nbf_pad / Tt, nt2 = nt * (nt + 1) / 2;
        const size_t xcd = local % kXcd, kq = local / kXcd;
        const size_t t = kq % nt2;
        const int split = (int)((kq / nt2) * kXcd + xcd);
        size_t tm, tn;
        upper_tile(t, &tm, &tn);
        const int Pb = split * P.pps, Pe = min(P.naux, Pb + P.pps);
        ksyrk_tile<T, F>((const T*)P.m, P.nbf_pad, P.nk, (int)tm * Tt, (int)tn * Tt, 0, P.ncb, Pb, Pe,
                         (const T*)P.sign, 0, P.kpart + (size_t)split * P.nbf_pad * P.nbf_pad);
    }
}

__global__ __launch_bounds__(256) void k_b_kreduce(const BProb* __restrict__ pr, int np, Tables tb, double scale) {
    const size_t total = tb.pairs[np];
    for (size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x; t < total; t += (size_t)gridDim.x * blockDim.x) {
        const int q = find_prob(tb.pairs, np, t);
        const BProb& P = pr[q];
        const size_t p = t - tb.pairs[q];
        size_t i, j;
        pair_ij(p, &i, &j);
        const size_t slot = (size_t)P.nbf_pad * P.nbf_pad;
        double acc = 0.0;
        for (int s = 0; s < P.nsplit; ++s) acc += P.kpart[s * slot + j * P.nbf_pad + i];
        P.vk[p] = scale * acc;
    }
}

size_t rup(size_t x, size_t a) { return (x + a - 1) / a * a; }
double now_s() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

}  // namespace
}  // namespace rjk

using namespace rjk;

struct rjk_batch {
    int device = 0, prec = RJK_PREC_FP64;
    hipStream_t own = nullptr, stream = nullptr;
    std::vector<BProb> h;            // host copy of the problem table
    BProb* d_prob = nullptr;
    size_t* d_tab = nullptr;         // the five prefix tables back to back
    int* d_sub = nullptr;            // [2][nprob] problem lists of the two half-transform variants
    size_t* d_suboff = nullptr;      // [2][nprob + 1]
    std::vector<size_t> t_pairs, t_cols, t_jseg, t_prep, t_syrk;
    std::vector<void*> allocs;
    double* scratch = nullptr;       // FP64 staging for uploads / the synthetic fill
    size_t scratch_elems = 0;
    size_t bytes_b = 0;
    // launch configuration (autotuned)
    int khalf_cut = 256;             // nbf_pad below this uses the 64×64 half-transform tile
    int waves = 2;                   // persistent work-groups per CU for the tile kernels
    int ncu = 256;
    int t_sub_n[2] = {0, 0};
    size_t t_sub_items[2] = {0, 0};
    double sj = 1.0, sk = 1.0;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    std::string err;
};

namespace {

int bfail(rjk_batch* b, int code, const std::string& m) {
    if (b) b->err = m;
    return code;
}
#define BHIP(bt, call)                                                                                 \
    do {                                                                                               \
        hipError_t e_ = (call);                                                                        \
        if (e_ != hipSuccess) return bfail((bt), RJK_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
    } while (0)

void drop_graph(rjk_batch* b) {
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    b->exec = nullptr;
    b->graph = nullptr;
}

Tables tables(const rjk_batch* b) {
    const size_t n1 = b->h.size() + 1;
    Tables t;
    t.pairs = b->d_tab;
    t.cols = b->d_tab + n1;
    t.jseg = b->d_tab + 2 * n1;
    t.prep = b->d_tab + 3 * n1;
    t.syrk = b->d_tab + 4 * n1;
    return t;
}

// half-transform subsets for the current cut: 0 = 64×64 tiles (small nbf), 1 = 128×64
int build_subsets(rjk_batch* b) {
    const int np = (int)b->h.size();
    std::vector<int> sub[2];
    std::vector<size_t> off[2];
    for (int v = 0; v < 2; ++v) off[v].push_back(0);
    for (int q = 0; q < np; ++q) {
        const BProb& P = b->h[q];
        const int v = P.nbf_pad < b->khalf_cut ? 0 : 1;
        const int TM = v == 0 ? 64 : 128, TN = 64;
        sub[v].push_back(q);
        off[v].push_back(off[v].back() + rup(P.naux, 8) * (size_t)(P.nbf_pad / TM) * (P.nk / TN));
    }
    std::vector<int> hs(2 * np, 0);
    std::vector<size_t> ho(2 * (np + 1), 0);
    for (int v = 0; v < 2; ++v) {
        std::copy(sub[v].begin(), sub[v].end(), hs.begin() + v * np);
        std::copy(off[v].begin(), off[v].end(), ho.begin() + v * (np + 1));
        // entries past the subset repeat the last offset (never searched: nsub bounds the search)
    }
    b->t_sub_n[0] = (int)sub[0].size();
    b->t_sub_n[1] = (int)sub[1].size();
    BHIP(b, hipMemcpyAsync(b->d_sub, hs.data(), hs.size() * sizeof(int), hipMemcpyHostToDevice, b->stream));
    BHIP(b, hipMemcpyAsync(b->d_suboff, ho.data(), ho.size() * sizeof(size_t), hipMemcpyHostToDevice, b->stream));
    b->t_sub_items[0] = off[0].back();
    b->t_sub_items[1] = off[1].back();
    BHIP(b, hipStreamSynchronize(b->stream));   // host vectors die here
    return RJK_OK;
}

}  // namespace

namespace {

unsigned grid_elem(size_t n) { return (unsigned)std::max<size_t>(1, std::min<size_t>((n + 255) / 256, 8192)); }
unsigned grid_items(const rjk_batch* b, size_t items, int per_cu) {
    size_t g = std::min<size_t>(std::max<size_t>(items, 1), (size_t)b->ncu * per_cu);
    return (unsigned)rup(g, 8);
}

template <typename T>
int enqueue_t(rjk_batch* b) {
    const int np = (int)b->h.size();
    const Tables tb = tables(b);
    const hipStream_t st = b->stream;
    hipLaunchKernelGGL(k_b_weights, dim3(grid_elem(b->t_pairs.back())), dim3(256), 0, st, b->d_prob, np, tb);
    hipLaunchKernelGGL(k_b_gamma<T>, dim3(grid_items(b, b->t_cols.back(), 8)), dim3(256), 0, st, b->d_prob, np, tb);
    hipLaunchKernelGGL(k_b_jacc<T>, dim3(grid_items(b, b->t_jseg.back(), 8)), dim3(256), 0, st, b->d_prob, np, tb,
                       b->sj);
    hipLaunchKernelGGL(k_b_prep<T>, dim3(grid_elem(b->t_prep.back())), dim3(256), 0, st, b->d_prob, np, tb);
    if (b->t_sub_items[0])
        hipLaunchKernelGGL((k_b_khalf<T, 2, 2>), dim3(grid_items(b, b->t_sub_items[0], b->waves)), dim3(kBlock), 0,
                           st, b->d_prob, b->d_sub, b->d_suboff, b->t_sub_n[0]);
    if (b->t_sub_items[1])
        hipLaunchKernelGGL((k_b_khalf<T, 4, 2>), dim3(grid_items(b, b->t_sub_items[1], b->waves)), dim3(kBlock), 0,
                           st, b->d_prob, b->d_sub + np, b->d_suboff + (np + 1), b->t_sub_n[1]);
    hipLaunchKernelGGL((k_b_ksyrk<T, 4>), dim3(grid_items(b, b->t_syrk.back(), b->waves)), dim3(kBlock), 0, st,
                       b->d_prob, np, tb);
    hipLaunchKernelGGL(k_b_kreduce, dim3(grid_elem(b->t_pairs.back())), dim3(256), 0, st, b->d_prob, np, tb, b->sk);
    BHIP(b, hipGetLastError());
    return RJK_OK;
}

int enqueue(rjk_batch* b) { return b->prec == RJK_PREC_FP64 ? enqueue_t<double>(b) : enqueue_t<float>(b); }

int capture(rjk_batch* b) {
    drop_graph(b);
    BHIP(b, hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
    const int rc = enqueue(b);
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(b->stream, &g);
    if (rc) {
        if (g) (void)hipGraphDestroy(g);
        return rc;
    }
    if (e != hipSuccess) return bfail(b, RJK_ERR_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
    b->graph = g;
    BHIP(b, hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
    return RJK_OK;
}

template <typename F>
int alloc_slots(rjk_batch* b, size_t elem, F count, std::vector<size_t>* offs, void** base) {
    offs->clear();
    size_t tot = 0;
    for (const BProb& P : b->h) {
        offs->push_back(tot);
        tot += rup(count(P) * elem, 256);
    }
    void* p = nullptr;
    if (hipMalloc(&p, std::max<size_t>(tot, 256)) != hipSuccess) {
        (void)hipGetLastError();
        return bfail(b, RJK_ERR_NOMEM, "rjk_batch_create: device allocation of " + std::to_string(tot) + " bytes failed");
    }
    b->allocs.push_back(p);
    *base = p;
    return RJK_OK;
}

}  // namespace

extern "C" {

int rjk_batch_create(int device, size_t nprob, const size_t* nbf, const size_t* naux, const size_t* nocc, int prec,
                     rjk_batch** out) {
    if (!out || !nbf || !naux || !nocc || nprob == 0 || nprob > (1u << 20)) return RJK_ERR_ARG;
    *out = nullptr;
    if (prec != RJK_PREC_FP64 && prec != RJK_PREC_FP32) return RJK_ERR_ARG;
    for (size_t i = 0; i < nprob; ++i)
        if (nbf[i] == 0 || naux[i] == 0 || nocc[i] == 0 || nocc[i] > nbf[i] || nbf[i] > 32768 || naux[i] > (1u << 24))
            return RJK_ERR_ARG;
    rjk_batch* b = new rjk_batch();
    b->device = device;
    b->prec = prec;
    auto fail_free = [&](int code) {
        rjk_batch_destroy(b);
        return code;
    };
    if (hipSetDevice(device) != hipSuccess || hipStreamCreate(&b->own) != hipSuccess) return fail_free(RJK_ERR_HIP);
    b->stream = b->own;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) == hipSuccess && prop.multiProcessorCount > 0)
        b->ncu = prop.multiProcessorCount;
    const size_t es = prec == RJK_PREC_FP64 ? 8 : 4;
    const size_t chain = 16384;   // products per FP32 (or FP64) SYRK accumulator before the FP64 partial sum
    b->h.resize(nprob);
    size_t max_compact = 0, max_pad = 0;
    for (size_t i = 0; i < nprob; ++i) {
        BProb& P = b->h[i];
        P = BProb{};
        P.nbf = (int)nbf[i];
        P.naux = (int)naux[i];
        P.nocc = (int)nocc[i];
        P.npair = nbf[i] * (nbf[i] + 1) / 2;
        P.ldb = rup(P.npair, kLdAlign);
        P.nbf_pad = (int)rup(nbf[i], kTileMN);
        P.ncb = (int)(rup(nocc[i], kTileK) / kTileK);
        P.nk = (int)rup((size_t)P.ncb * kTileK, kColPad);
        P.pps = (int)std::max<size_t>(1, chain / P.nk);
        P.nsplit = (int)rup((naux[i] + P.pps - 1) / P.pps, 8);
        P.pps = (int)((naux[i] + P.nsplit - 1) / P.nsplit);
        max_compact = std::max(max_compact, P.npair * naux[i]);
        max_pad = std::max(max_pad, P.ldb * naux[i]);
        b->bytes_b += P.ldb * naux[i] * es;
    }
    std::vector<size_t> o;
    void* base = nullptr;
    int rc = RJK_OK;
#define SLOTS(elem, expr, field, type)                                                              \
    rc = alloc_slots(b, elem, [&](const BProb& P) { return (size_t)(expr); }, &o, &base);          \
    if (rc) return fail_free(rc);                                                                    \
    for (size_t i = 0; i < nprob; ++i) b->h[i].field = (type)((char*)base + o[i]);
    SLOTS(es, P.ldb * P.naux, b, const void*)
    SLOTS(8, P.nbf * (size_t)P.nbf, dm, const double*)
    SLOTS(8, P.nbf * (size_t)P.nocc, c, const double*)
    SLOTS(8, P.nocc, occ, const double*)
    SLOTS(8, P.ldb, w, double*)
    SLOTS(8, P.naux, gamma, double*)
    SLOTS(8, P.npair, vj, double*)
    SLOTS(8, P.npair, vk, double*)
    SLOTS(es, (size_t)P.nbf_pad * P.nk, cocc, void*)
    SLOTS(es, P.nk, sign, void*)
    SLOTS(es, (size_t)P.naux * P.nk * P.nbf_pad, m, void*)
    SLOTS(8, (size_t)P.nsplit * P.nbf_pad * P.nbf_pad, kpart, double*)
#undef SLOTS
    // zero the padding of every input/weight slot once (graphs hold no memsets)
    for (void* p : b->allocs) (void)p;
    for (const BProb& P : b->h) {
        if (hipMemsetAsync(P.w, 0, P.ldb * 8, b->stream) != hipSuccess ||
            hipMemsetAsync((void*)P.b, 0, P.ldb * P.naux * es, b->stream) != hipSuccess ||
            hipMemsetAsync((void*)P.dm, 0, (size_t)P.nbf * P.nbf * 8, b->stream) != hipSuccess ||
            hipMemsetAsync((void*)P.c, 0, (size_t)P.nbf * P.nocc * 8, b->stream) != hipSuccess ||
            hipMemsetAsync((void*)P.occ, 0, (size_t)P.nocc * 8, b->stream) != hipSuccess)
            return fail_free(RJK_ERR_HIP);
    }
    // prefix tables
    auto prefix = [&](std::vector<size_t>& t, auto f) {
        t.assign(1, 0);
        for (const BProb& P : b->h) t.push_back(t.back() + f(P));
    };
    prefix(b->t_pairs, [](const BProb& P) { return P.npair; });
    prefix(b->t_cols, [](const BProb& P) { return (size_t)P.naux; });
    prefix(b->t_jseg, [&](const BProb& P) { return (P.ldb + 256 * (16 / es) - 1) / (256 * (16 / es)); });
    prefix(b->t_prep, [](const BProb& P) { return (size_t)P.nbf_pad * P.nk; });
    prefix(b->t_syrk, [](const BProb& P) {
        const size_t nt = (size_t)P.nbf_pad / 128;
        return (size_t)P.nsplit * (nt * (nt + 1) / 2);
    });
    std::vector<size_t> tab;
    for (auto* t : {&b->t_pairs, &b->t_cols, &b->t_jseg, &b->t_prep, &b->t_syrk}) tab.insert(tab.end(), t->begin(), t->end());
    if (hipMalloc((void**)&b->d_prob, nprob * sizeof(BProb)) != hipSuccess ||
        hipMalloc((void**)&b->d_tab, tab.size() * sizeof(size_t)) != hipSuccess ||
        hipMalloc((void**)&b->d_sub, 2 * nprob * sizeof(int)) != hipSuccess ||
        hipMalloc((void**)&b->d_suboff, 2 * (nprob + 1) * sizeof(size_t)) != hipSuccess)
        return fail_free(RJK_ERR_NOMEM);
    if (hipMemcpyAsync(b->d_prob, b->h.data(), nprob * sizeof(BProb), hipMemcpyHostToDevice, b->stream) != hipSuccess ||
        hipMemcpyAsync(b->d_tab, tab.data(), tab.size() * sizeof(size_t), hipMemcpyHostToDevice, b->stream) !=
            hipSuccess)
        return fail_free(RJK_ERR_HIP);
    if (hipStreamSynchronize(b->stream) != hipSuccess) return fail_free(RJK_ERR_HIP);
    if (build_subsets(b)) return fail_free(RJK_ERR_HIP);
    b->scratch_elems = max_compact + (prec == RJK_PREC_FP64 ? 0 : max_pad);
    *out = b;
    return RJK_OK;
}

int rjk_batch_destroy(rjk_batch* b) {
    if (!b) return RJK_OK;
    (void)hipSetDevice(b->device);
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    drop_graph(b);
    for (void* p : b->allocs) (void)hipFree(p);
    for (void* p : {(void*)b->d_prob, (void*)b->d_tab, (void*)b->d_sub, (void*)b->d_suboff, (void*)b->scratch})
        if (p) (void)hipFree(p);
    if (b->own) (void)hipStreamDestroy(b->own);
    delete b;
    return RJK_OK;
}

const char* rjk_batch_last_error(const rjk_batch* b) { return b ? b->err.c_str() : "NULL batch"; }

size_t rjk_batch_size(const rjk_batch* b) { return b ? b->h.size() : 0; }

size_t rjk_batch_bytes_b(const rjk_batch* b) { return b ? b->bytes_b : 0; }

int rjk_batch_set_stream(rjk_batch* b, void* stream) {
    if (!b) return RJK_ERR_ARG;
    b->stream = stream ? (hipStream_t)stream : b->own;
    drop_graph(b);
    return RJK_OK;
}

static int stage_b(rjk_batch* b, size_t i, const double* host_compact, uint64_t seed, bool synth) {
    const BProb& P = b->h[i];
    BHIP(b, hipSetDevice(b->device));
    if (!b->scratch) {
        if (hipMalloc((void**)&b->scratch, std::max<size_t>(b->scratch_elems, 1) * 8) != hipSuccess)
            return bfail(b, RJK_ERR_NOMEM, "rjk_batch: staging allocation failed");
    }
    double* compact = b->scratch;
    double* pad64 = b->prec == RJK_PREC_FP64 ? (double*)P.b : b->scratch + (b->scratch_elems - P.ldb * P.naux);
    if (synth) {
        BHIP(b, launch_fill_synth(b->stream, seed, P.npair, P.naux, P.ldb, pad64));
    } else {
        BHIP(b, hipMemcpyAsync(compact, host_compact, P.npair * P.naux * 8, hipMemcpyHostToDevice, b->stream));
        BHIP(b, launch_pad_rimatr(b->stream, compact, P.npair, P.naux, P.ldb, pad64));
    }
    if (b->prec != RJK_PREC_FP64) BHIP(b, launch_to_fp32(b->stream, pad64, P.ldb * P.naux, (float*)P.b));
    BHIP(b, hipStreamSynchronize(b->stream));
    return RJK_OK;
}

int rjk_batch_set_rimatr(rjk_batch* b, size_t i, const double* rimatr) {
    if (!b || !rimatr || i >= b->h.size()) return b ? bfail(b, RJK_ERR_ARG, "rjk_batch_set_rimatr: bad argument") : 1;
    return stage_b(b, i, rimatr, 0, false);
}

int rjk_batch_fill_synthetic(rjk_batch* b, uint64_t seed) {
    if (!b) return RJK_ERR_ARG;
    for (size_t i = 0; i < b->h.size(); ++i) {
        const int rc = stage_b(b, i, nullptr, seed + i, true);
        if (rc) return rc;
    }
    return RJK_OK;
}

int rjk_batch_set_inputs(rjk_batch* b, size_t i, const double* dm, const double* eigv, const double* occ) {
    if (!b || i >= b->h.size() || !dm || !eigv || !occ)
        return b ? bfail(b, RJK_ERR_ARG, "rjk_batch_set_inputs: bad argument") : RJK_ERR_ARG;
    const BProb& P = b->h[i];
    BHIP(b, hipSetDevice(b->device));
    BHIP(b, hipMemcpyAsync((void*)P.dm, dm, (size_t)P.nbf * P.nbf * 8, hipMemcpyHostToDevice, b->stream));
    BHIP(b, hipMemcpyAsync((void*)P.c, eigv, (size_t)P.nbf * P.nocc * 8, hipMemcpyHostToDevice, b->stream));
    BHIP(b, hipMemcpyAsync((void*)P.occ, occ, (size_t)P.nocc * 8, hipMemcpyHostToDevice, b->stream));
    BHIP(b, hipStreamSynchronize(b->stream));
    return RJK_OK;
}

int rjk_batch_get_outputs(rjk_batch* b, size_t i, double* vj, double* vk) {
    if (!b || i >= b->h.size()) return b ? bfail(b, RJK_ERR_ARG, "rjk_batch_get_outputs: bad argument") : 1;
    const BProb& P = b->h[i];
    BHIP(b, hipSetDevice(b->device));
    if (vj) BHIP(b, hipMemcpyAsync(vj, P.vj, P.npair * 8, hipMemcpyDeviceToHost, b->stream));
    if (vk) BHIP(b, hipMemcpyAsync(vk, P.vk, P.npair * 8, hipMemcpyDeviceToHost, b->stream));
    BHIP(b, hipStreamSynchronize(b->stream));
    return RJK_OK;
}

int rjk_batch_run(rjk_batch* b, double scale_j, double scale_k) {
    if (!b) return RJK_ERR_ARG;
    BHIP(b, hipSetDevice(b->device));
    if (!b->exec || scale_j != b->sj || scale_k != b->sk) {
        b->sj = scale_j;
        b->sk = scale_k;
        const int rc = capture(b);
        if (rc) return rc;
    }
    BHIP(b, hipGraphLaunch(b->exec, b->stream));
    return RJK_OK;
}

int rjk_batch_synchronize(rjk_batch* b) {
    if (!b) return RJK_ERR_ARG;
    BHIP(b, hipStreamSynchronize(b->stream));
    return RJK_OK;
}

int rjk_batch_autotune(rjk_batch* b, int reps, double* best_ms) {
    if (!b || reps < 1) return b ? bfail(b, RJK_ERR_ARG, "rjk_batch_autotune: reps must be >= 1") : 1;
    BHIP(b, hipSetDevice(b->device));
    drop_graph(b);
    const int cuts[] = {0, 256, 512, 1 << 30};
    const int waves[] = {1, 2, 4};
    hipEvent_t e0, e1;
    BHIP(b, hipEventCreate(&e0));
    BHIP(b, hipEventCreate(&e1));
    double best = 1e300;
    int bc = b->khalf_cut, bw = b->waves;
    int rc = RJK_OK;
    for (int cut : cuts) {
        b->khalf_cut = cut;
        rc = build_subsets(b);
        if (rc) break;
        for (int wv : waves) {
            b->waves = wv;
            rc = enqueue(b);   // warm-up of this configuration
            if (rc) break;
            (void)hipEventRecord(e0, b->stream);
            for (int r = 0; r < reps && !rc; ++r) rc = enqueue(b);
            (void)hipEventRecord(e1, b->stream);
            if (rc || hipEventSynchronize(e1) != hipSuccess) {
                rc = rc ? rc : bfail(b, RJK_ERR_HIP, "rjk_batch_autotune: timing failed");
                break;
            }
            float ms = 0.f;
            (void)hipEventElapsedTime(&ms, e0, e1);
            if (ms / reps < best) {
                best = ms / reps;
                bc = cut;
                bw = wv;
            }
        }
        if (rc) break;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    b->khalf_cut = bc;
    b->waves = bw;
    const int rc2 = build_subsets(b);
    if (rc) return rc;
    if (rc2) return rc2;
    if (best_ms) *best_ms = best;
    return RJK_OK;
}

int rjk_batch_config(const rjk_batch* b, int* khalf_cut, int* waves_per_cu) {
    if (!b) return RJK_ERR_ARG;
    if (khalf_cut) *khalf_cut = b->khalf_cut;
    if (waves_per_cu) *waves_per_cu = b->waves;
    return RJK_OK;
}

int rjk_batch_set_config(rjk_batch* b, int khalf_cut, int waves_per_cu) {
    if (!b || waves_per_cu < 1 || waves_per_cu > 16 || khalf_cut < 0)
        return b ? bfail(b, RJK_ERR_ARG, "rjk_batch_set_config: waves_per_cu must be 1..16, cut >= 0") : 1;
    b->khalf_cut = khalf_cut;
    b->waves = waves_per_cu;
    drop_graph(b);
    return build_subsets(b);
}

}  // extern "C"

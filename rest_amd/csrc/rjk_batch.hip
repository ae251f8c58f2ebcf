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
// (FP64, or FP32 with FP64 J accumulation and FP32 MFMA K as in rjk_k.hip, FP64 partial
// sums every kFp32Chain products).
//
// BF16 precisions (RJK_PREC_BF16X1/X2/X3): B is kept once, packed in FP32, and
// J runs on it as at FP32.  The K half-transform reads the packed FP32 B and
// splits every element into S BF16 planes while staging it into LDS (no plane
// copy of B is stored), then runs the split products of rjk_kbf.hip on the
// BF16 matrix cores; M' and the SYRK are those of rjk_kbf.hip per problem.  Both
// stages keep the leading product (plane 0 × plane 0) and the correction
// products in separate FP32 accumulators, added at the end, so the corrections'
// accumulation rounding stays at their own scale.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <string>
#include <vector>

#include "../../include/rest_batch.h"
#include "rjk_bftile.hpp"
#include "rjk_ktiles.hpp"

namespace rjk {
namespace {

using namespace kt;

struct BProb {
    const void* b;        // packed B [naux][ldb], batch precision
    size_t ldb, npair;
    int nbf, naux, nbf_pad, nk, nocc, ncb, nsplit, pps;
    int active;           // 0: every kernel skips this problem (its outputs keep their last values)
    const double* dm;     // [nbf][nbf] column-major
    const double* c;      // [nbf][nocc] column-major
    const double* occ;    // [nocc]
    double* w;            // [ldb], padding zero
    double* gamma;        // [naux]
    double* vj;           // [npair]
    double* vk;           // [npair]
    void* cocc;           // [nbf_pad][nk]; BF16: C'ᵀ planes [S][nk][nbf_pad]
    void* sign;           // [nk]; BF16: sign masks (0x8000 for a negative weight) [nk]
    void* m;              // [naux][nk][nbf_pad]; BF16: M' planes [S][naux][nbf_pad][nk]
    double* kpart;        // [nsplit][nbf_pad][nbf_pad]
};

// Prefix tables of the work lists (nprob + 1 entries each, device).
struct Tables {
    const size_t* pairs;   // Σ npair
    const size_t* cols;    // Σ ceil(naux / kGammaGroup): γ work items
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

// ---- J -----------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_b_weights(const BProb* __restrict__ pr, int np, Tables tb) {
    const size_t total = tb.pairs[np];
    for (size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x; t < total; t += (size_t)gridDim.x * blockDim.x) {
        const int q = find_prob(tb.pairs, np, t);
        const BProb& P = pr[q];
        if (!P.active) continue;
        const size_t p = t - tb.pairs[q];
        size_t i, j;
        pair_ij(p, &i, &j);
        const double d = glb(P.dm)[i + j * (size_t)P.nbf];   // upper triangle (row <= column)
        glb(P.w)[p] = i == j ? d : 2.0 * d;
    }
}

template <typename T>
struct Vec16;
template <> struct Vec16<double> { typedef double v __attribute__((ext_vector_type(2))); static constexpr int n = 2; };
template <> struct Vec16<float> { typedef float v __attribute__((ext_vector_type(4))); static constexpr int n = 4; };

// γ_P for one (problem, group of kGammaGroup auxiliary columns) per work item: each thread loads
// its weights once and applies them to the group's columns of B (streamed with non-temporal loads)
constexpr int kGammaGroup = 8;
template <typename T>
__global__ __launch_bounds__(256) void k_b_gamma(const BProb* __restrict__ pr, int np, Tables tb) {
    using V = typename Vec16<T>::v;
    using W = typename Vec16<double>::v;
    constexpr int NV = Vec16<T>::n, G = kGammaGroup;
    __shared__ double sh[4][G];
    const size_t total = tb.cols[np];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    for (size_t it = blockIdx.x; it < total; it += gridDim.x) {
        const int q = find_prob(tb.cols, np, it);
        const BProb& P = pr[q];
        if (!P.active) continue;
        const int c0 = (int)(it - tb.cols[q]) * G, cnt = min(G, P.naux - c0);
        const RJK_GLB T* __restrict__ bc = glb((const T*)P.b) + (size_t)c0 * P.ldb;
        const RJK_GLB double* __restrict__ w = glb((const double*)P.w);
        double acc[G];
#pragma unroll
        for (int c = 0; c < G; ++c) acc[c] = 0.0;
        for (size_t p = (size_t)threadIdx.x * NV; p < P.ldb; p += 256 * NV) {
            double wl[NV];
#pragma unroll
            for (int h = 0; h < NV / 2; ++h) {
                const W t = *(const RJK_GLB W*)(w + p + 2 * h);
                wl[2 * h] = t[0];
                wl[2 * h + 1] = t[1];
            }
#pragma unroll
            for (int c = 0; c < G; ++c)
                if (c < cnt) {
                    const V bv = __builtin_nontemporal_load((const RJK_GLB V*)(bc + (size_t)c * P.ldb + p));
#pragma unroll
                    for (int e = 0; e < NV; ++e) acc[c] = fma((double)bv[e], wl[e], acc[c]);
                }
        }
#pragma unroll
        for (int c = 0; c < G; ++c) {
            double v = acc[c];
            for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
            if (lane == 0) sh[wave][c] = v;
        }
        __syncthreads();
        if (threadIdx.x < cnt)
            glb(P.gamma)[c0 + threadIdx.x] =
                sh[0][threadIdx.x] + sh[1][threadIdx.x] + sh[2][threadIdx.x] + sh[3][threadIdx.x];
        __syncthreads();
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
        if (!P.active) continue;
        const size_t p0 = (it - tb.jseg[q]) * (256 * NV) + (size_t)threadIdx.x * NV;
        if (p0 >= P.ldb) continue;
        double acc[NV];
#pragma unroll
        for (int e = 0; e < NV; ++e) acc[e] = 0.0;
        const RJK_GLB T* __restrict__ bp = glb((const T*)P.b) + p0;
        const RJK_GLB double* __restrict__ gam = glb((const double*)P.gamma);
        constexpr int U = 8;   // columns whose loads are in flight together
        int col = 0;
        for (; col + U <= P.naux; col += U) {
            V v[U];
#pragma unroll
            for (int u = 0; u < U; ++u)
                v[u] = __builtin_nontemporal_load((const RJK_GLB V*)(bp + (size_t)(col + u) * P.ldb));
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const double g = gam[col + u];
#pragma unroll
                for (int e = 0; e < NV; ++e) acc[e] += (double)v[u][e] * g;
            }
        }
        for (; col < P.naux; ++col) {
            const V v = __builtin_nontemporal_load((const RJK_GLB V*)(bp + (size_t)col * P.ldb));
            const double g = gam[col];
#pragma unroll
            for (int e = 0; e < NV; ++e) acc[e] += (double)v[e] * g;
        }
#pragma unroll
        for (int e = 0; e < NV; ++e)
            if (p0 + e < P.npair) glb(P.vj)[p0 + e] = scale * acc[e];
    }
}

// ---- K -----------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(256) void k_b_prep(const BProb* __restrict__ pr, int np, Tables tb) {
    const size_t total = tb.prep[np];
    for (size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x; t < total; t += (size_t)gridDim.x * blockDim.x) {
        const int q = find_prob(tb.prep, np, t);
        const BProb& P = pr[q];
        if (!P.active) continue;
        const size_t e = t - tb.prep[q];
        const int c = (int)(e % P.nk), nu = (int)(e / P.nk);
        double v = 0.0;
        if (c < P.nocc && nu < P.nbf) v = glb(P.c)[(size_t)c * P.nbf + nu] * sqrt(fabs(glb(P.occ)[c]));
        glb((T*)P.cocc)[e] = (T)v;
        if (nu == 0) {
            const double o = c < P.nocc ? glb(P.occ)[c] : 0.0;
            glb((T*)P.sign)[c] = (T)(o > 0 ? 1.0 : (o < 0 ? -1.0 : 0.0));
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
        if (!P.active) continue;
        const size_t local = it - off[j];
        const int ntm = P.nbf_pad / TM, ntc = P.nk / TN, per = ntm * ntc;
        const size_t xcd = local % kXcd, kq = local / kXcd;
        const int Pl = (int)((kq / per) * kXcd + xcd);
        if (Pl >= P.naux) continue;   // padding items (uniform per work-group)
        const int rem = (int)(kq % per), mt = rem / ntc, ct = rem % ntc;
        khalf_tile<T, FM, FN>(glb((const T*)P.b) + (size_t)Pl * P.ldb, P.nbf, mt * TM, ct * TN, glb((const T*)P.cocc),
                              P.nk, P.nbf_pad, glb((T*)P.m) + (size_t)Pl * P.nk * P.nbf_pad);
    }
}

template <typename T, int F>
__global__ __launch_bounds__(kBlock) void k_b_ksyrk(const BProb* __restrict__ pr, int np, Tables tb) {
    constexpr int Tt = 2 * F * 16;
    const size_t total = tb.syrk[np];
    for (size_t it = blockIdx.x; it < total; it += gridDim.x) {
        const int q = find_prob(tb.syrk, np, it);
        const BProb& P = pr[q];
        if (!P.active) continue;
        const size_t local = it - tb.syrk[q];
        const size_t nt = (size_t)P.nbf_pad / Tt, nt2 = nt * (nt + 1) / 2;
        const size_t xcd = local % kXcd, kq = local / kXcd;
        const size_t t = kq % nt2;
        const int split = (int)((kq / nt2) * kXcd + xcd);
        size_t tm, tn;
        upper_tile(t, &tm, &tn);
        const int Pb = split * P.pps, Pe = min(P.naux, Pb + P.pps);
        ksyrk_tile<T, F>(glb((const T*)P.m), P.nbf_pad, P.nk, (int)tm * Tt, (int)tn * Tt, 0, P.ncb, Pb, Pe,
                         glb((const T*)P.sign), 0, glb(P.kpart) + (size_t)split * P.nbf_pad * P.nbf_pad);
    }
}

__global__ __launch_bounds__(256) void k_b_kreduce(const BProb* __restrict__ pr, int np, Tables tb, double scale) {
    const size_t total = tb.pairs[np];
    for (size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x; t < total; t += (size_t)gridDim.x * blockDim.x) {
        const int q = find_prob(tb.pairs, np, t);
        const BProb& P = pr[q];
        if (!P.active) continue;
        const size_t p = t - tb.pairs[q];
        size_t i, j;
        pair_ij(p, &i, &j);
        const size_t slot = (size_t)P.nbf_pad * P.nbf_pad;
        double acc = 0.0;
        for (int s = 0; s < P.nsplit; ++s) acc += glb((const double*)P.kpart)[s * slot + j * P.nbf_pad + i];
        glb(P.vk)[p] = scale * acc;
    }
}

// ---- K on split BF16 planes -----------------------------------------------------
template <int S>
__device__ inline void split_to(double x, __bf16* out) {
    double r = x;
#pragma unroll
    for (int k = 0; k < S; ++k) {
        const __bf16 h = (__bf16)(float)r;
        out[k] = h;
        r -= (double)(float)h;
    }
}

template <int S>
__device__ inline void split_to_f(float x, __bf16* out) {
    float r = x;
#pragma unroll
    for (int k = 0; k < S; ++k) {
        const __bf16 h = (__bf16)r;
        out[k] = h;
        r -= (float)h;
    }
}

// C'ᵀ planes (column c of C diag(sqrt|occ|) as a row, ν fastest) and sign masks
template <int S>
__global__ __launch_bounds__(256) void k_b_bf_prep(const BProb* __restrict__ pr, int np, Tables tb) {
    const size_t total = tb.prep[np];
    for (size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x; t < total; t += (size_t)gridDim.x * blockDim.x) {
        const int q = find_prob(tb.prep, np, t);
        const BProb& P = pr[q];
        if (!P.active) continue;
        const size_t e = t - tb.prep[q], plane = (size_t)P.nk * P.nbf_pad;
        const int c = (int)(e / P.nbf_pad), nu = (int)(e % P.nbf_pad);
        const double w = c < P.nocc ? glb(P.occ)[c] : 0.0;
        const double v = (c < P.nocc && nu < P.nbf) ? glb(P.c)[(size_t)c * P.nbf + nu] * sqrt(fabs(w)) : 0.0;
        __bf16 h[S];
        split_to<S>(v, h);
#pragma unroll
        for (int k = 0; k < S; ++k) glb((__bf16*)P.cocc)[k * plane + e] = h[k];
        if (nu == 0) glb((unsigned short*)P.sign)[c] = w < 0 ? 0x8000 : 0;
    }
}

// Half-transform tile D[c][μ] = Σ_ν C'ᵀ[c][ν] B[μ][ν] of TN c × TM μ on 2 × 2 waves, k-blocks of TK
// ν.  B comes from the packed FP32 column (element (i<=j) at j(j+1)/2 + i); each thread stages
// runs of 8 consecutive ν of one μ: along ν (contiguous in memory) when the tile lies wholly below
// the diagonal, else one μ per lane (contiguous across lanes above the diagonal).  LDS rows are TK
// BF16 values, 16-byte piece p of row r at p ^ sw(r) (the swizzle of rjk_bftile.hpp).
template <int S, int TM, int TK, int TN = 64>
struct BfHalf {
    static constexpr int NT = 256, WX = 2, WY = 2, WGX = TN, WGY = TM, NPC = TK / 8;
    static constexpr int FX = WGX / WX / 16, FY = WGY / WY / 16;
    static constexpr int LX = WGX * NPC / NT, LY = WGY * NPC / NT;
    static constexpr int LDS_ELEMS = S * (WGX + WGY) * TK;
    static constexpr int S_ = S, TK_ = TK;
    static_assert(LX >= 1 && LY >= 1 && LX * NT == WGX * NPC && LY * NT == WGY * NPC, "load shape");
    __device__ static constexpr int sw(int r) { return NPC == 4 ? ((r >> 1) & 3) : ((r >> 1) & 7); }
};

typedef float F4a __attribute__((ext_vector_type(4), aligned(4)));   // 4 floats at 4-byte alignment

template <class H>
__device__ __forceinline__ void bf_half_tile(__bf16* __restrict__ lds, const RJK_GLB float* __restrict__ bcol, int nbf,
                                             int mu0, int c0, const RJK_GLB __bf16* __restrict__ ct, size_t pct,
                                             int nbf_pad, int nk, RJK_GLB __bf16* __restrict__ mP, size_t pm) {
    using namespace bft;
    constexpr int S = H::S_, NP = npairs<S>(), TK = H::TK_, NPC = H::NPC;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wx = wave % H::WX, wy = wave / H::WX, l16 = lane & 15;
    auto at = [](int r, int p) { return r * TK + ((p ^ H::sw(r)) << 3); };
    auto Xs = [&](int k) { return lds + (size_t)k * (H::WGX + H::WGY) * TK; };
    auto Ys = [&](int k) { return lds + ((size_t)k * (H::WGX + H::WGY) + H::WGX) * TK; };
    f4 acc[H::FX][H::FY], acc2[H::FX][H::FY];   // leading product; corrections (added at the end)
#pragma unroll
    for (int i = 0; i < H::FX; ++i)
#pragma unroll
        for (int j = 0; j < H::FY; ++j) {
            acc[i][j] = f4{0.f, 0.f, 0.f, 0.f};
            acc2[i][j] = f4{0.f, 0.f, 0.f, 0.f};
        }
    u4 rx[S][H::LX];
    float rb[H::LY][8];
    bool below = false;
    auto task_rp = [&](int j, int* r, int* p) {   // row (μ offset) and piece of B task j
        const int e = tid + j * H::NT;
        if (below) {
            *r = e / NPC;
            *p = e % NPC;
        } else {
            *r = e % H::WGY;
            *p = e / H::WGY;
        }
    };
    auto fetch = [&](int nu0) {
#pragma unroll
        for (int k = 0; k < S; ++k)
#pragma unroll
            for (int j = 0; j < H::LX; ++j) {
                const int e = tid + j * H::NT;
                rx[k][j] = ld16(ct + k * pct + (size_t)(c0 + e / NPC) * nbf_pad + nu0 + (e % NPC) * 8);
            }
        below = mu0 >= nu0 + TK - 1;
#pragma unroll
        for (int j = 0; j < H::LY; ++j) {
            int r, p;
            task_rp(j, &r, &p);
            const int mu = mu0 + r, nub = nu0 + p * 8;
            if (below) {
                // 8 consecutive ν of row μ: two 16-byte loads at the row's 4-byte alignment
                const size_t base = mu < nbf ? tri((size_t)mu) + nub : 0;
                const RJK_GLB F4a* src = (const RJK_GLB F4a*)(bcol + base);
                const F4a lo = src[0], hi = src[1];
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    rb[j][i] = mu < nbf ? lo[i] : 0.f;
                    rb[j][4 + i] = mu < nbf ? hi[i] : 0.f;
                }
            } else {
#pragma unroll
                for (int i = 0; i < 8; ++i) {
                    const int nu = nub + i;
                    const bool ok = mu < nbf && nu < nbf;
                    const int lo = mu < nu ? mu : nu, hi = mu < nu ? nu : mu;
                    const float v = bcol[ok ? tri((size_t)hi) + lo : 0];
                    rb[j][i] = ok ? v : 0.f;
                }
            }
        }
    };
    auto stash = [&]() {
#pragma unroll
        for (int k = 0; k < S; ++k)
#pragma unroll
            for (int j = 0; j < H::LX; ++j) {
                const int e = tid + j * H::NT;
                *(u4*)(Xs(k) + at(e / NPC, e % NPC)) = rx[k][j];
            }
#pragma unroll
        for (int j = 0; j < H::LY; ++j) {
            int r, p;
            task_rp(j, &r, &p);
            bf8 v[S];
#pragma unroll
            for (int i = 0; i < 8; ++i) {
                __bf16 h[S];
                split_to_f<S>(rb[j][i], h);
#pragma unroll
                for (int k = 0; k < S; ++k) v[k][i] = h[k];
            }
#pragma unroll
            for (int k = 0; k < S; ++k) *(bf8*)(Ys(k) + at(r, p)) = v[k];
        }
    };
    auto compute = [&]() {
#pragma unroll
        for (int ks = 0; ks < TK / KM; ++ks) {
            const int pr = ks * 4 + (lane >> 4);
            bf8 fx[H::FX][S], fy[H::FY][S];
#pragma unroll
            for (int k = 0; k < S; ++k) {
#pragma unroll
                for (int i = 0; i < H::FX; ++i) fx[i][k] = *(const bf8*)(Xs(k) + at((wx * H::FX + i) * 16 + l16, pr));
#pragma unroll
                for (int j = 0; j < H::FY; ++j) fy[j][k] = *(const bf8*)(Ys(k) + at((wy * H::FY + j) * 16 + l16, pr));
            }
#pragma unroll
            for (int p = 0; p < NP; ++p)
#pragma unroll
                for (int i = 0; i < H::FX; ++i)
#pragma unroll
                    for (int j = 0; j < H::FY; ++j) {
                        if (p > 0)
                            acc2[i][j] = mfma(fx[i][pa<S>(p)], fy[j][pb<S>(p)], acc2[i][j]);
                        else
                            acc[i][j] = mfma(fx[i][pa<S>(p)], fy[j][pb<S>(p)], acc[i][j]);
                    }
        }
    };
    const int nq = (nbf + TK - 1) / TK;
    fetch(0);
    for (int q = 0; q < nq; ++q) {
        stash();
        __syncthreads();
        if (q + 1 < nq) fetch((q + 1) * TK);
        __builtin_amdgcn_s_setprio(1);   // MFMA phase first (batch of 512 clusters: 874 -> 853 ms at BF16X3)
        compute();
        __builtin_amdgcn_s_setprio(0);
        __syncthreads();
    }
    // M' planes: lane holds c = c0 + 16(wx FX + i) + 4(l/16) + r for μ = mu0 + 16(wy FY + j) + l%16
#pragma unroll
    for (int i = 0; i < H::FX; ++i)
#pragma unroll
        for (int j = 0; j < H::FY; ++j) {
            const int c = c0 + (wx * H::FX + i) * 16 + 4 * (lane >> 4);
            const int mu = mu0 + (wy * H::FY + j) * 16 + l16;
            __bf16 h[4][S];
#pragma unroll
            for (int r = 0; r < 4; ++r) split_to_f<S>(acc[i][j][r] + acc2[i][j][r], h[r]);
#pragma unroll
            for (int k = 0; k < S; ++k)
                *(RJK_GLB bft::bf4*)(mP + (size_t)k * pm + (size_t)mu * nk + c) = bft::bf4{h[0][k], h[1][k], h[2][k], h[3][k]};
        }
}

// half-transform over the problems of one tile class: round_up(naux, 8) · ntm · ntc items per problem
template <class H>
__global__ __launch_bounds__(256) void k_b_bf_khalf(const BProb* __restrict__ pr, const int* __restrict__ sub,
                                                    const size_t* __restrict__ off, int nsub) {
    __shared__ __bf16 lds[H::LDS_ELEMS];
    const size_t total = off[nsub];
    for (size_t it = blockIdx.x; it < total; it += gridDim.x) {
        const int j = find_prob(off, nsub, it);
        const BProb& P = pr[sub[j]];
        if (!P.active) continue;
        const size_t local = it - off[j];
        const int ntm = P.nbf_pad / H::WGY, ntc = P.nk / H::WGX, per = ntm * ntc;
        const size_t xcd = local % kXcd, kq = local / kXcd;
        const int Pl = (int)((kq / per) * kXcd + xcd);
        if (Pl >= P.naux) continue;   // padding items (uniform per work-group)
        const int rem = (int)(kq % per), mt = rem / ntc, ctile = rem % ntc;
        const size_t pm = (size_t)P.naux * P.nbf_pad * P.nk;
        bf_half_tile<H>(lds, glb((const float*)P.b) + (size_t)Pl * P.ldb, P.nbf, mt * H::WGY, ctile * H::WGX,
                        glb((const __bf16*)P.cocc), (size_t)P.nk * P.nbf_pad, P.nbf_pad, P.nk,
                        glb((__bf16*)P.m) + (size_t)Pl * P.nbf_pad * P.nk, pm);
    }
}

// SYRK over the problems of one tile class: nsplit · ntile2 items per problem (split s on XCD s % 8),
// Kpart[split][ν][μ] = Σ_{P in split, c} sign_c M'_P[ν][c] M'_P[μ][c] on upper tile pairs
template <class C>
__global__ __launch_bounds__(C::NT) void k_b_bf_ksyrk(const BProb* __restrict__ pr, const int* __restrict__ sub,
                                                      const size_t* __restrict__ off, int nsub) {
    __shared__ __bf16 lds[C::LDS_ELEMS];
    const size_t total = off[nsub];
    for (size_t it = blockIdx.x; it < total; it += gridDim.x) {
        const int jj = find_prob(off, nsub, it);
        const BProb& P = pr[sub[jj]];
        if (!P.active) continue;
        const size_t local = it - off[jj];
        const size_t nt = (size_t)P.nbf_pad / C::WGX, nt2 = nt * (nt + 1) / 2;
        const size_t xcd = local % kXcd, kq = local / kXcd;
        const size_t t = kq % nt2;
        const int split = (int)((kq / nt2) * kXcd + xcd);
        size_t tm, tn;
        upper_tile(t, &tm, &tn);
        const int mu0 = (int)tm * C::WGY, nu0 = (int)tn * C::WGX;
        const int ncb = (P.nocc + C::TK - 1) / C::TK;
        const int Pb = split * P.pps, Pe = min(P.naux, Pb + P.pps);
        const int nq = (Pe > Pb ? Pe - Pb : 0) * ncb;
        const size_t mplane = (size_t)P.naux * P.nbf_pad * P.nk, mPs = (size_t)P.nbf_pad * P.nk;
        const RJK_GLB __bf16* mpl = glb((const __bf16*)P.m);
        const RJK_GLB unsigned short* smask = glb((const unsigned short*)P.sign);
        bft::f4 acc[C::FX][C::FY], acc2[C::FX][C::FY];   // leading product; corrections
        bft::nt_tile2<C, true>(
            lds, [&](int q) { return mpl + (Pb + q / ncb) * mPs + (size_t)nu0 * P.nk + (q % ncb) * C::TK; },
            (size_t)P.nk, mplane,
            [&](int q) { return mpl + (Pb + q / ncb) * mPs + (size_t)mu0 * P.nk + (q % ncb) * C::TK; },
            (size_t)P.nk, mplane,
            [&](int q, int pc) {
                const RJK_GLB unsigned short* sm = smask + (q % ncb) * C::TK + pc;
                return bft::u4{(unsigned)sm[0] | ((unsigned)sm[1] << 16), (unsigned)sm[2] | ((unsigned)sm[3] << 16),
                               (unsigned)sm[4] | ((unsigned)sm[5] << 16), (unsigned)sm[6] | ((unsigned)sm[7] << 16)};
            },
            nq, acc, acc2);
        const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, wx = wave % C::WX, wy = wave / C::WX;
        RJK_GLB double* __restrict__ kp = glb(P.kpart) + (size_t)split * P.nbf_pad * P.nbf_pad;
#pragma unroll
        for (int i = 0; i < C::FX; ++i)
#pragma unroll
            for (int j = 0; j < C::FY; ++j) {
                const int nb = nu0 + (wx * C::FX + i) * 16 + 4 * (lane >> 4);
                const int mu = mu0 + (wy * C::FY + j) * 16 + (lane & 15);
#pragma unroll
                for (int r = 0; r < 4; ++r)
                    kp[(size_t)(nb + r) * P.nbf_pad + mu] = (double)acc[i][j][r] + (double)acc2[i][j][r];
            }
    }
}

// B of one problem from a packed FP64 device array [naux][ld] into the batch slot (precision T)
template <typename T>
__global__ __launch_bounds__(256) void k_b_from_dev(const double* __restrict__ src, size_t ld, size_t npair,
                                                    size_t naux, size_t ldb, T* __restrict__ dst) {
    const size_t total = naux * ldb;
    for (size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x; t < total; t += (size_t)gridDim.x * blockDim.x) {
        const size_t P = t / ldb, p = t % ldb;
        dst[t] = p < npair ? (T)src[P * ld + p] : T(0);
    }
}

size_t rup(size_t x, size_t a) { return (x + a - 1) / a * a; }

bool is_bf(int prec) { return prec == RJK_PREC_BF16X1 || prec == RJK_PREC_BF16X2 || prec == RJK_PREC_BF16X3; }
int planes_of(int prec) { return prec == RJK_PREC_BF16X3 ? 3 : (prec == RJK_PREC_BF16X1 ? 1 : 2); }
size_t elem_b(int prec) { return prec == RJK_PREC_FP64 ? 8 : 4; }

// Products per SYRK accumulator before the FP64 split sum: kFp32Chain (as the single-problem K) where
// the accumulator is FP32, 16384 at FP64 (there only the split count, i.e. the parallelism, follows).
size_t batch_chain(int prec) { return prec == RJK_PREC_FP64 ? 16384 : kFp32Chain; }

// Geometry of one problem at a precision (pointers left null).
BProb geometry(size_t nbf, size_t naux, size_t nocc, int prec) {
    BProb P = BProb{};
    P.nbf = (int)nbf;
    P.naux = (int)naux;
    P.nocc = (int)nocc;
    P.npair = nbf * (nbf + 1) / 2;
    P.ldb = rup(P.npair, kLdAlign);
    if (is_bf(prec)) {
        P.nbf_pad = (int)rup(nbf, 64);     // the 64-wide BF16 tiles; 128-wide ones where nbf_pad % 128 == 0
        P.ncb = 0;
        P.nk = (int)rup(nocc, kBfColGran);
    } else {
        P.nbf_pad = (int)rup(nbf, kTileMN);
        P.ncb = (int)(rup(nocc, kTileK) / kTileK);
        P.nk = (int)rup((size_t)P.ncb * kTileK, kColPad);
    }
    P.pps = (int)std::max<size_t>(1, batch_chain(prec) / P.nk);
    P.nsplit = (int)rup((naux + P.pps - 1) / P.pps, 8);
    P.pps = (int)((naux + P.nsplit - 1) / P.nsplit);
    P.active = 1;
    return P;
}

// Device slots of one problem: {element bytes, element count} in allocation order (the order of the
// BProb fields b, dm, c, occ, w, gamma, vj, vk, cocc, sign, m, kpart).
constexpr int kNumSlots = 12;
void slot_sizes(const BProb& P, int prec, size_t (&elem)[kNumSlots], size_t (&count)[kNumSlots]) {
    const size_t es = elem_b(prec), S = is_bf(prec) ? (size_t)planes_of(prec) : 1;
    const size_t ek = is_bf(prec) ? 2 : es;   // C', sign and M' elements
    const size_t nb = P.nbf, pad = P.nbf_pad, nk = P.nk, na = P.naux;
    const size_t e[kNumSlots] = {es, 8, 8, 8, 8, 8, 8, 8, ek, ek, ek, 8};
    const size_t c[kNumSlots] = {P.ldb * na, nb * nb,      nb * P.nocc, (size_t)P.nocc, P.ldb, na,
                                 P.npair,    P.npair,      S * pad * nk, nk,            S * na * nk * pad,
                                 (size_t)P.nsplit * pad * pad};
    for (int k = 0; k < kNumSlots; ++k) {
        elem[k] = e[k];
        count[k] = c[k];
    }
}

size_t prob_bytes(const BProb& P, int prec) {
    size_t e[kNumSlots], c[kNumSlots], tot = 0;
    slot_sizes(P, prec, e, c);
    for (int k = 0; k < kNumSlots; ++k) tot += rup(e[k] * c[k], 256);
    return tot;
}

bool valid_prec(int prec) { return prec == RJK_PREC_FP64 || prec == RJK_PREC_FP32 || is_bf(prec); }

// Half-transform groups: (μ rows, c columns) of the tile.  BF16 uses all five (one c tile covers
// every occupied column up to 192 at 64 rows, 128 at 128 rows, so each B element is staged once);
// FP64/FP32 use 0 and 3 (rjk_ktiles.hpp's 64 × 64 and 128 × 64 tiles).
constexpr int kHalfGroups = 5;
constexpr int kGroupTM[kHalfGroups] = {64, 64, 64, 128, 128};
constexpr int kGroupTN[kHalfGroups] = {64, 128, 192, 64, 128};

bool valid_shape(size_t nbf, size_t naux, size_t nocc) {
    return nbf != 0 && naux != 0 && nocc != 0 && nocc <= nbf && nbf <= 32768 && naux <= (1u << 24);
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
    int* d_sub = nullptr;            // [kHalfGroups][nprob] problem lists of the half-transform groups
    size_t* d_suboff = nullptr;      // [kHalfGroups][nprob + 1] their item offsets
    int* d_ssub = nullptr;           // [2][nprob] problem lists of the two SYRK tile classes (BF16)
    size_t* d_syrkoff = nullptr;     // [2][nprob + 1] their item offsets
    std::vector<size_t> t_pairs, t_cols, t_jseg, t_prep, t_syrk;
    std::vector<void*> allocs;
    double* scratch = nullptr;       // FP64 staging for uploads / the synthetic fill
    size_t scratch_elems = 0;
    size_t bytes_b = 0, bytes_total = 0;
    // launch configuration (autotuned)
    int khalf_cut = 256;             // nbf_pad below this uses the narrow tile class
    int waves = 2;                   // persistent work-groups per CU for the tile kernels
    int ncu = 256;
    int t_sub_n[kHalfGroups] = {};
    size_t t_sub_items[kHalfGroups] = {};
    int t_ssub_n[2] = {0, 0};
    size_t t_syrk_items[2] = {0, 0};
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

// Tile class of a problem under the current cut: 0 = narrow (FP: 64 × 64 half-transform tiles;
// BF16: 64-wide half-transform and SYRK tiles), 1 = wide (128); BF16 needs nbf_pad % 128 == 0 for 1.
int tile_class(const rjk_batch* b, const BProb& P) {
    if (P.nbf_pad < b->khalf_cut) return 0;
    return (is_bf(b->prec) && P.nbf_pad % 128) ? 0 : 1;
}

int half_group(const rjk_batch* b, const BProb& P) {
    const int v = tile_class(b, P);
    if (!is_bf(b->prec)) return v == 0 ? 0 : 3;
    if (v == 0) return P.nk == 128 ? 1 : (P.nk == 192 ? 2 : 0);
    return P.nk == 128 ? 4 : 3;
}

int build_subsets(rjk_batch* b) {
    const int np = (int)b->h.size();
    const bool bf = is_bf(b->prec);
    std::vector<int> sub[kHalfGroups], ssub[2];
    std::vector<size_t> off[kHalfGroups], soff[2];
    for (int g = 0; g < kHalfGroups; ++g) off[g].push_back(0);
    for (int v = 0; v < 2; ++v) soff[v].push_back(0);
    for (int q = 0; q < np; ++q) {
        const BProb& P = b->h[q];
        const int g = half_group(b, P), TM = kGroupTM[g], TN = kGroupTN[g];
        sub[g].push_back(q);
        off[g].push_back(off[g].back() + rup(P.naux, 8) * (size_t)(P.nbf_pad / TM) * (P.nk / TN));
        if (bf) {
            const int v = tile_class(b, P);
            const size_t nt = (size_t)P.nbf_pad / (v == 0 ? 64 : 128);
            ssub[v].push_back(q);
            soff[v].push_back(soff[v].back() + (size_t)P.nsplit * (nt * (nt + 1) / 2));
        }
    }
    // entries past a list's end stay zero (never searched: its length bounds the search)
    std::vector<int> hs(kHalfGroups * np, 0), hss(2 * np, 0);
    std::vector<size_t> ho(kHalfGroups * (np + 1), 0), hso(2 * (np + 1), 0);
    for (int g = 0; g < kHalfGroups; ++g) {
        std::copy(sub[g].begin(), sub[g].end(), hs.begin() + g * np);
        std::copy(off[g].begin(), off[g].end(), ho.begin() + g * (np + 1));
        b->t_sub_n[g] = (int)sub[g].size();
        b->t_sub_items[g] = off[g].back();
    }
    for (int v = 0; v < 2; ++v) {
        std::copy(ssub[v].begin(), ssub[v].end(), hss.begin() + v * np);
        std::copy(soff[v].begin(), soff[v].end(), hso.begin() + v * (np + 1));
        b->t_ssub_n[v] = (int)ssub[v].size();
        b->t_syrk_items[v] = soff[v].back();
    }
    BHIP(b, hipMemcpyAsync(b->d_sub, hs.data(), hs.size() * sizeof(int), hipMemcpyHostToDevice, b->stream));
    BHIP(b, hipMemcpyAsync(b->d_suboff, ho.data(), ho.size() * sizeof(size_t), hipMemcpyHostToDevice, b->stream));
    BHIP(b, hipMemcpyAsync(b->d_ssub, hss.data(), hss.size() * sizeof(int), hipMemcpyHostToDevice, b->stream));
    BHIP(b, hipMemcpyAsync(b->d_syrkoff, hso.data(), hso.size() * sizeof(size_t), hipMemcpyHostToDevice, b->stream));
    BHIP(b, hipStreamSynchronize(b->stream));   // host vectors die here
    return RJK_OK;
}

unsigned grid_elem(size_t n) { return (unsigned)std::max<size_t>(1, std::min<size_t>((n + 255) / 256, 8192)); }
unsigned grid_items(const rjk_batch* b, size_t items, int per_cu) {
    size_t g = std::min<size_t>(std::max<size_t>(items, 1), (size_t)b->ncu * per_cu);
    return (unsigned)rup(g, 8);
}

// Records ev[k] before stage k and ev[kStages] after the last (ev may be null).  Stages: 0 weights,
// 1 γ, 2 J accumulation, 3 C' preparation, 4-5 half-transform (narrow, wide), 6-7 SYRK (narrow /
// all for FP, wide), 8 split reduction.
template <typename T>
int enqueue_fp(rjk_batch* b, hipEvent_t* ev) {
    const int np = (int)b->h.size();
    const Tables tb = tables(b);
    const hipStream_t st = b->stream;
    auto mark = [&](int k) {
        if (ev) (void)hipEventRecord(ev[k], st);
    };
    mark(0);
    hipLaunchKernelGGL(k_b_weights, dim3(grid_elem(b->t_pairs.back())), dim3(256), 0, st, b->d_prob, np, tb);
    mark(1);
    hipLaunchKernelGGL(k_b_gamma<T>, dim3(grid_items(b, b->t_cols.back(), 8)), dim3(256), 0, st, b->d_prob, np, tb);
    mark(2);
    hipLaunchKernelGGL(k_b_jacc<T>, dim3(grid_items(b, b->t_jseg.back(), 8)), dim3(256), 0, st, b->d_prob, np, tb,
                       b->sj);
    mark(3);
    hipLaunchKernelGGL(k_b_prep<T>, dim3(grid_elem(b->t_prep.back())), dim3(256), 0, st, b->d_prob, np, tb);
    mark(4);
    if (b->t_sub_items[0])
        hipLaunchKernelGGL((k_b_khalf<T, 2, 2>), dim3(grid_items(b, b->t_sub_items[0], b->waves)), dim3(kBlock), 0,
                           st, b->d_prob, b->d_sub, b->d_suboff, b->t_sub_n[0]);
    mark(5);
    if (b->t_sub_items[3])
        hipLaunchKernelGGL((k_b_khalf<T, 4, 2>), dim3(grid_items(b, b->t_sub_items[3], b->waves)), dim3(kBlock), 0,
                           st, b->d_prob, b->d_sub + 3 * np, b->d_suboff + 3 * (np + 1), b->t_sub_n[3]);
    mark(6);
    hipLaunchKernelGGL((k_b_ksyrk<T, 4>), dim3(grid_items(b, b->t_syrk.back(), b->waves)), dim3(kBlock), 0, st,
                       b->d_prob, np, tb);
    mark(7);
    mark(8);
    hipLaunchKernelGGL(k_b_kreduce, dim3(grid_elem(b->t_pairs.back())), dim3(256), 0, st, b->d_prob, np, tb, b->sk);
    mark(9);
    BHIP(b, hipGetLastError());
    return RJK_OK;
}

template <int S>
int enqueue_bf(rjk_batch* b, hipEvent_t* ev) {
    constexpr int TK = S == 3 ? 32 : 64;
    using H0 = BfHalf<S, 64, TK, 64>;
    using H1 = BfHalf<S, 64, TK, 128>;
    using H2 = BfHalf<S, 64, TK, 192>;
    using H3 = BfHalf<S, 128, TK, 64>;
    using H4 = BfHalf<S, 128, TK, 128>;
    using Y0 = bft::Cfg<S, 64, 64, 2, 2, false, TK>;
    using Y1 = bft::Cfg<S, 128, 128, 2, 4, false, TK>;
    const int np = (int)b->h.size();
    const Tables tb = tables(b);
    const hipStream_t st = b->stream;
    auto mark = [&](int k) {
        if (ev) (void)hipEventRecord(ev[k], st);
    };
    mark(0);
    hipLaunchKernelGGL(k_b_weights, dim3(grid_elem(b->t_pairs.back())), dim3(256), 0, st, b->d_prob, np, tb);
    mark(1);
    hipLaunchKernelGGL(k_b_gamma<float>, dim3(grid_items(b, b->t_cols.back(), 8)), dim3(256), 0, st, b->d_prob, np,
                       tb);
    mark(2);
    hipLaunchKernelGGL(k_b_jacc<float>, dim3(grid_items(b, b->t_jseg.back(), 8)), dim3(256), 0, st, b->d_prob, np,
                       tb, b->sj);
    mark(3);
    hipLaunchKernelGGL(k_b_bf_prep<S>, dim3(grid_elem(b->t_prep.back())), dim3(256), 0, st, b->d_prob, np, tb);
    mark(4);
    auto half = [&](auto kern, int g) {
        if (b->t_sub_items[g])
            hipLaunchKernelGGL(kern, dim3(grid_items(b, b->t_sub_items[g], b->waves)), dim3(256), 0, st, b->d_prob,
                               b->d_sub + g * np, b->d_suboff + g * (np + 1), b->t_sub_n[g]);
    };
    half(k_b_bf_khalf<H0>, 0);
    half(k_b_bf_khalf<H1>, 1);
    half(k_b_bf_khalf<H2>, 2);
    mark(5);
    half(k_b_bf_khalf<H3>, 3);
    half(k_b_bf_khalf<H4>, 4);
    mark(6);
    if (b->t_syrk_items[0])
        hipLaunchKernelGGL(k_b_bf_ksyrk<Y0>, dim3(grid_items(b, b->t_syrk_items[0], b->waves)), dim3(Y0::NT), 0, st,
                           b->d_prob, b->d_ssub, b->d_syrkoff, b->t_ssub_n[0]);
    mark(7);
    if (b->t_syrk_items[1])
        hipLaunchKernelGGL(k_b_bf_ksyrk<Y1>, dim3(grid_items(b, b->t_syrk_items[1], b->waves)), dim3(Y1::NT), 0, st,
                           b->d_prob, b->d_ssub + np, b->d_syrkoff + (np + 1), b->t_ssub_n[1]);
    mark(8);
    hipLaunchKernelGGL(k_b_kreduce, dim3(grid_elem(b->t_pairs.back())), dim3(256), 0, st, b->d_prob, np, tb, b->sk);
    mark(9);
    BHIP(b, hipGetLastError());
    return RJK_OK;
}

int enqueue(rjk_batch* b, hipEvent_t* ev = nullptr) {
    switch (b->prec) {
        case RJK_PREC_FP64: return enqueue_fp<double>(b, ev);
        case RJK_PREC_FP32: return enqueue_fp<float>(b, ev);
        case RJK_PREC_BF16X1: return enqueue_bf<1>(b, ev);
        case RJK_PREC_BF16X3: return enqueue_bf<3>(b, ev);
        default: return enqueue_bf<2>(b, ev);
    }
}

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

}  // namespace

extern "C" {

size_t rjk_batch_device_bytes(size_t nprob, const size_t* nbf, const size_t* naux, const size_t* nocc, int prec) {
    if (!nbf || !naux || !nocc || !valid_prec(prec)) return 0;
    size_t tot = 0;
    for (size_t i = 0; i < nprob; ++i) {
        if (!valid_shape(nbf[i], naux[i], nocc[i])) return 0;
        tot += prob_bytes(geometry(nbf[i], naux[i], nocc[i], prec), prec);
    }
    return tot;
}

int rjk_batch_device_memory(int device, size_t* free_bytes, size_t* total_bytes) {
    if (!free_bytes || !total_bytes) return RJK_ERR_ARG;
    int cur = 0;
    if (hipGetDevice(&cur) != hipSuccess || hipSetDevice(device) != hipSuccess) {
        (void)hipGetLastError();
        return RJK_ERR_HIP;
    }
    const hipError_t e = hipMemGetInfo(free_bytes, total_bytes);
    (void)hipSetDevice(cur);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        return RJK_ERR_HIP;
    }
    return RJK_OK;
}

int rjk_batch_plan_waves(size_t nprob, const size_t* nbf, const size_t* naux, const size_t* nocc, int prec,
                         size_t budget_bytes, size_t* wave_start, size_t* nwaves) {
    if (!nbf || !naux || !nocc || !wave_start || !nwaves || nprob == 0 || !valid_prec(prec)) return RJK_ERR_ARG;
    for (size_t i = 0; i < nprob; ++i)
        if (!valid_shape(nbf[i], naux[i], nocc[i])) return RJK_ERR_ARG;
    if (budget_bytes == 0) {
        size_t fr = 0, tot = 0;
        if (hipMemGetInfo(&fr, &tot) != hipSuccess) {
            (void)hipGetLastError();
            return RJK_ERR_HIP;
        }
        budget_bytes = fr / 10 * 9;
    }
    size_t w = 0, used = 0;
    wave_start[0] = 0;
    for (size_t i = 0; i < nprob; ++i) {
        const size_t pb = prob_bytes(geometry(nbf[i], naux[i], nocc[i], prec), prec);
        if (i > wave_start[w] && used + pb > budget_bytes) {
            wave_start[++w] = i;
            used = 0;
        }
        used += pb;
    }
    wave_start[w + 1] = nprob;
    *nwaves = w + 1;
    return RJK_OK;
}

int rjk_batch_create(int device, size_t nprob, const size_t* nbf, const size_t* naux, const size_t* nocc, int prec,
                     rjk_batch** out) {
    if (!out || !nbf || !naux || !nocc || nprob == 0 || nprob > (1u << 20)) return RJK_ERR_ARG;
    *out = nullptr;
    if (!valid_prec(prec)) return RJK_ERR_ARG;
    for (size_t i = 0; i < nprob; ++i)
        if (!valid_shape(nbf[i], naux[i], nocc[i])) return RJK_ERR_ARG;
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
    const size_t es = elem_b(prec);
    b->h.resize(nprob);
    size_t max_compact = 0, max_pad = 0;
    for (size_t i = 0; i < nprob; ++i) {
        b->h[i] = geometry(nbf[i], naux[i], nocc[i], prec);
        const BProb& P = b->h[i];
        max_compact = std::max(max_compact, P.npair * naux[i]);
        max_pad = std::max(max_pad, P.ldb * naux[i]);
        b->bytes_b += P.ldb * naux[i] * es;
        b->bytes_total += prob_bytes(P, prec);
    }
    // one allocation per slot kind; problem i's slot at the 256-byte aligned running offset
    for (int k = 0; k < kNumSlots; ++k) {
        std::vector<size_t> o;
        size_t tot = 0;
        for (const BProb& P : b->h) {
            size_t e[kNumSlots], c[kNumSlots];
            slot_sizes(P, prec, e, c);
            o.push_back(tot);
            tot += rup(e[k] * c[k], 256);
        }
        void* base = nullptr;
        if (hipMalloc(&base, std::max<size_t>(tot, 256)) != hipSuccess) {
            (void)hipGetLastError();
            b->err = "rjk_batch_create: device allocation of " + std::to_string(tot) + " bytes failed";
            return fail_free(RJK_ERR_NOMEM);
        }
        b->allocs.push_back(base);
        for (size_t i = 0; i < nprob; ++i) {
            void* a = (char*)base + o[i];
            BProb& P = b->h[i];
            switch (k) {
                case 0: P.b = a; break;
                case 1: P.dm = (const double*)a; break;
                case 2: P.c = (const double*)a; break;
                case 3: P.occ = (const double*)a; break;
                case 4: P.w = (double*)a; break;
                case 5: P.gamma = (double*)a; break;
                case 6: P.vj = (double*)a; break;
                case 7: P.vk = (double*)a; break;
                case 8: P.cocc = a; break;
                case 9: P.sign = a; break;
                case 10: P.m = a; break;
                default: P.kpart = (double*)a; break;
            }
        }
    }
    // zero the padding of every input/weight slot once (graphs hold no memsets)
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
    prefix(b->t_cols, [](const BProb& P) { return (size_t)(P.naux + kGammaGroup - 1) / kGammaGroup; });
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
        hipMalloc((void**)&b->d_sub, kHalfGroups * nprob * sizeof(int)) != hipSuccess ||
        hipMalloc((void**)&b->d_suboff, kHalfGroups * (nprob + 1) * sizeof(size_t)) != hipSuccess ||
        hipMalloc((void**)&b->d_ssub, 2 * nprob * sizeof(int)) != hipSuccess ||
        hipMalloc((void**)&b->d_syrkoff, 2 * (nprob + 1) * sizeof(size_t)) != hipSuccess)
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
    for (void* p : {(void*)b->d_prob, (void*)b->d_tab, (void*)b->d_sub, (void*)b->d_suboff, (void*)b->d_ssub,
                    (void*)b->d_syrkoff,
                    (void*)b->scratch})
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
    if (!b || !rimatr || i >= b->h.size()) return b ? bfail(b, RJK_ERR_ARG, "rjk_batch_set_rimatr: bad argument") : RJK_ERR_ARG;
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

int rjk_batch_set_rimatr_dev(rjk_batch* b, size_t i, const double* rimatr_dev, size_t ld) {
    if (!b || !rimatr_dev || i >= b->h.size() || ld < b->h[i].npair)
        return b ? bfail(b, RJK_ERR_ARG, "rjk_batch_set_rimatr_dev: bad argument (ld must be >= npair)") : RJK_ERR_ARG;
    const BProb& P = b->h[i];
    BHIP(b, hipSetDevice(b->device));
    const unsigned g = grid_elem(P.ldb * P.naux);
    if (b->prec == RJK_PREC_FP64)
        hipLaunchKernelGGL(k_b_from_dev<double>, dim3(g), dim3(256), 0, b->stream, rimatr_dev, ld, P.npair,
                           (size_t)P.naux, P.ldb, (double*)P.b);
    else
        hipLaunchKernelGGL(k_b_from_dev<float>, dim3(g), dim3(256), 0, b->stream, rimatr_dev, ld, P.npair,
                           (size_t)P.naux, P.ldb, (float*)P.b);
    BHIP(b, hipGetLastError());
    BHIP(b, hipStreamSynchronize(b->stream));
    return RJK_OK;
}

int rjk_batch_set_active(rjk_batch* b, size_t i, int active) {
    if (!b || i >= b->h.size()) return b ? bfail(b, RJK_ERR_ARG, "rjk_batch_set_active: bad argument") : RJK_ERR_ARG;
    BHIP(b, hipSetDevice(b->device));
    BProb& P = b->h[i];
    P.active = active ? 1 : 0;
    // stream-ordered after earlier runs; the captured graph reads the table at every replay
    BHIP(b, hipMemcpyAsync(&b->d_prob[i].active, &P.active, sizeof(int), hipMemcpyHostToDevice, b->stream));
    BHIP(b, hipStreamSynchronize(b->stream));
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
    if (!b || i >= b->h.size()) return b ? bfail(b, RJK_ERR_ARG, "rjk_batch_get_outputs: bad argument") : RJK_ERR_ARG;
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

int rjk_batch_profile(rjk_batch* b, double scale_j, double scale_k, double* stage_ms) {
    if (!b || !stage_ms) return b ? bfail(b, RJK_ERR_ARG, "rjk_batch_profile: bad argument") : RJK_ERR_ARG;
    BHIP(b, hipSetDevice(b->device));
    b->sj = scale_j;
    b->sk = scale_k;
    drop_graph(b);   // the next rjk_batch_run captures again with these scales
    hipEvent_t ev[RJK_BATCH_NUM_STAGES + 1] = {};
    int rc = RJK_OK;
    for (int k = 0; k <= RJK_BATCH_NUM_STAGES && !rc; ++k)
        if (hipEventCreate(&ev[k]) != hipSuccess) rc = bfail(b, RJK_ERR_HIP, "rjk_batch_profile: hipEventCreate");
    if (!rc) rc = enqueue(b, ev);
    if (!rc && hipEventSynchronize(ev[RJK_BATCH_NUM_STAGES]) != hipSuccess)
        rc = bfail(b, RJK_ERR_HIP, "rjk_batch_profile: hipEventSynchronize");
    for (int k = 0; k < RJK_BATCH_NUM_STAGES && !rc; ++k) {
        float ms = 0.f;
        (void)hipEventElapsedTime(&ms, ev[k], ev[k + 1]);
        stage_ms[k] = ms;
    }
    for (int k = 0; k <= RJK_BATCH_NUM_STAGES; ++k)
        if (ev[k]) (void)hipEventDestroy(ev[k]);
    return rc;
}

int rjk_batch_synchronize(rjk_batch* b) {
    if (!b) return RJK_ERR_ARG;
    BHIP(b, hipStreamSynchronize(b->stream));
    return RJK_OK;
}

int rjk_batch_autotune(rjk_batch* b, int reps, double* best_ms) {
    if (!b || reps < 1) return b ? bfail(b, RJK_ERR_ARG, "rjk_batch_autotune: reps must be >= 1") : RJK_ERR_ARG;
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
        return b ? bfail(b, RJK_ERR_ARG, "rjk_batch_set_config: waves_per_cu must be 1..16, cut >= 0") : RJK_ERR_ARG;
    b->khalf_cut = khalf_cut;
    b->waves = waves_per_cu;
    drop_graph(b);
    return build_subsets(b);
}

}  // extern "C"

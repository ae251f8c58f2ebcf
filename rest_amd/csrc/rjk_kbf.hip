// rjk_kbf.hip — exchange (K) on the BF16 matrix cores with split operands.
//
// Every operand x is carried as S BF16 planes, x ≈ x_0 + … + x_{S-1} with
// x_0 = bf16(x) and x_k = bf16(x − x_0 − … − x_{k-1}) (round to nearest even),
// and a product uses the plane pairs whose plane indices sum to less than S:
//   S = 2:  a0·b0 + a0·b1 + a1·b0                        (3 MFMAs, ~2^-16 relative)
//   S = 3:  a0·b0 + a0·b1 + a1·b0 + a0·b2 + a1·b1 + a2·b0 (6 MFMAs, ~2^-24 relative)
// accumulated in FP32 by v_mfma_f32_16x16x32_bf16.  Lane l supplies
// A[l%16][8(l/16)+i] and B[8(l/16)+i][l%16]; register r of lane l receives
// D[4(l/16)+r][l%16] (probed on gfx950).  Split-K partials join an FP64 sum.
//
// Layouts (BF16 planes, padding zero; nbf_pad a multiple of 128, nk of 64):
//   B planes  [S][naux][nbf_pad][nbf_pad]  full symmetric squares (aligned 16-byte row pieces)
//   C' planes [S][nk][nbf_pad]             C'ᵀ: column c of C' as a row, ν fastest
//   M' planes [S][Pc][nbf_pad][nk]         c fastest, so both SYRK operands load as rows
// Stage 1: M'_P[μ][c] = Σ_ν C'ᵀ[c][ν] B_P[ν][μ]                  (128 μ × 64 c tiles)
// Stage 2: Kpart[split][s][ν][μ] (+)= Σ_{P,c} sign_c M'_P[ν][c] M'_P[μ][c]  (128 × 128 upper tiles)
// Stage 3: the FP64 split reduction of rjk_k.hip.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "rjk_internal.hpp"

namespace rjk {
namespace {

typedef __bf16 bf8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));
typedef unsigned int u4 __attribute__((ext_vector_type(4)));

constexpr int kThreads = 256;   // 4 waves, 2 × 2
constexpr int TKB = 32;         // k-block: one MFMA deep
constexpr int RS = TKB + 8;     // LDS row stride in elements (80 bytes: rows stay 16-byte aligned)
constexpr int kXcdB = 8;

__device__ inline size_t tri_b(size_t j) { return j * (j + 1) / 2; }

template <int S>
__device__ inline void split_d(double x, __bf16* out) {
    double r = x;
#pragma unroll
    for (int k = 0; k < S; ++k) {
        const __bf16 h = (__bf16)(float)r;
        out[k] = h;
        r -= (double)(float)h;
    }
}

template <int S>
__device__ inline void split_f(float x, __bf16* out) {
    float r = x;
#pragma unroll
    for (int k = 0; k < S; ++k) {
        const __bf16 h = (__bf16)r;
        out[k] = h;
        r -= (float)h;
    }
}

// plane pair (first operand plane, second operand plane) of product p
template <int S> __device__ constexpr int pa(int p);
template <int S> __device__ constexpr int pb(int p);
template <> __device__ constexpr int pa<2>(int p) { return p == 2 ? 1 : 0; }
template <> __device__ constexpr int pb<2>(int p) { return p == 1 ? 1 : 0; }
template <> __device__ constexpr int pa<3>(int p) { return p == 2 || p == 4 ? 1 : (p == 5 ? 2 : 0); }
template <> __device__ constexpr int pb<3>(int p) { return p == 1 || p == 4 ? 1 : (p == 3 ? 2 : 0); }
template <int S> constexpr int npairs() { return S * (S + 1) / 2; }

__device__ inline f4 mfma(bf8 a, bf8 b, f4 c) { return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0); }

// ---- conversions ---------------------------------------------------------------
// B planes of auxiliary columns [P0, P0 + Pc) from the packed FP64 B (upper triangle,
// element (i<=j) at j(j+1)/2 + i)
template <int S>
__global__ void k_bf_planes_b(const double* __restrict__ b64, size_t ldb, int nbf, int nbf_pad, size_t naux,
                              size_t P0, size_t Pc, __bf16* __restrict__ planes) {
    const size_t sq = (size_t)nbf_pad * nbf_pad, total = Pc * sq;
    for (size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x; t < total; t += (size_t)gridDim.x * blockDim.x) {
        const size_t P = t / sq, e = t % sq;
        const int mu = (int)(e / nbf_pad), nu = (int)(e % nbf_pad);
        double x = 0.0;
        if (mu < nbf && nu < nbf) {
            const size_t lo = mu < nu ? mu : nu, hi = mu < nu ? nu : mu;
            x = b64[(P0 + P) * ldb + tri_b(hi) + lo];
        }
        __bf16 h[S];
        split_d<S>(x, h);
#pragma unroll
        for (int k = 0; k < S; ++k) planes[((size_t)k * naux + P0 + P) * sq + e] = h[k];
    }
}

// C'ᵀ planes and sign masks: column c of spin s is orbital (c − off_s) · sqrt|occ|
template <int S>
__global__ void k_bf_prep_c(const double* __restrict__ eigv, int nbf, int nstate, int nw0, int nw1,
                            const double* __restrict__ occ, int nspin, int off0, int off1, int nk, int nbf_pad,
                            __bf16* __restrict__ ct, unsigned short* __restrict__ smask) {
    const size_t total = (size_t)nk * nbf_pad;
    for (size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x; t < total; t += (size_t)gridDim.x * blockDim.x) {
        const int c = (int)(t / nbf_pad), nu = (int)(t % nbf_pad);
        double v = 0.0, w = 0.0;
        int s = -1, o = 0;
        if (c >= off0 && c < off0 + nw0) {
            s = 0;
            o = c - off0;
        } else if (nspin == 2 && c >= off1 && c < off1 + nw1) {
            s = 1;
            o = c - off1;
        }
        if (s >= 0) {
            w = occ[(size_t)s * nstate + o];
            if (nu < nbf) v = eigv[(size_t)s * nbf * nstate + (size_t)o * nbf + nu] * sqrt(fabs(w));
        }
        __bf16 h[S];
        split_d<S>(v, h);
#pragma unroll
        for (int k = 0; k < S; ++k) ct[(size_t)k * total + t] = h[k];
        if (nu == 0) smask[c] = w < 0 ? 0x8000 : 0;
    }
}

// C'ᵀ planes of gathered eigenvectors (density-only route): column col_off + q is
// vecs[:, idx[q]] · sqrt|lam[q]|; the other columns must be zero beforehand.
template <int S>
__global__ void k_bf_prep_c_idx(const double* __restrict__ vecs, int nbf, const int* __restrict__ idx,
                                const double* __restrict__ lam, int ncol, int col_off, int nk, int nbf_pad,
                                __bf16* __restrict__ ct, unsigned short* __restrict__ smask) {
    const size_t total = (size_t)ncol * nbf_pad, plane = (size_t)nk * nbf_pad;
    for (size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x; t < total; t += (size_t)gridDim.x * blockDim.x) {
        const int q = (int)(t / nbf_pad), nu = (int)(t % nbf_pad);
        const double w = lam[q];
        const double v = nu < nbf ? vecs[(size_t)idx[q] * nbf + nu] * sqrt(fabs(w)) : 0.0;
        __bf16 h[S];
        split_d<S>(v, h);
        const size_t e = (size_t)(col_off + q) * nbf_pad + nu;
#pragma unroll
        for (int k = 0; k < S; ++k) ct[(size_t)k * plane + e] = h[k];
        if (nu == 0) smask[col_off + q] = w < 0 ? 0x8000 : 0;
    }
}

__device__ inline bf8 frag(const __bf16* row) { return *(const bf8*)row; }

// ---- stage 1 -------------------------------------------------------------------
// 1-D grid, XCD-aware: every tile of auxiliary column P runs on XCD P % 8 and the
// column tiles of one μ tile are consecutive there, so they share B's rows in L2.
// Work-group tile 128 μ × 64 c; wave (wm, wn) computes 64 μ × 32 c.
template <int S>
__global__ __launch_bounds__(kThreads) void k_bf_khalf(const __bf16* __restrict__ bpl, size_t naux, int nbf,
                                                       int nbf_pad, size_t P0, int Pc, int ntm, int ntc,
                                                       const __bf16* __restrict__ ct, int nk,
                                                       __bf16* __restrict__ mpl) {
    constexpr int NP = npairs<S>();
    __shared__ __bf16 Bs[S][128][RS];   // [μ][ν]
    __shared__ __bf16 Cs[S][64][RS];    // [c][ν]
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = wave & 1, wn = wave >> 1;
    const int w = blockIdx.x, xcd = w % kXcdB, kq = w / kXcdB, per = ntm * ntc;
    const int Pl = (kq / per) * kXcdB + xcd;
    if (Pl >= Pc) return;   // padding work-groups (uniform exit)
    const int rem = kq % per, mt = rem / ntc, ctile = rem % ntc;
    const int mu0 = mt * 128, c0 = ctile * 64;
    const size_t sq = (size_t)nbf_pad * nbf_pad, bstride = naux * sq, cplane = (size_t)nk * nbf_pad;
    const __bf16* __restrict__ bP = bpl + (P0 + Pl) * sq + (size_t)mu0 * nbf_pad;
    const __bf16* __restrict__ cT = ct + (size_t)c0 * nbf_pad;
    // staging: row = tid / 4 (0..63), piece = 8 elements at (tid % 4) * 8;
    // B rows row and row + 64, C' row row
    const int row = tid >> 2, pc = (tid & 3) * 8;
    u4 rb0[S], rb1[S], rc[S];
    auto fetch = [&](int nu0) {
#pragma unroll
        for (int k = 0; k < S; ++k) {
            const __bf16* b = bP + (size_t)k * bstride + nu0 + pc;
            rb0[k] = *(const u4*)(b + (size_t)row * nbf_pad);
            rb1[k] = *(const u4*)(b + (size_t)(row + 64) * nbf_pad);
            rc[k] = *(const u4*)(cT + (size_t)k * cplane + (size_t)row * nbf_pad + nu0 + pc);
        }
    };
    f4 acc[4][2];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) acc[i][j] = f4{0.f, 0.f, 0.f, 0.f};
    const int nkb = (nbf + TKB - 1) / TKB;
    const int kk = 8 * (lane >> 4), l16 = lane & 15;
    fetch(0);
    for (int kb = 0; kb < nkb; ++kb) {
#pragma unroll
        for (int k = 0; k < S; ++k) {
            *(u4*)&Bs[k][row][pc] = rb0[k];
            *(u4*)&Bs[k][row + 64][pc] = rb1[k];
            *(u4*)&Cs[k][row][pc] = rc[k];
        }
        __syncthreads();
        if (kb + 1 < nkb) fetch((kb + 1) * TKB);
        bf8 fb[4][S], fc[2][S];
#pragma unroll
        for (int k = 0; k < S; ++k) {
#pragma unroll
            for (int i = 0; i < 4; ++i) fb[i][k] = frag(&Bs[k][wm * 64 + i * 16 + l16][kk]);
#pragma unroll
            for (int j = 0; j < 2; ++j) fc[j][k] = frag(&Cs[k][wn * 32 + j * 16 + l16][kk]);
        }
        // D[c][μ] = Σ_ν C'ᵀ[c][ν] B[ν][μ]; products outermost so consecutive MFMAs
        // write different accumulators
#pragma unroll
        for (int p = 0; p < NP; ++p)
#pragma unroll
            for (int i = 0; i < 4; ++i)
#pragma unroll
                for (int j = 0; j < 2; ++j) acc[i][j] = mfma(fc[j][pa<S>(p)], fb[i][pb<S>(p)], acc[i][j]);
        __syncthreads();
    }
    // M' planes: lane holds c = cb + 4(l/16) + r for μ = mb + l%16
    const size_t mplane = (size_t)Pc * nbf_pad * nk;
    __bf16* __restrict__ mP = mpl + (size_t)Pl * nbf_pad * nk;
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) {
            const int mu = mu0 + wm * 64 + i * 16 + l16;
            const int c = c0 + wn * 32 + j * 16 + 4 * (lane >> 4);
            __bf16 h[4][S];
#pragma unroll
            for (int r = 0; r < 4; ++r) split_f<S>(acc[i][j][r], h[r]);
#pragma unroll
            for (int k = 0; k < S; ++k)
                *(bf4*)(mP + (size_t)k * mplane + (size_t)mu * nk + c) = bf4{h[0][k], h[1][k], h[2][k], h[3][k]};
        }
}

// ---- stage 2 -------------------------------------------------------------------
// 1-D grid of nsplit × nspin × ntile2 work-groups (nsplit a multiple of 8; split s
// on XCD s % 8).  Work-group tile 128 ν × 128 μ (μ tile <= ν tile); wave (wm, wn)
// computes 64 ν × 64 μ.
template <int S>
__global__ __launch_bounds__(kThreads) void k_bf_ksyrk(const __bf16* __restrict__ mpl, int Pc, int nbf_pad, int nk,
                                                       int ntile2, int nspin, int off0, int cnt0, int off1, int cnt1,
                                                       const unsigned short* __restrict__ smask, int p_per_split,
                                                       int beta, double* __restrict__ kpart, size_t slot) {
    constexpr int NP = npairs<S>();
    __shared__ __bf16 As[S][128][RS];   // [ν][c], signs applied
    __shared__ __bf16 Bs[S][128][RS];   // [μ][c]
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = wave & 1, wn = wave >> 1;
    const int w = blockIdx.x, xcd = w % kXcdB, kq = w / kXcdB;
    const size_t t = (size_t)(kq % ntile2);
    const int rest = kq / ntile2, spin = rest % nspin, split = (rest / nspin) * kXcdB + xcd;
    size_t tn = (size_t)((sqrt(8.0 * (double)t + 1.0) - 1.0) * 0.5);
    while (tri_b(tn) > t) --tn;
    while (tri_b(tn + 1) <= t) ++tn;
    const size_t tm = t - tri_b(tn);
    const int mu0 = (int)tm * 128, nu0 = (int)tn * 128;
    const int cb = spin == 0 ? off0 : off1, cnt = spin == 0 ? cnt0 : cnt1;
    const int ncb = (cnt + TKB - 1) / TKB;
    const int Pb = split * p_per_split, Pe = min(Pc, Pb + p_per_split);
    const int nq = (Pe > Pb ? Pe - Pb : 0) * ncb;
    const size_t mplane = (size_t)Pc * nbf_pad * nk;
    const int row = tid >> 2, pc = (tid & 3) * 8;
    u4 ra0[S], ra1[S], rb0[S], rb1[S], rm;
    auto fetch = [&](int q) {
        const int P = Pb + q / ncb, c = cb + (q % ncb) * TKB + pc;
        const __bf16* base = mpl + (size_t)P * nbf_pad * nk + c;
#pragma unroll
        for (int k = 0; k < S; ++k) {
            const __bf16* bk = base + (size_t)k * mplane;
            ra0[k] = *(const u4*)(bk + (size_t)(nu0 + row) * nk);
            ra1[k] = *(const u4*)(bk + (size_t)(nu0 + row + 64) * nk);
            rb0[k] = *(const u4*)(bk + (size_t)(mu0 + row) * nk);
            rb1[k] = *(const u4*)(bk + (size_t)(mu0 + row + 64) * nk);
        }
        const unsigned short* sm = smask + c;
        rm = u4{(unsigned)sm[0] | ((unsigned)sm[1] << 16), (unsigned)sm[2] | ((unsigned)sm[3] << 16),
                (unsigned)sm[4] | ((unsigned)sm[5] << 16), (unsigned)sm[6] | ((unsigned)sm[7] << 16)};
    };
    f4 acc[4][4];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] = f4{0.f, 0.f, 0.f, 0.f};
    const int kk = 8 * (lane >> 4), l16 = lane & 15;
    if (nq > 0) fetch(0);
    for (int q = 0; q < nq; ++q) {
#pragma unroll
        for (int k = 0; k < S; ++k) {
            *(u4*)&As[k][row][pc] = ra0[k] ^ rm;   // sign_c: flip the sign bits exactly
            *(u4*)&As[k][row + 64][pc] = ra1[k] ^ rm;
            *(u4*)&Bs[k][row][pc] = rb0[k];
            *(u4*)&Bs[k][row + 64][pc] = rb1[k];
        }
        __syncthreads();
        if (q + 1 < nq) fetch(q + 1);
        bf8 fa[4][S], fb[4][S];
#pragma unroll
        for (int k = 0; k < S; ++k)
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                fa[i][k] = frag(&As[k][wn * 64 + i * 16 + l16][kk]);
                fb[i][k] = frag(&Bs[k][wm * 64 + i * 16 + l16][kk]);
            }
        // D[ν][μ] = Σ_c sign_c M'[ν][c] M'[μ][c]
#pragma unroll
        for (int p = 0; p < NP; ++p)
#pragma unroll
            for (int i = 0; i < 4; ++i)
#pragma unroll
                for (int j = 0; j < 4; ++j) acc[i][j] = mfma(fa[j][pa<S>(p)], fb[i][pb<S>(p)], acc[i][j]);
        __syncthreads();
    }
    double* __restrict__ kp = kpart + ((size_t)split * nspin + spin) * slot;
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int mu = mu0 + wm * 64 + i * 16 + l16;
            const int nb = nu0 + wn * 64 + j * 16 + 4 * (lane >> 4);
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                double* dst = kp + (size_t)(nb + r) * nbf_pad + mu;
                double v = (double)acc[i][j][r];
                if (beta) v += *dst;
                *dst = v;
            }
        }
}

unsigned gs_blocks(size_t n) { return (unsigned)std::max<size_t>(1, std::min<size_t>((n + 255) / 256, 65536)); }

}  // namespace

int bf_planes(int prec) { return prec == RJK_PREC_BF16X3 ? 3 : 2; }

hipError_t launch_bf_planes_b(hipStream_t st, const double* b64, size_t ldb, size_t nbf, size_t nbf_pad,
                              size_t naux, void* planes) {
    // in slabs of auxiliary columns so each launch's grid-stride loop stays short
    const size_t sq = nbf_pad * nbf_pad, slab = std::max<size_t>(1, (size_t(1) << 28) / sq);
    for (size_t P0 = 0; P0 < naux; P0 += slab) {
        const size_t Pc = std::min(slab, naux - P0);
        hipLaunchKernelGGL(k_bf_planes_b<kBfStoredPlanes>, dim3(gs_blocks(Pc * sq)), dim3(256), 0, st, b64, ldb,
                           (int)nbf, (int)nbf_pad, naux, P0, Pc, (__bf16*)planes);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

hipError_t launch_bf_prep_c(hipStream_t st, int prec, const double* eigv, size_t nbf, size_t nstate,
                            const size_t* nw, const double* occ, int nspin, const KPlan& kp, void* ct, void* smask) {
    const size_t n = kp.nk * kp.nbf_pad;
    const int nw1 = nspin == 2 ? (int)nw[1] : 0;
    if (bf_planes(prec) == 3)
        hipLaunchKernelGGL(k_bf_prep_c<3>, dim3(gs_blocks(n)), dim3(256), 0, st, eigv, (int)nbf, (int)nstate,
                           (int)nw[0], nw1, occ, nspin, (int)kp.col_off[0], (int)kp.col_off[1], (int)kp.nk,
                           (int)kp.nbf_pad, (__bf16*)ct, (unsigned short*)smask);
    else
        hipLaunchKernelGGL(k_bf_prep_c<2>, dim3(gs_blocks(n)), dim3(256), 0, st, eigv, (int)nbf, (int)nstate,
                           (int)nw[0], nw1, occ, nspin, (int)kp.col_off[0], (int)kp.col_off[1], (int)kp.nk,
                           (int)kp.nbf_pad, (__bf16*)ct, (unsigned short*)smask);
    return hipGetLastError();
}

hipError_t launch_bf_prep_c_idx(hipStream_t st, int prec, const double* vecs, size_t nbf, const int* idx,
                                const double* lam, size_t ncol, size_t col_off, const KPlan& kp, void* ct,
                                void* smask) {
    if (ncol == 0) return hipSuccess;
    const size_t n = ncol * kp.nbf_pad;
    if (bf_planes(prec) == 3)
        hipLaunchKernelGGL(k_bf_prep_c_idx<3>, dim3(gs_blocks(n)), dim3(256), 0, st, vecs, (int)nbf, idx, lam,
                           (int)ncol, (int)col_off, (int)kp.nk, (int)kp.nbf_pad, (__bf16*)ct,
                           (unsigned short*)smask);
    else
        hipLaunchKernelGGL(k_bf_prep_c_idx<2>, dim3(gs_blocks(n)), dim3(256), 0, st, vecs, (int)nbf, idx, lam,
                           (int)ncol, (int)col_off, (int)kp.nk, (int)kp.nbf_pad, (__bf16*)ct,
                           (unsigned short*)smask);
    return hipGetLastError();
}

hipError_t launch_bf_khalf(hipStream_t st, int prec, const void* bpl, size_t naux, size_t nbf, size_t P0, size_t Pc,
                           const KPlan& kp, const void* ct, void* mpl) {
    const int ntm = (int)(kp.nbf_pad / 128), ntc = (int)(kp.nk / 64);
    const size_t nwg = ((Pc + kXcdB - 1) / kXcdB * kXcdB) * (size_t)ntm * ntc;
    if (bf_planes(prec) == 3)
        hipLaunchKernelGGL(k_bf_khalf<3>, dim3((unsigned)nwg), dim3(kThreads), 0, st, (const __bf16*)bpl, naux,
                           (int)nbf, (int)kp.nbf_pad, P0, (int)Pc, ntm, ntc, (const __bf16*)ct, (int)kp.nk,
                           (__bf16*)mpl);
    else
        hipLaunchKernelGGL(k_bf_khalf<2>, dim3((unsigned)nwg), dim3(kThreads), 0, st, (const __bf16*)bpl, naux,
                           (int)nbf, (int)kp.nbf_pad, P0, (int)Pc, ntm, ntc, (const __bf16*)ct, (int)kp.nk,
                           (__bf16*)mpl);
    return hipGetLastError();
}

hipError_t launch_bf_ksyrk(hipStream_t st, int prec, const void* mpl, size_t Pc, const KPlan& kp, int nspin,
                           const void* smask, int beta, double* kpart) {
    const size_t nt = kp.nbf_pad / 128, nt2 = nt * (nt + 1) / 2;
    const int pps = (int)((Pc + kp.nsplit - 1) / kp.nsplit);
    const size_t nwg = kp.nsplit * (size_t)nspin * nt2;
    if (bf_planes(prec) == 3)
        hipLaunchKernelGGL(k_bf_ksyrk<3>, dim3((unsigned)nwg), dim3(kThreads), 0, st, (const __bf16*)mpl, (int)Pc,
                           (int)kp.nbf_pad, (int)kp.nk, (int)nt2, nspin, (int)kp.col_off[0], (int)kp.col_cnt[0],
                           (int)kp.col_off[1], (int)kp.col_cnt[1], (const unsigned short*)smask, pps, beta, kpart,
                           kp.nbf_pad * kp.nbf_pad);
    else
        hipLaunchKernelGGL(k_bf_ksyrk<2>, dim3((unsigned)nwg), dim3(kThreads), 0, st, (const __bf16*)mpl, (int)Pc,
                           (int)kp.nbf_pad, (int)kp.nk, (int)nt2, nspin, (int)kp.col_off[0], (int)kp.col_cnt[0],
                           (int)kp.col_off[1], (int)kp.col_cnt[1], (const unsigned short*)smask, pps, beta, kpart,
                           kp.nbf_pad * kp.nbf_pad);
    return hipGetLastError();
}

}  // namespace rjk

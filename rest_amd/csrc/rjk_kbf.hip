// rjk_kbf.hip — exchange (K) on the BF16 matrix cores with split operands.
//
// Every operand x is carried as S BF16 planes, x ≈ x_0 + … + x_{S-1} with
// x_0 = bf16(x) and x_k = bf16(x − x_0 − … − x_{k-1}) (round to nearest even),
// and a product uses the plane pairs whose plane indices sum to less than S:
//   S = 2:  a0·b0 + a0·b1 + a1·b0                        (3 MFMAs, ~2^-16 relative)
//   S = 3:  a0·b0 + a0·b1 + a1·b0 + a0·b2 + a1·b1 + a2·b0 (6 MFMAs, ~2^-24 relative)
//   S = 1:  a0·b0                                         (1 MFMA, ~2^-8 relative)
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

#include "rjk_bftile.hpp"
#include "rjk_internal.hpp"

namespace rjk {
namespace {

typedef __bf16 bf4 __attribute__((ext_vector_type(4)));

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

// Tile shapes from a sweep at nbf 1536, nk 320, 256 auxiliary columns per launch (executed
// BF16 MFMA rate, accumulators in VGPRs, the conflict-free LDS swizzle of rjk_bftile.hpp), then
// checked in the K build itself (about 1000 auxiliary columns per launch):
//   half-transform, 2 planes: 64 c × 256 μ on 1 × 4 waves, 64-deep: 5.39 ms per launch in the
//     build against 5.46 for 64 × 128 on 2 × 2 waves (the sweep, at a quarter of the columns,
//     ranks them the other way: 747 against 857 TFLOP/s); needs nbf_pad % 256 == 0, else 64 × 128;
//   half-transform, 3 planes: 64 c × 128 μ on 2 × 4 waves, 64-deep: 9.34 ms per 1024 columns
//     (2 × 2 waves: 9.64; 64 × 256 on 2 × 4: 10.24);
//   SYRK, 2 planes: 128 × 128 on 2 × 4 waves, 64-deep: 988 (2 × 2 waves: 951; in the build 2.66
//     against 2.93 ms per launch);
//   SYRK, 3 planes: 128 × 128 on 2 × 4 waves, 32-deep: 1110 (2 × 2 waves: 1077; 64-deep: 1043).
template <int S> struct KbfTiles;
template <> struct KbfTiles<2> {
    using Half = bft::Cfg<2, 64, 256, 1, 4, false, 64>;    // X = 64 c, Y = 256 μ
    using HalfNarrow = bft::Cfg<2, 64, 128, 2, 2, false, 64>;
    using Syrk = bft::Cfg<2, 128, 128, 2, 4, false, 64>;   // X = 128 ν, Y = 128 μ
};
template <> struct KbfTiles<1> {
    using Half = bft::Cfg<1, 64, 256, 1, 4, false, 64>;
    using HalfNarrow = bft::Cfg<1, 64, 128, 2, 2, false, 64>;
    using Syrk = bft::Cfg<1, 128, 128, 2, 4, false, 64>;
};
template <> struct KbfTiles<3> {
    using Half = bft::Cfg<3, 64, 128, 2, 4, false, 64>;
    using Syrk = bft::Cfg<3, 128, 128, 2, 4, false, 32>;
};

// ---- stage 1 -------------------------------------------------------------------
// 1-D grid, XCD-aware: every tile of auxiliary column P runs on XCD P % 8 and the
// column tiles of one μ tile are consecutive there, so they share B's rows in L2.
template <class C>
__global__ __launch_bounds__(C::NT) void k_bf_khalf(const __bf16* __restrict__ bpl, size_t naux, int nbf,
                                                    int nbf_pad, size_t P0, int Pc, int ntm, int ntc,
                                                    const __bf16* __restrict__ ct, int nk,
                                                    __bf16* __restrict__ mpl) {
    constexpr int S = C::S;
    __shared__ __bf16 lds[C::LDS_ELEMS];
    const int w = blockIdx.x, xcd = w % kXcdB, kq = w / kXcdB, per = ntm * ntc;
    const int Pl = (kq / per) * kXcdB + xcd;
    if (Pl >= Pc) return;   // padding work-groups (uniform exit)
    const int rem = kq % per, mt = rem / ntc, ctile = rem % ntc;
    const int mu0 = mt * C::WGY, c0 = ctile * C::WGX;
    const size_t sq = (size_t)nbf_pad * nbf_pad;
    const __bf16* __restrict__ bP = bpl + (P0 + Pl) * sq + (size_t)mu0 * nbf_pad;
    const __bf16* __restrict__ cT = ct + (size_t)c0 * nbf_pad;
    bft::f4 acc[C::FX][C::FY];
    // D[c][μ] = Σ_ν C'ᵀ[c][ν] B[μ][ν]  (B_P symmetric; padding rows and columns are zero)
    bft::nt_tile<C>(lds, [&](int q) { return cT + q * C::TK; }, (size_t)nbf_pad, (size_t)nk * nbf_pad,
                    [&](int q) { return bP + q * C::TK; }, (size_t)nbf_pad, naux * sq,
                    [](int, int) { return bft::u4{0, 0, 0, 0}; }, (nbf + C::TK - 1) / C::TK, acc);
    // M' planes: lane holds c = cb + 4(l/16) + r for μ = mb + l%16
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, wx = wave % C::WX, wy = wave / C::WX;
    const size_t mplane = (size_t)Pc * nbf_pad * nk;
    __bf16* __restrict__ mP = mpl + (size_t)Pl * nbf_pad * nk;
#pragma unroll
    for (int i = 0; i < C::FX; ++i)
#pragma unroll
        for (int j = 0; j < C::FY; ++j) {
            const int c = c0 + (wx * C::FX + i) * 16 + 4 * (lane >> 4);
            const int mu = mu0 + (wy * C::FY + j) * 16 + (lane & 15);
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
// on XCD s % 8).  Upper tile pairs (μ tile <= ν tile) of 128 × 128.
template <class C>
__global__ __launch_bounds__(C::NT) void k_bf_ksyrk(const __bf16* __restrict__ mpl, int Pc, int nbf_pad,
                                                    int nk, int ntile2, int nspin, int off0, int cnt0,
                                                    int off1, int cnt1,
                                                    const unsigned short* __restrict__ smask,
                                                    int p_per_split, int beta, double* __restrict__ kpart,
                                                    size_t slot) {
    __shared__ __bf16 lds[C::LDS_ELEMS];
    const int w = blockIdx.x, xcd = w % kXcdB, kq = w / kXcdB;
    const size_t t = (size_t)(kq % ntile2);
    const int rest = kq / ntile2, spin = rest % nspin, split = (rest / nspin) * kXcdB + xcd;
    size_t tn = (size_t)((sqrt(8.0 * (double)t + 1.0) - 1.0) * 0.5);
    while (tri_b(tn) > t) --tn;
    while (tri_b(tn + 1) <= t) ++tn;
    const size_t tm = t - tri_b(tn);
    const int mu0 = (int)tm * C::WGY, nu0 = (int)tn * C::WGX;
    const int cb = spin == 0 ? off0 : off1, cnt = spin == 0 ? cnt0 : cnt1;
    const int ncb = (cnt + C::TK - 1) / C::TK;
    const int Pb = split * p_per_split, Pe = min(Pc, Pb + p_per_split);
    const int nq = (Pe > Pb ? Pe - Pb : 0) * ncb;
    const size_t mplane = (size_t)Pc * nbf_pad * nk, mP = (size_t)nbf_pad * nk;
    bft::f4 acc[C::FX][C::FY];
    // D[ν][μ] = Σ_c sign_c M'[ν][c] M'[μ][c]
    bft::nt_tile<C>(
        lds, [&](int q) { return mpl + (Pb + q / ncb) * mP + (size_t)nu0 * nk + cb + (q % ncb) * C::TK; }, (size_t)nk,
        mplane, [&](int q) { return mpl + (Pb + q / ncb) * mP + (size_t)mu0 * nk + cb + (q % ncb) * C::TK; },
        (size_t)nk, mplane,
        [&](int q, int pc) {
            const unsigned short* sm = smask + cb + (q % ncb) * C::TK + pc;
            return bft::u4{(unsigned)sm[0] | ((unsigned)sm[1] << 16), (unsigned)sm[2] | ((unsigned)sm[3] << 16),
                           (unsigned)sm[4] | ((unsigned)sm[5] << 16), (unsigned)sm[6] | ((unsigned)sm[7] << 16)};
        },
        nq, acc);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, wx = wave % C::WX, wy = wave / C::WX;
    double* __restrict__ kp = kpart + ((size_t)split * nspin + spin) * slot;
#pragma unroll
    for (int i = 0; i < C::FX; ++i)
#pragma unroll
        for (int j = 0; j < C::FY; ++j) {
            const int nb = nu0 + (wx * C::FX + i) * 16 + 4 * (lane >> 4);
            const int mu = mu0 + (wy * C::FY + j) * 16 + (lane & 15);
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

int bf_planes(int prec) { return prec == RJK_PREC_BF16X3 ? 3 : (prec == RJK_PREC_BF16X1 ? 1 : 2); }

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
    auto go = [&](auto kern) {
        hipLaunchKernelGGL(kern, dim3(gs_blocks(n)), dim3(256), 0, st, eigv, (int)nbf, (int)nstate, (int)nw[0], nw1,
                           occ, nspin, (int)kp.col_off[0], (int)kp.col_off[1], (int)kp.nk, (int)kp.nbf_pad,
                           (__bf16*)ct, (unsigned short*)smask);
    };
    switch (bf_planes(prec)) {
        case 3: go(k_bf_prep_c<3>); break;
        case 1: go(k_bf_prep_c<1>); break;
        default: go(k_bf_prep_c<2>); break;
    }
    return hipGetLastError();
}

hipError_t launch_bf_prep_c_idx(hipStream_t st, int prec, const double* vecs, size_t nbf, const int* idx,
                                const double* lam, size_t ncol, size_t col_off, const KPlan& kp, void* ct,
                                void* smask) {
    if (ncol == 0) return hipSuccess;
    const size_t n = ncol * kp.nbf_pad;
    auto go = [&](auto kern) {
        hipLaunchKernelGGL(kern, dim3(gs_blocks(n)), dim3(256), 0, st, vecs, (int)nbf, idx, lam, (int)ncol,
                           (int)col_off, (int)kp.nk, (int)kp.nbf_pad, (__bf16*)ct, (unsigned short*)smask);
    };
    switch (bf_planes(prec)) {
        case 3: go(k_bf_prep_c_idx<3>); break;
        case 1: go(k_bf_prep_c_idx<1>); break;
        default: go(k_bf_prep_c_idx<2>); break;
    }
    return hipGetLastError();
}

namespace {

template <class C>
hipError_t bf_khalf_t(hipStream_t st, const void* bpl, size_t naux, size_t nbf, size_t P0, size_t Pc, const KPlan& kp,
                      const void* ct, void* mpl) {
    const int ntm = (int)(kp.nbf_pad / C::WGY), ntc = (int)(kp.nk / C::WGX);
    const size_t nwg = ((Pc + kXcdB - 1) / kXcdB * kXcdB) * (size_t)ntm * ntc;
    hipLaunchKernelGGL(k_bf_khalf<C>, dim3((unsigned)nwg), dim3(C::NT), 0, st, (const __bf16*)bpl, naux, (int)nbf,
                       (int)kp.nbf_pad, P0, (int)Pc, ntm, ntc, (const __bf16*)ct, (int)kp.nk, (__bf16*)mpl);
    return hipGetLastError();
}

template <class C>
hipError_t bf_ksyrk_t(hipStream_t st, const void* mpl, size_t Pc, const KPlan& kp, int nspin, const void* smask,
                      int beta, double* kpart) {
    const size_t nt = kp.nbf_pad / C::WGX, nt2 = nt * (nt + 1) / 2;
    const int pps = (int)((Pc + kp.nsplit - 1) / kp.nsplit);
    const size_t nwg = kp.nsplit * (size_t)nspin * nt2;
    hipLaunchKernelGGL(k_bf_ksyrk<C>, dim3((unsigned)nwg), dim3(C::NT), 0, st, (const __bf16*)mpl, (int)Pc,
                       (int)kp.nbf_pad, (int)kp.nk, (int)nt2, nspin, (int)kp.col_off[0], (int)kp.col_cnt[0],
                       (int)kp.col_off[1], (int)kp.col_cnt[1], (const unsigned short*)smask, pps, beta, kpart,
                       kp.nbf_pad * kp.nbf_pad);
    return hipGetLastError();
}

}  // namespace

// Tile preconditions: nbf_pad a multiple of 128 (both SYRK tiles, the narrow half-transform
// tile) or of 256 (the wide one, chosen only then); nk a multiple of 64 (the c tiles); every
// spin's column count a multiple of kBfColGran = 64 (the deepest k-block).
hipError_t launch_bf_khalf(hipStream_t st, int prec, const void* bpl, size_t naux, size_t nbf, size_t P0, size_t Pc,
                           const KPlan& kp, const void* ct, void* mpl) {
    const int S = bf_planes(prec);
    if (S == 3) return bf_khalf_t<KbfTiles<3>::Half>(st, bpl, naux, nbf, P0, Pc, kp, ct, mpl);
    const bool wide = kp.nbf_pad % KbfTiles<2>::Half::WGY == 0;
    if (S == 1)
        return wide ? bf_khalf_t<KbfTiles<1>::Half>(st, bpl, naux, nbf, P0, Pc, kp, ct, mpl)
                    : bf_khalf_t<KbfTiles<1>::HalfNarrow>(st, bpl, naux, nbf, P0, Pc, kp, ct, mpl);
    return wide ? bf_khalf_t<KbfTiles<2>::Half>(st, bpl, naux, nbf, P0, Pc, kp, ct, mpl)
                : bf_khalf_t<KbfTiles<2>::HalfNarrow>(st, bpl, naux, nbf, P0, Pc, kp, ct, mpl);
}

hipError_t launch_bf_ksyrk(hipStream_t st, int prec, const void* mpl, size_t Pc, const KPlan& kp, int nspin,
                           const void* smask, int beta, double* kpart) {
    switch (bf_planes(prec)) {
        case 3: return bf_ksyrk_t<KbfTiles<3>::Syrk>(st, mpl, Pc, kp, nspin, smask, beta, kpart);
        case 1: return bf_ksyrk_t<KbfTiles<1>::Syrk>(st, mpl, Pc, kp, nspin, smask, beta, kpart);
        default: return bf_ksyrk_t<KbfTiles<2>::Syrk>(st, mpl, Pc, kp, nspin, smask, beta, kpart);
    }
}

}  // namespace rjk

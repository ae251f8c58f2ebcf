// rjk_ktiles.hpp — the K tiles (half-transform and SYRK) as device functions,
// shared by the single-problem kernels (rjk_k.hip) and the grouped persistent
// kernels of a batch of problems (rjk_batch.hip).
//
// MFMA: v_mfma_f64_16x16x4f64.  Lane l supplies A[l%16][l/16] and
// B[l/16][l%16]; register r of lane l receives D[l/16 + 4r][l%16] (probed on
// gfx950).  Both tiles put the memory-contiguous index (μ) on the D column
// so that 16 consecutive lanes store 128 contiguous bytes.
#pragma once
#include "rjk_internal.hpp"

namespace rjk {
namespace kt {

typedef double d4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));

constexpr int kBlock = 256;           // 4 waves, 2×2
constexpr int TK = kTileK;            // 16
// LDS row padding (elements).  FP64: rows of 128 + 2 (or 64 + 2) doubles put the 16 rows that
// one transposed stash instruction writes (one element each, 16 lanes at a time) on 16 distinct
// 4-bank groups, and the 16 consecutive doubles of a fragment read on consecutive banks; a pad of
// 4 leaves rows k and k + 8 on the same banks.  FP32 keeps 4.
template <typename T> struct Pad { static constexpr int v = sizeof(T) == 8 ? 2 : 4; };
constexpr int kXcd = 8;               // MI355X: 8 XCDs, work-groups dealt round-robin by linear id

__device__ inline size_t tri(size_t j) { return j * (j + 1) / 2; }

// A device pointer read from a table in memory (not a kernel argument) is a flat pointer to the
// compiler, and flat loads count against the LDS wait counter as well as the vector-memory one,
// so every LDS wait of a pipelined tile would also wait for the next block's loads.  glb() gives
// such a pointer the global address space; the tiles below take their pointers as template types,
// so these accesses stay global loads and stores.
#define RJK_GLB __attribute__((address_space(1)))
template <typename T> __device__ __forceinline__ RJK_GLB T* glb(T* p) { return (RJK_GLB T*)p; }

// One 16x16x4 MFMA per element type.  Lane l supplies A[l%16][l/16] and
// B[l/16][l%16]; register r of lane l holds D[row(l, r)][l%16] (probed on
// gfx950: FP64 interleaves rows by 4, FP32 keeps 4 consecutive rows).
template <typename T> struct Mma;
template <> struct Mma<double> {
    using acc_t = d4;
    __device__ static acc_t zero() { return d4{0.0, 0.0, 0.0, 0.0}; }
    __device__ static acc_t op(double a, double b, acc_t c) {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
    __device__ static int row(int lane, int r) { return (lane >> 4) + 4 * r; }
};
template <> struct Mma<float> {
    using acc_t = f4;
    __device__ static acc_t zero() { return f4{0.f, 0.f, 0.f, 0.f}; }
    __device__ static acc_t op(float a, float b, acc_t c) {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
    __device__ static int row(int lane, int r) { return 4 * (lane >> 4) + r; }
};

// upper tile pair (tm <= tn) of packed tile index t
__device__ inline void upper_tile(size_t t, size_t* tm, size_t* tn) {
    size_t n = (size_t)((sqrt(8.0 * (double)t + 1.0) - 1.0) * 0.5);
    while (tri(n) > t) --n;
    while (tri(n + 1) <= t) ++n;
    *tn = n;
    *tm = t - tri(n);
}

// ---------------------------------------------------------------------------
// Stage 1: M'[P][c][μ] = Σ_ν Bsym_P[μ,ν] C'[ν][c], one (μ-tile, c-tile) of one
//   auxiliary column.  Software pipeline: the next k-block is fetched into
//   registers while the MFMAs consume the current one from LDS; LDS is
//   double-buffered, one barrier per k-block.
// ---------------------------------------------------------------------------
template <typename T, int FM, int FN, typename BP, typename CP, typename MP>
__device__ __forceinline__ void khalf_tile(BP __restrict__ bcol, int nbf, int mu0, int c0, CP __restrict__ cocc,
                                           int nk, int ldm, MP __restrict__ mP) {
    using MM = Mma<T>;
    constexpr int TM = 2 * FM * 16, TN = 2 * FN * 16;
    constexpr int NA = TK * TM / kBlock, NB = TK * TN / kBlock;
    __shared__ T As[2][TK][TM + Pad<T>::v];
    __shared__ T Bs[2][TK][TN + Pad<T>::v];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = wave & 1, wn = wave >> 1;
    T ra[NA], rb[NB];
    unsigned live = 0;   // bit r: element r of the A fetch lies inside nbf×nbf

    // A tile As[k][mm] = Bsym(mu0+mm, nu0+k).  Tiles wholly below the diagonal
    // are read along ν (contiguous there), all others along μ.  Out-of-range
    // elements read a valid address and are zeroed when stashed, so nothing
    // waits on the loads before the MFMAs of the current block.
    auto fetch = [&](int nu0) {
        live = 0;
        if (mu0 >= nu0 + TK - 1) {
#pragma unroll
            for (int r = 0; r < NA; ++r) {
                const int e = tid + r * kBlock, k = e % TK, mm = e / TK;
                const int mu = mu0 + mm, nu = nu0 + k;
                const bool ok = mu < nbf && nu < nbf;
                live |= (unsigned)ok << r;
                ra[r] = bcol[ok ? tri((size_t)mu) + nu : 0];
            }
        } else {
#pragma unroll
            for (int r = 0; r < NA; ++r) {
                const int e = tid + r * kBlock, mm = e % TM, k = e / TM;
                const int mu = mu0 + mm, nu = nu0 + k;
                const int lo = mu < nu ? mu : nu, hi = mu < nu ? nu : mu;
                const bool ok = mu < nbf && nu < nbf;
                live |= (unsigned)ok << r;
                ra[r] = bcol[ok ? tri((size_t)hi) + lo : 0];
            }
        }
#pragma unroll
        for (int r = 0; r < NB; ++r) {
            const int e = tid + r * kBlock, cc = e % TN, k = e / TN;
            rb[r] = cocc[(size_t)(nu0 + k) * nk + c0 + cc];   // C' rows are zero-padded to nbf_pad
        }
    };
    auto stash = [&](int buf, int nu0) {
        if (mu0 >= nu0 + TK - 1) {
#pragma unroll
            for (int r = 0; r < NA; ++r) {
                const int e = tid + r * kBlock;
                As[buf][e % TK][e / TK] = ((live >> r) & 1u) ? ra[r] : T(0);
            }
        } else {
#pragma unroll
            for (int r = 0; r < NA; ++r) {
                const int e = tid + r * kBlock;
                As[buf][e / TM][e % TM] = ((live >> r) & 1u) ? ra[r] : T(0);
            }
        }
#pragma unroll
        for (int r = 0; r < NB; ++r) {
            const int e = tid + r * kBlock;
            Bs[buf][e / TN][e % TN] = rb[r];
        }
    };

    typename MM::acc_t acc[FM][FN];
#pragma unroll
    for (int i = 0; i < FM; ++i)
#pragma unroll
        for (int j = 0; j < FN; ++j) acc[i][j] = MM::zero();

    const int nkb = (nbf + TK - 1) / TK;
    fetch(0);
    stash(0, 0);
    __syncthreads();
    for (int kb = 0; kb < nkb; ++kb) {
        const int cur = kb & 1;
        const bool more = kb + 1 < nkb;
        if (more) fetch((kb + 1) * TK);
#pragma unroll
        for (int kk = 0; kk < TK; kk += 4) {
            T bmu[FM], ac[FN];
#pragma unroll
            for (int i = 0; i < FM; ++i) bmu[i] = As[cur][kk + (lane >> 4)][wm * FM * 16 + i * 16 + (lane & 15)];
#pragma unroll
            for (int j = 0; j < FN; ++j) ac[j] = Bs[cur][kk + (lane >> 4)][wn * FN * 16 + j * 16 + (lane & 15)];
            // D[c][μ] = Σ_ν C'[ν][c] Bsym[ν][μ]
#pragma unroll
            for (int i = 0; i < FM; ++i)
#pragma unroll
                for (int j = 0; j < FN; ++j) acc[i][j] = MM::op(ac[j], bmu[i], acc[i][j]);
        }
        if (more) stash(cur ^ 1, (kb + 1) * TK);
        __syncthreads();
    }

#pragma unroll
    for (int i = 0; i < FM; ++i)
#pragma unroll
        for (int j = 0; j < FN; ++j) {
            const int mu = mu0 + wm * FM * 16 + i * 16 + (lane & 15);
            const int cb = c0 + wn * FN * 16 + j * 16;
#pragma unroll
            for (int r = 0; r < 4; ++r) mP[(size_t)(cb + MM::row(lane, r)) * ldm + mu] = acc[i][j][r];
        }
}


// ---------------------------------------------------------------------------
// Stage 2: Kpart (+)= Σ_{P,c} M'[P][c][μ] sign_c M'[P][c][ν] over the upper tiles.
//   1-D grid of nsplit × nspin × ntile(ntile+1)/2 work-groups, nsplit a
//   multiple of 8; split s runs entirely on XCD s % 8, so the M' slab of a
//   split is streamed into one L2 and shared by all tiles there.  Same
//   pipeline as stage 1 over the flattened (P, 16-column block) index;
//   accumulation in the MFMA type, partial tiles kept in FP64.
// ---------------------------------------------------------------------------
struct SyrkCols {
    int off[2], cnt[2];
};

template <typename T, int F, typename MP, typename SP, typename KP>
__device__ __forceinline__ void ksyrk_tile(MP __restrict__ m, int ldm, int nk, int mu0, int nu0, int cb, int ncb,
                                           int Pb, int Pe, SP __restrict__ sign, int beta, KP __restrict__ kp) {
    using MM = Mma<T>;
    constexpr int Tt = 2 * F * 16;
    constexpr int NL = TK * Tt / kBlock;
    __shared__ T As[2][TK][Tt + Pad<T>::v];
    __shared__ T Bs[2][TK][Tt + Pad<T>::v];
    __shared__ T Ss[2][TK];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = wave & 1, wn = wave >> 1;
    const int nq = (Pe > Pb ? Pe - Pb : 0) * ncb;
    T ra[NL], rb[NL], rsg = T(0);

    // raw loads only (the k-block's signs go to LDS and are applied to the
    // fragments), so the loads stay in flight across the current block's MFMAs
    auto fetch = [&](int q) {
        const int P = Pb + q / ncb, c = cb + (q % ncb) * TK;
        const auto base = m + ((size_t)P * nk + c) * ldm;
#pragma unroll
        for (int r = 0; r < NL; ++r) {
            const int e = tid + r * kBlock, mm = e % Tt, k = e / Tt;
            ra[r] = base[(size_t)k * ldm + mu0 + mm];
            rb[r] = base[(size_t)k * ldm + nu0 + mm];
        }
        if (tid < TK) rsg = sign[c + tid];
    };
    auto stash = [&](int buf) {
#pragma unroll
        for (int r = 0; r < NL; ++r) {
            const int e = tid + r * kBlock;
            As[buf][e / Tt][e % Tt] = ra[r];
            Bs[buf][e / Tt][e % Tt] = rb[r];
        }
        if (tid < TK) Ss[buf][tid] = rsg;
    };

    typename MM::acc_t acc[F][F];
#pragma unroll
    for (int i = 0; i < F; ++i)
#pragma unroll
        for (int j = 0; j < F; ++j) acc[i][j] = MM::zero();

    if (nq > 0) {
        fetch(0);
        stash(0);
    }
    __syncthreads();
    // branch-free body (one basic block keeps the accumulators in place):
    // the last iteration re-fetches its own block and stashes it into the
    // idle buffer, which nothing reads afterwards
    for (int q = 0; q < nq; ++q) {
        const int cur = q & 1;
        fetch(q + 1 < nq ? q + 1 : q);
#pragma unroll
        for (int kk = 0; kk < TK; kk += 4) {
            T bmu[F], anu[F];
            const T sg = Ss[cur][kk + (lane >> 4)];
#pragma unroll
            for (int i = 0; i < F; ++i) bmu[i] = As[cur][kk + (lane >> 4)][wm * F * 16 + i * 16 + (lane & 15)];
#pragma unroll
            for (int j = 0; j < F; ++j) anu[j] = sg * Bs[cur][kk + (lane >> 4)][wn * F * 16 + j * 16 + (lane & 15)];
            // D[ν][μ] = Σ_k sign_k M'[k][ν] M'[k][μ]
#pragma unroll
            for (int i = 0; i < F; ++i)
#pragma unroll
                for (int j = 0; j < F; ++j) acc[i][j] = MM::op(anu[j], bmu[i], acc[i][j]);
        }
        stash(cur ^ 1);
        __syncthreads();
    }

#pragma unroll
    for (int i = 0; i < F; ++i)
#pragma unroll
        for (int j = 0; j < F; ++j) {
            const int mu = mu0 + wm * F * 16 + i * 16 + (lane & 15);
            const int nb = nu0 + wn * F * 16 + j * 16;
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const auto dst = kp + (size_t)(nb + MM::row(lane, r)) * ldm + mu;
                double v = (double)acc[i][j][r];
                if (beta) v += *dst;
                *dst = v;
            }
        }
}


}  // namespace kt
}  // namespace rjk

// rjk_bftile.hpp — one work-group tile of D[x][y] += Σ_k X[x][k] Y[y][k] on the BF16
// matrix cores with operands split into S planes (both operands k-contiguous rows).
//
// Plane products: the pairs (a, b) with a + b < S, one v_mfma_f32_16x16x32_bf16 each
// (lane l supplies A[l%16][8(l/16)+i] = X and B[8(l/16)+i][l%16] = Yᵀ; register r of
// lane l receives D[4(l/16)+r][l%16], probed on gfx950).  Accumulation in FP32.
//
// A work-group of WX × WY waves covers WGX × WGY outputs; wave (wx, wy) owns
// FX × FY fragments of 16 × 16.  Per k-block of TK (32 or 64), every thread moves
// 16-byte row pieces of both operands from global memory into registers, then into
// LDS (one buffer, or two alternating with DB so that one barrier per k-block
// suffices); the next k-block's global loads are in flight while the MFMAs run.
// LDS rows are TK elements without padding; piece p of row r sits at piece
// p ^ sw(r), which keeps both the row-piece writes (consecutive rows) and the
// fragment reads (16 rows, one piece) free of bank conflicts (SQ_LDS_BANK_CONFLICT
// zero for both patterns in a probe on gfx950).
#pragma once
#include <hip/hip_runtime.h>

namespace rjk {
namespace bft {

typedef __bf16 bf8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));
typedef unsigned int u4 __attribute__((ext_vector_type(4)));

constexpr int KM = 32;        // one MFMA deep

// plane pair (X plane, Y plane) of product p
template <int S> __device__ constexpr int pa(int p);
template <int S> __device__ constexpr int pb(int p);
template <> __device__ constexpr int pa<1>(int) { return 0; }
template <> __device__ constexpr int pb<1>(int) { return 0; }
template <> __device__ constexpr int pa<2>(int p) { return p == 2 ? 1 : 0; }
template <> __device__ constexpr int pb<2>(int p) { return p == 1 ? 1 : 0; }
template <> __device__ constexpr int pa<3>(int p) { return p == 2 || p == 4 ? 1 : (p == 5 ? 2 : 0); }
template <> __device__ constexpr int pb<3>(int p) { return p == 1 || p == 4 ? 1 : (p == 3 ? 2 : 0); }
template <int S> constexpr int npairs() { return S * (S + 1) / 2; }

template <int S_, int WGX_, int WGY_, int WX_, int WY_, bool DB_, int TK_ = 32>
struct Cfg {
    static constexpr int S = S_, WGX = WGX_, WGY = WGY_, WX = WX_, WY = WY_, TK = TK_;
    static constexpr bool DB = DB_;
    static constexpr int NT = WX * WY * 64;
    static constexpr int NPC = TK / 8;                         // 16-byte pieces per row
    static constexpr int FX = WGX / WX / 16, FY = WGY / WY / 16;
    static constexpr int LX = WGX * NPC / NT, LY = WGY * NPC / NT;   // pieces per thread per plane
    static constexpr int NBUF = DB ? 2 : 1;
    static constexpr int LDS_ELEMS = NBUF * S * (WGX + WGY) * TK;
    static_assert(TK == 32 || TK == 64, "k-block");
    static_assert(FX * WX * 16 == WGX && FY * WY * 16 == WGY, "tile shape");
    static_assert(LX * NT == WGX * NPC && LY * NT == WGY * NPC, "load shape");
    // piece swizzle, one value per pair of rows: measured free of LDS bank conflicts on gfx950 for
    // both the fragment reads (16 rows, one piece) and the row-piece writes, where (r ^ r/8) % 8
    // costs 2 conflict cycles per 16-byte fragment read at 8 pieces a row (1.33 at 4)
    __device__ static constexpr int sw(int r) { return NPC == 4 ? ((r >> 1) & 3) : ((r >> 1) & 7); }
    // MFMA phase at raised wave priority: measured faster with 64-deep k-blocks (two-plane half-
    // transform 5.42 -> 5.08 ms, SYRK 2.65 -> 2.51 ms; three-plane half-transform 6.75 -> 6.64 ms) and
    // slower with 32-deep ones (three-plane SYRK 3.54 -> 3.91 ms)
    static constexpr bool PRIO = TK == 64;
};

__device__ inline f4 mfma(bf8 a, bf8 b, f4 c) { return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0); }

// 16-byte load that keeps the pointer's address space (a global pointer stays a global load)
__device__ __forceinline__ u4 ld16(const __bf16* p) { return *(const u4*)p; }
__device__ __forceinline__ u4 ld16(const __attribute__((address_space(1))) __bf16* p) {
    return *(const __attribute__((address_space(1))) u4*)p;
}

// Runs nq k-blocks.  xaddr(q) / yaddr(q): pointer to row 0, plane 0, first element of
// k-block q; ldx / ldy row strides, px / py plane strides (elements); xmask(q, col):
// the 8 sign masks (0x8000 flips) of the X piece at k-block column col, as a u4 of
// BF16 pairs.  With SPLIT the leading product (plane 0 × plane 0) accumulates in acc and
// the correction products in acc2, so that FP32 rounding of the accumulators acts on the
// corrections at their own (2^-8 smaller) scale; the caller adds the two in FP64.
template <class C, bool SPLIT, class XA, class YA, class XM>
__device__ __forceinline__ void nt_tile2(__bf16* __restrict__ lds, XA xaddr, size_t ldx, size_t px, YA yaddr,
                                         size_t ldy, size_t py, XM xmask, int nq, f4 (&acc)[C::FX][C::FY],
                                         f4 (&acc2)[C::FX][C::FY]) {
    constexpr int S = C::S, NP = npairs<S>(), TK = C::TK, NPC = C::NPC;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wx = wave % C::WX, wy = wave / C::WX;
    const int row = tid / NPC, piece = tid % NPC, pc = piece * 8;
    const int l16 = lane & 15;
    constexpr int RSTEP = C::NT / NPC;   // rows covered by one piece per thread
    // element offset of piece p of row r
    auto at = [](int r, int p) { return r * TK + ((p ^ C::sw(r)) << 3); };
    auto Xs = [&](int buf, int k) { return lds + (size_t)(buf * S + k) * (C::WGX + C::WGY) * TK; };
    auto Ys = [&](int buf, int k) { return lds + ((size_t)(buf * S + k) * (C::WGX + C::WGY) + C::WGX) * TK; };
#pragma unroll
    for (int i = 0; i < C::FX; ++i)
#pragma unroll
        for (int j = 0; j < C::FY; ++j) {
            acc[i][j] = f4{0.f, 0.f, 0.f, 0.f};
            acc2[i][j] = f4{0.f, 0.f, 0.f, 0.f};
        }
    u4 rx[S][C::LX], ry[S][C::LY], rm;
    auto fetch = [&](int q) {
        const auto xb = xaddr(q);
        const auto yb = yaddr(q);
#pragma unroll
        for (int k = 0; k < S; ++k) {
#pragma unroll
            for (int j = 0; j < C::LX; ++j) rx[k][j] = ld16(xb + (size_t)k * px + (size_t)(row + j * RSTEP) * ldx + pc);
#pragma unroll
            for (int j = 0; j < C::LY; ++j) ry[k][j] = ld16(yb + (size_t)k * py + (size_t)(row + j * RSTEP) * ldy + pc);
        }
        rm = xmask(q, pc);
    };
    auto stash = [&](int buf) {
#pragma unroll
        for (int k = 0; k < S; ++k) {
#pragma unroll
            for (int j = 0; j < C::LX; ++j) *(u4*)(Xs(buf, k) + at(row + j * RSTEP, piece)) = rx[k][j] ^ rm;
#pragma unroll
            for (int j = 0; j < C::LY; ++j) *(u4*)(Ys(buf, k) + at(row + j * RSTEP, piece)) = ry[k][j];
        }
    };
    auto compute = [&](int buf) {
#pragma unroll
        for (int ks = 0; ks < TK / KM; ++ks) {
            const int pr = ks * 4 + (lane >> 4);   // the piece this lane reads
            bf8 fx[C::FX][S], fy[C::FY][S];
#pragma unroll
            for (int k = 0; k < S; ++k) {
#pragma unroll
                for (int i = 0; i < C::FX; ++i) fx[i][k] = *(const bf8*)(Xs(buf, k) + at((wx * C::FX + i) * 16 + l16, pr));
#pragma unroll
                for (int j = 0; j < C::FY; ++j) fy[j][k] = *(const bf8*)(Ys(buf, k) + at((wy * C::FY + j) * 16 + l16, pr));
            }
            // products outermost: consecutive MFMAs write different accumulators
#pragma unroll
            for (int p = 0; p < NP; ++p)
#pragma unroll
                for (int i = 0; i < C::FX; ++i)
#pragma unroll
                    for (int j = 0; j < C::FY; ++j) {
                        if (SPLIT && p > 0)
                            acc2[i][j] = mfma(fx[i][pa<S>(p)], fy[j][pb<S>(p)], acc2[i][j]);
                        else
                            acc[i][j] = mfma(fx[i][pa<S>(p)], fy[j][pb<S>(p)], acc[i][j]);
                    }
        }
    };
    // with C::PRIO the MFMA phase runs at raised wave priority, so that a wave ready to issue MFMAs
    // goes before the other waves' loads and LDS stashes (results unchanged)
    auto compute_prio = [&](int buf) {
        if constexpr (C::PRIO) __builtin_amdgcn_s_setprio(1);
        compute(buf);
        if constexpr (C::PRIO) __builtin_amdgcn_s_setprio(0);
    };
    if (nq <= 0) return;
    fetch(0);
    if constexpr (C::DB) {
        stash(0);
        __syncthreads();
        for (int q = 0; q < nq; ++q) {
            const int cur = q & 1;
            if (q + 1 < nq) fetch(q + 1);
            compute_prio(cur);
            if (q + 1 < nq) stash(cur ^ 1);   // the other buffer: last read before the previous barrier
            __syncthreads();
        }
    } else {
        for (int q = 0; q < nq; ++q) {
            stash(0);
            __syncthreads();
            if (q + 1 < nq) fetch(q + 1);
            compute_prio(0);
            __syncthreads();
        }
    }
}

// One accumulator for every product.
template <class C, class XA, class YA, class XM>
__device__ __forceinline__ void nt_tile(__bf16* __restrict__ lds, XA xaddr, size_t ldx, size_t px, YA yaddr,
                                        size_t ldy, size_t py, XM xmask, int nq, f4 (&acc)[C::FX][C::FY]) {
    nt_tile2<C, false>(lds, xaddr, ldx, px, yaddr, ldy, py, xmask, nq, acc, acc);
}

}  // namespace bft
}  // namespace rjk

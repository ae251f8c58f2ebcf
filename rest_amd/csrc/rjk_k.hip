// rjk_k.hip — exchange (K) build on packed B, occupied-orbital route.
//
//   K_s = scale · upper( Σ_P M_P diag(occ_s) M_Pᵀ ),  M_P = B_P C_s
//
// which is vk_upper_with_ri_v_sync (src/scf_io/mod.rs:2758-2834) evaluated on
// the packed B of vk_upper_with_rimatr_use_dm_only_sync (:2703-2756).
//
// Columns: the occupied columns of all spins are concatenated (one B stream
// serves both spins); column c of spin s holds C[:,c]·sqrt|occ_c| and a sign
// sign_c = sgn(occ_c), so K = Σ_{P,c} sign_c M'_{P,c} M'_{P,c}ᵀ also for the
// signed weights of the density-only route (D = U Λ Uᵀ).
//
// Stage 1 (K-half): M'[P][c][μ] = Σ_ν B_P[μ,ν] C'[ν][c] — a batched FP64-MFMA
//   GEMM.  The A tile is read straight from the packed column and mirrored
//   through the triangle while it is staged into LDS.
// Stage 2 (K-SYRK): Kpart[split][s][ν][μ] (+)= Σ_{P,c} M'[P][c][μ] sign_c M'[P][c][ν]
//   over upper tile pairs (μ-tile <= ν-tile), split-K over P so that the
//   small nbf×nbf output still fills 256 CUs; deterministic (no atomics).
// Stage 3 (K-reduce): vk_s[p(i,j)] = scale · Σ_split Kpart[split][s][j][i].
//
// MFMA: v_mfma_f64_16x16x4f64.  Lane l supplies A[l%16][l/16] and
// B[l/16][l%16]; register r of lane l receives D[l/16 + 4r][l%16] (probed on
// gfx950).  Both kernels put the memory-contiguous index (μ) on the D column
// so that 16 consecutive lanes store 128 contiguous bytes.
#include "rjk_internal.hpp"

namespace rjk {
namespace {

typedef double d4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));

constexpr int kBlock = 256;           // 4 waves, 2×2
constexpr int TK = kTileK;            // 16
constexpr int kPadA = 4;              // LDS row padding (elements)
constexpr int kXcd = 8;               // MI355X: 8 XCDs, work-groups dealt round-robin by linear id

__device__ inline size_t tri(size_t j) { return j * (j + 1) / 2; }

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

// ---------------------------------------------------------------------------
// Stage 1: M'[P][c][μ] = Σ_ν Bsym_P[μ,ν] C'[ν][c]
//   1-D grid, XCD-aware: every work-group of auxiliary column P runs on XCD
//   P % 8 (the 12 μ-tiles × column tiles of one packed column share it in
//   that XCD's L2).  Software pipeline: the next k-block is fetched into
//   registers while the MFMAs consume the current one from LDS; LDS is
//   double-buffered, one barrier per k-block.
// ---------------------------------------------------------------------------
template <typename T, int FM, int FN>
__global__ __launch_bounds__(kBlock) void k_khalf(const T* __restrict__ b, size_t ldb, int nbf, size_t P0, int Pc,
                                                  int ntm, int ntc, const T* __restrict__ cocc, int nk, int ldm,
                                                  T* __restrict__ m) {
    using MM = Mma<T>;
    constexpr int TM = 2 * FM * 16, TN = 2 * FN * 16;
    constexpr int NA = TK * TM / kBlock, NB = TK * TN / kBlock;
    __shared__ T As[2][TK][TM + kPadA];
    __shared__ T Bs[2][TK][TN + kPadA];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = wave & 1, wn = wave >> 1;
    // XCD-aware decode of the linear work-group id
    const int w = blockIdx.x, xcd = w % kXcd, kq = w / kXcd, per = ntm * ntc;
    const int Pl = (kq / per) * kXcd + xcd;
    if (Pl >= Pc) return;                    // padding work-groups (uniform exit)
    const int rem = kq % per, mt = rem / ntc, ct = rem % ntc;
    const int mu0 = mt * TM, c0 = ct * TN;
    const T* __restrict__ bcol = b + (P0 + Pl) * ldb;
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

    T* __restrict__ mP = m + (size_t)Pl * nk * ldm;
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

template <typename T, int F>
__global__ __launch_bounds__(kBlock) void k_ksyrk(const T* __restrict__ m, int ldm, int nk, int Pc, int ntile2,
                                                  int nspin, SyrkCols cols, const T* __restrict__ sign,
                                                  int p_per_split, int beta, double* __restrict__ kpart,
                                                  size_t slot) {
    using MM = Mma<T>;
    constexpr int Tt = 2 * F * 16;
    constexpr int NL = TK * Tt / kBlock;
    __shared__ T As[2][TK][Tt + kPadA];
    __shared__ T Bs[2][TK][Tt + kPadA];
    __shared__ T Ss[2][TK];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = wave & 1, wn = wave >> 1;
    // XCD-aware decode: w -> (split on XCD w%8, spin, tile)
    const int w = blockIdx.x, xcd = w % kXcd, kq = w / kXcd;
    const size_t t = (size_t)(kq % ntile2);
    const int rest = kq / ntile2, spin = rest % nspin, split = (rest / nspin) * kXcd + xcd;
    // upper tile pair (tm <= tn) from the packed tile index
    size_t tn = (size_t)((sqrt(8.0 * (double)t + 1.0) - 1.0) * 0.5);
    while (tri(tn) > t) --tn;
    while (tri(tn + 1) <= t) ++tn;
    const size_t tm = t - tri(tn);
    const int mu0 = (int)tm * Tt, nu0 = (int)tn * Tt;
    const int cb = cols.off[spin];
    const int ncb = cols.cnt[spin] / TK;
    const int Pb = split * p_per_split, Pe = min(Pc, Pb + p_per_split);
    const int nq = (Pe > Pb ? Pe - Pb : 0) * ncb;
    T ra[NL], rb[NL], rsg = T(0);

    // raw loads only (the k-block's signs go to LDS and are applied to the
    // fragments), so the loads stay in flight across the current block's MFMAs
    auto fetch = [&](int q) {
        const int P = Pb + q / ncb, c = cb + (q % ncb) * TK;
        const T* __restrict__ base = m + ((size_t)P * nk + c) * ldm;
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

    double* __restrict__ kp = kpart + ((size_t)split * nspin + spin) * slot;
#pragma unroll
    for (int i = 0; i < F; ++i)
#pragma unroll
        for (int j = 0; j < F; ++j) {
            const int mu = mu0 + wm * F * 16 + i * 16 + (lane & 15);
            const int nb = nu0 + wn * F * 16 + j * 16;
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                double* dst = kp + (size_t)(nb + MM::row(lane, r)) * ldm + mu;
                double v = (double)acc[i][j][r];
                if (beta) v += *dst;
                *dst = v;
            }
        }
}

__global__ void k_kreduce(const double* __restrict__ kpart, int nsplit, int nspin, size_t slot, int ldm,
                          size_t npair, double scale, double* __restrict__ vk) {
    size_t p = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= npair) return;
    size_t j = (size_t)((sqrt(8.0 * (double)p + 1.0) - 1.0) * 0.5);
    while (tri(j) > p) --j;
    while (tri(j + 1) <= p) ++j;
    size_t i = p - tri(j);
    for (int s = 0; s < nspin; ++s) {
        double acc = 0;
        for (int q = 0; q < nsplit; ++q) acc += kpart[((size_t)q * nspin + s) * slot + j * ldm + i];
        vk[(size_t)s * npair + p] = acc * scale;
    }
}

// C'[ν][c] and sign[c] from the orbitals: column c of spin s is orbital
// (c - off_s) scaled by sqrt|occ|; padding columns and rows are zero.
template <typename T>
__global__ void k_prep_cocc(const double* __restrict__ eigv, int nbf, int nstate, int nw0, int nw1,
                            const double* __restrict__ occ, int nspin, int off0, int off1, int nk, int nrow,
                            T* __restrict__ cocc, T* __restrict__ sign) {
    size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (size_t)nrow * nk) return;
    int c = (int)(t % nk), nu = (int)(t / nk);
    double v = 0.0, sg = 0.0;
    int s = -1, o = 0;
    if (c >= off0 && c < off0 + nw0) { s = 0; o = c - off0; }
    else if (nspin == 2 && c >= off1 && c < off1 + nw1) { s = 1; o = c - off1; }
    if (s >= 0) {
        double w = occ[(size_t)s * nstate + o];
        sg = (w > 0) ? 1.0 : ((w < 0) ? -1.0 : 0.0);
        if (nu < nbf) v = eigv[(size_t)s * nbf * nstate + (size_t)o * nbf + nu] * sqrt(fabs(w));
    }
    cocc[t] = (T)v;
    if (nu == 0) sign[c] = (T)sg;
}

// Columns gathered by index (density-only route): column col_off+q is
// vecs[:, idx[q]] · sqrt|lam[q]|.
template <typename T>
__global__ void k_prep_cocc_idx(const double* __restrict__ vecs, int nbf, const int* __restrict__ idx,
                                const double* __restrict__ lam, int ncol, int col_off, int nk, int nrow,
                                T* __restrict__ cocc, T* __restrict__ sign) {
    size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (size_t)nrow * ncol) return;
    int q = (int)(t % ncol), nu = (int)(t / ncol);
    double w = lam[q];
    double v = (nu < nbf) ? vecs[(size_t)idx[q] * nbf + nu] * sqrt(fabs(w)) : 0.0;
    cocc[(size_t)nu * nk + col_off + q] = (T)v;
    if (nu == 0) sign[col_off + q] = (T)((w > 0) ? 1.0 : ((w < 0) ? -1.0 : 0.0));
}

template <typename T>
hipError_t khalf_t(hipStream_t st, const T* b, size_t ldb, size_t nbf, size_t P0, size_t Pc, const KPlan& kp,
                   const T* cocc, T* m) {
    constexpr int FM = 4, FN = 2;                       // 128 (μ) × 64 (c) tiles
    constexpr int TM = 2 * FM * 16, TN = 2 * FN * 16;
    const int ntm = (int)(kp.nbf_pad / TM), ntc = (int)(kp.nk / TN);
    const size_t nwg = round_up(Pc, kXcd) * (size_t)ntm * ntc;
    hipLaunchKernelGGL((k_khalf<T, FM, FN>), dim3((unsigned)nwg), dim3(kBlock), 0, st, b, ldb, (int)nbf, P0,
                       (int)Pc, ntm, ntc, cocc, (int)kp.nk, (int)kp.nbf_pad, m);
    return hipGetLastError();
}

template <typename T>
hipError_t ksyrk_t(hipStream_t st, const T* m, size_t Pc, const KPlan& kp, int nspin, const T* sign, int beta,
                   double* kpart) {
    constexpr int F = 4;                                // 128 × 128 tiles
    constexpr int Tt = 2 * F * 16;
    const size_t nt = kp.nbf_pad / Tt, nt2 = nt * (nt + 1) / 2;
    SyrkCols cols;
    cols.off[0] = (int)kp.col_off[0];
    cols.cnt[0] = (int)kp.col_cnt[0];
    cols.off[1] = (int)kp.col_off[1];
    cols.cnt[1] = (int)kp.col_cnt[1];
    const int pps = (int)((Pc + kp.nsplit - 1) / kp.nsplit);
    const size_t nwg = kp.nsplit * (size_t)nspin * nt2;   // nsplit is a multiple of 8
    hipLaunchKernelGGL((k_ksyrk<T, F>), dim3((unsigned)nwg), dim3(kBlock), 0, st, m, (int)kp.nbf_pad, (int)kp.nk,
                       (int)Pc, (int)nt2, nspin, cols, sign, pps, beta, kpart, kp.nbf_pad * kp.nbf_pad);
    return hipGetLastError();
}

}  // namespace

size_t k_elem_size(int prec) { return prec == RJK_PREC_FP64 ? sizeof(double) : sizeof(float); }

hipError_t launch_prep_cocc(hipStream_t st, int prec, const double* eigv, size_t nbf, size_t nstate,
                            const size_t* nw, const double* occ, int nspin, const KPlan& kp, void* cocc,
                            void* sign) {
    size_t n = kp.nbf_pad * kp.nk;
    dim3 g((unsigned)((n + 255) / 256));
    if (prec == RJK_PREC_FP64)
        hipLaunchKernelGGL(k_prep_cocc<double>, g, dim3(256), 0, st, eigv, (int)nbf, (int)nstate, (int)nw[0],
                           nspin == 2 ? (int)nw[1] : 0, occ, nspin, (int)kp.col_off[0], (int)kp.col_off[1],
                           (int)kp.nk, (int)kp.nbf_pad, (double*)cocc, (double*)sign);
    else
        hipLaunchKernelGGL(k_prep_cocc<float>, g, dim3(256), 0, st, eigv, (int)nbf, (int)nstate, (int)nw[0],
                           nspin == 2 ? (int)nw[1] : 0, occ, nspin, (int)kp.col_off[0], (int)kp.col_off[1],
                           (int)kp.nk, (int)kp.nbf_pad, (float*)cocc, (float*)sign);
    return hipGetLastError();
}

hipError_t launch_prep_cocc_idx(hipStream_t st, int prec, const double* vecs, size_t nbf, const int* idx,
                                const double* lam, size_t ncol, size_t col_off, const KPlan& kp, void* cocc,
                                void* sign) {
    if (ncol == 0) return hipSuccess;
    size_t n = kp.nbf_pad * ncol;
    dim3 g((unsigned)((n + 255) / 256));
    if (prec == RJK_PREC_FP64)
        hipLaunchKernelGGL(k_prep_cocc_idx<double>, g, dim3(256), 0, st, vecs, (int)nbf, idx, lam, (int)ncol,
                           (int)col_off, (int)kp.nk, (int)kp.nbf_pad, (double*)cocc, (double*)sign);
    else
        hipLaunchKernelGGL(k_prep_cocc_idx<float>, g, dim3(256), 0, st, vecs, (int)nbf, idx, lam, (int)ncol,
                           (int)col_off, (int)kp.nk, (int)kp.nbf_pad, (float*)cocc, (float*)sign);
    return hipGetLastError();
}

hipError_t launch_khalf(hipStream_t st, int prec, const void* b, size_t ldb, size_t nbf, size_t P0, size_t Pc,
                        const KPlan& kp, const void* cocc, void* m) {
    if (prec == RJK_PREC_FP64)
        return khalf_t<double>(st, (const double*)b, ldb, nbf, P0, Pc, kp, (const double*)cocc, (double*)m);
    return khalf_t<float>(st, (const float*)b, ldb, nbf, P0, Pc, kp, (const float*)cocc, (float*)m);
}

hipError_t launch_ksyrk(hipStream_t st, int prec, const void* m, size_t Pc, const KPlan& kp, int nspin,
                        const void* sign, int beta, double* kpart) {
    if (prec == RJK_PREC_FP64)
        return ksyrk_t<double>(st, (const double*)m, Pc, kp, nspin, (const double*)sign, beta, kpart);
    return ksyrk_t<float>(st, (const float*)m, Pc, kp, nspin, (const float*)sign, beta, kpart);
}

hipError_t launch_kreduce(hipStream_t st, const double* kpart, const KPlan& kp, int nspin, size_t nbf, double scale,
                          double* vk) {
    size_t npair = nbf * (nbf + 1) / 2;
    hipLaunchKernelGGL(k_kreduce, dim3((unsigned)((npair + 255) / 256)), dim3(256), 0, st, kpart, (int)kp.nsplit,
                       nspin, kp.nbf_pad * kp.nbf_pad, (int)kp.nbf_pad, npair, scale, vk);
    return hipGetLastError();
}

}  // namespace rjk

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
#include "rjk_ktiles.hpp"

namespace rjk {
namespace {

using namespace kt;

// ---------------------------------------------------------------------------
// Stage 1 kernel: 1-D grid, XCD-aware — every work-group of auxiliary column P
// runs on XCD P % 8, so the 12 μ-tiles × column tiles of one packed column
// share it in that XCD's L2.
// ---------------------------------------------------------------------------
template <typename T, int FM, int FN>
__global__ __launch_bounds__(kBlock) void k_khalf(const T* __restrict__ b, size_t ldb, int nbf, size_t P0, int Pc,
                                                  int ntm, int ntc, const T* __restrict__ cocc, int nk, int ldm,
                                                  T* __restrict__ m) {
    constexpr int TM = 2 * FM * 16, TN = 2 * FN * 16;
    // XCD-aware decode of the linear work-group id
    const int w = blockIdx.x, xcd = w % kXcd, kq = w / kXcd, per = ntm * ntc;
    const int Pl = (kq / per) * kXcd + xcd;
    if (Pl >= Pc) return;                    // padding work-groups (uniform exit)
    const int rem = kq % per, mt = rem / ntc, ct = rem % ntc;
    khalf_tile<T, FM, FN>(b + (P0 + Pl) * ldb, nbf, mt * TM, ct * TN, cocc, nk, ldm, m + (size_t)Pl * nk * ldm);
}

// ---------------------------------------------------------------------------
// Stage 2 kernel: 1-D grid of nsplit × nspin × ntile(ntile+1)/2 work-groups,
// nsplit a multiple of 8; split s runs entirely on XCD s % 8, so the M' slab
// of a split is streamed into one L2 and shared by all tiles there.
// ---------------------------------------------------------------------------
template <typename T, int F>
__global__ __launch_bounds__(kBlock) void k_ksyrk(const T* __restrict__ m, int ldm, int nk, int Pc, int ntile2,
                                                  int nspin, SyrkCols cols, const T* __restrict__ sign,
                                                  int p_per_split, int beta, double* __restrict__ kpart,
                                                  size_t slot) {
    constexpr int Tt = 2 * F * 16;
    // XCD-aware decode: w -> (split on XCD w%8, spin, tile)
    const int w = blockIdx.x, xcd = w % kXcd, kq = w / kXcd;
    const size_t t = (size_t)(kq % ntile2);
    const int rest = kq / ntile2, spin = rest % nspin, split = (rest / nspin) * kXcd + xcd;
    size_t tm, tn;
    upper_tile(t, &tm, &tn);
    const int Pb = split * p_per_split, Pe = min(Pc, Pb + p_per_split);
    ksyrk_tile<T, F>(m, ldm, nk, (int)tm * Tt, (int)tn * Tt, cols.off[spin], cols.cnt[spin] / TK, Pb, Pe, sign, beta,
                     kpart + ((size_t)split * nspin + spin) * slot);
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

size_t k_elem_size(int prec) {
    if (is_bf16(prec)) return 2 * (size_t)bf_planes(prec);
    return prec == RJK_PREC_FP64 ? sizeof(double) : sizeof(float);
}

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

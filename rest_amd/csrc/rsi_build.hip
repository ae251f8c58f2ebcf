// rsi_build.cpp — packed B = (μν|Q)·V^{-1/2} written straight into the
// HBM-resident storage of an rjk_ctx (Molecule::prepare_rimatr_for_ri_v_rayon_v02,
// src/molecule_io/mod.rs:1563-1721).
//
//   1. V = (P|Q) on the host, eigen-decomposed on the device (rocSOLVER dsyevd);
//      V^{-1/2} = W Wᵀ with W = U diag(λ^{-1/4}) over λ > threshold
//      (lapack_power(-0.5, 1e-8), :1582).
//   2. (μν|Q) in shell-aligned pair slabs on the host (OpenMP), uploaded, and
//      B[:, slab] = X_slab · V^{-1/2} with one rocBLAS dgemm per slab into
//      the padded [naux][ldb] layout.  Host work on slab k+1 overlaps the
//      upload and GEMM of slab k.
#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <vector>

#include "../../include/rest_ri.h"
#include "rjk_internal.hpp"
#include "rsi_gpu3c.hpp"
#include "rsi_internal.hpp"

namespace {

struct Dev {
    void* p = nullptr;
    ~Dev() {
        if (p) (void)hipFree(p);
    }
};
struct Host {
    void* p = nullptr;
    ~Host() {
        if (p) (void)hipHostFree(p);
    }
};
struct GpuInts {
    rsi::Gpu3c g;
    ~GpuInts() { rsi::gpu3c_free(&g); }
};
struct Blas {
    rocblas_handle h = nullptr;
    ~Blas() {
        if (h) rocblas_destroy_handle(h);
    }
};

double now() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

__global__ void k_scale_cols(double* U, int n, const double* w) {
    size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (size_t)n * n) return;
    U[t] *= w[t / n];   // column t / n (column-major)
}

}  // namespace

#define RSI_HIP(ctx, call)                                                                     \
    do {                                                                                       \
        hipError_t e_ = (call);                                                                \
        if (e_ != hipSuccess) {                                                                \
            rsi::set_error(std::string(#call) + ": " + hipGetErrorString(e_));                 \
            return rjk::fail((ctx), RJK_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
        }                                                                                      \
    } while (0)
#define RSI_BLAS(ctx, call)                                                        \
    do {                                                                           \
        rocblas_status s_ = (call);                                                \
        if (s_ != rocblas_status_success) {                                        \
            rsi::set_error(std::string(#call) + " failed: " + std::to_string((int)s_)); \
            return rjk::fail((ctx), RJK_ERR_HIP, std::string(#call) + " failed");  \
        }                                                                          \
    } while (0)

extern "C" int rsi_build_rimatr(rjk_ctx* ctx, const rsi_basis* orb, const rsi_basis* aux, double thr,
                                unsigned store_mask, size_t* nkept, double* seconds) {
    if (!ctx || !orb || !aux) {
        rsi::set_error("rsi_build_rimatr: NULL argument");
        return RJK_ERR_ARG;
    }
    if (orb->nbf != ctx->nbf || aux->nbf != ctx->naux) {
        rsi::set_error("rsi_build_rimatr: basis sizes do not match the context");
        return rjk::fail(ctx, RJK_ERR_ARG, "rsi_build_rimatr: basis sizes (" + std::to_string(orb->nbf) + ", " +
                                               std::to_string(aux->nbf) + ") do not match the context");
    }
    const size_t nbf = ctx->nbf, naux = ctx->naux, ldb = ctx->ldb;
    const hipStream_t st = ctx->stream;
    double t_3c = 0, t_v = 0, t_gemm = 0;
    RSI_HIP(ctx, hipSetDevice(ctx->device));
    int rc = rjk::alloc_b(ctx, store_mask);
    if (rc) return rc;
    Blas blas;
    RSI_BLAS(ctx, rocblas_create_handle(&blas.h));
    RSI_BLAS(ctx, rocblas_set_stream(blas.h, st));

    // ---- V^{-1/2} ----------------------------------------------------------
    double t0 = now();
    std::vector<double> V(naux * naux);
    rc = rsi_int2c2e(aux, V.data());
    if (rc) return rjk::fail(ctx, rc, rsi::last_error());
    Dev dU, dVm, dlam, dE, dinfo, dw;
    RSI_HIP(ctx, hipMalloc(&dU.p, naux * naux * sizeof(double)));
    RSI_HIP(ctx, hipMalloc(&dVm.p, naux * naux * sizeof(double)));
    RSI_HIP(ctx, hipMalloc(&dlam.p, naux * sizeof(double)));
    RSI_HIP(ctx, hipMalloc(&dE.p, naux * sizeof(double)));
    RSI_HIP(ctx, hipMalloc(&dw.p, naux * sizeof(double)));
    RSI_HIP(ctx, hipMalloc(&dinfo.p, sizeof(int)));
    RSI_HIP(ctx, hipMemcpyAsync(dU.p, V.data(), naux * naux * sizeof(double), hipMemcpyHostToDevice, st));
    RSI_BLAS(ctx, rocsolver_dsyevd(blas.h, rocblas_evect_original, rocblas_fill_upper, (rocblas_int)naux,
                                   (double*)dU.p, (rocblas_int)naux, (double*)dlam.p, (double*)dE.p,
                                   (rocblas_int*)dinfo.p));
    std::vector<double> lam(naux), w(naux);
    int info = 0;
    RSI_HIP(ctx, hipMemcpyAsync(lam.data(), dlam.p, naux * sizeof(double), hipMemcpyDeviceToHost, st));
    RSI_HIP(ctx, hipMemcpyAsync(&info, dinfo.p, sizeof(int), hipMemcpyDeviceToHost, st));
    RSI_HIP(ctx, hipStreamSynchronize(st));
    if (info != 0) return rjk::fail(ctx, RJK_ERR_NOSOLVER, "rocsolver_dsyevd(V) info = " + std::to_string(info));
    size_t kept = 0;
    for (size_t i = 0; i < naux; ++i) {
        if (lam[i] > thr) {
            w[i] = std::pow(lam[i], -0.25);
            ++kept;
        } else {
            w[i] = 0.0;
        }
    }
    if (nkept) *nkept = kept;
    RSI_HIP(ctx, hipMemcpyAsync(dw.p, w.data(), naux * sizeof(double), hipMemcpyHostToDevice, st));
    hipLaunchKernelGGL(k_scale_cols, dim3((unsigned)((naux * naux + 255) / 256)), dim3(256), 0, st, (double*)dU.p,
                       (int)naux, (const double*)dw.p);
    RSI_HIP(ctx, hipGetLastError());
    const double one_ = 1.0, zero_ = 0.0;
    RSI_BLAS(ctx, rocblas_dgemm(blas.h, rocblas_operation_none, rocblas_operation_transpose, (rocblas_int)naux,
                                (rocblas_int)naux, (rocblas_int)naux, &one_, (const double*)dU.p, (rocblas_int)naux,
                                (const double*)dU.p, (rocblas_int)naux, &zero_, (double*)dVm.p, (rocblas_int)naux));
    RSI_HIP(ctx, hipStreamSynchronize(st));
    (void)hipFree(dU.p);
    dU.p = nullptr;
    t_v = now() - t0;

    // ---- slabs of pair columns j, shell aligned, ≈ 256 MiB of (μν|Q) each --------
    std::vector<size_t> cuts{0};
    {
        const size_t budget = (size_t(256) << 20) / (naux * sizeof(double));   // pairs per slab
        size_t j0 = 0;
        for (const rsi::Shell& s : orb->sh) {
            const size_t jend = s.off + s.nsph;
            if (jend * (jend + 1) / 2 - j0 * (j0 + 1) / 2 > budget && (size_t)s.off > j0) {
                cuts.push_back(s.off);
                j0 = s.off;
            }
        }
        cuts.push_back(nbf);
    }
    size_t maxq = 0;
    for (size_t k = 0; k + 1 < cuts.size(); ++k)
        maxq = std::max(maxq, cuts[k + 1] * (cuts[k + 1] + 1) / 2 - cuts[k] * (cuts[k] + 1) / 2);
    RSI_HIP(ctx, hipMemsetAsync(ctx->b64, 0, naux * ldb * sizeof(double), st));
    const double one = 1.0, zero = 0.0;
    int status = RJK_OK;
    if (rsi::gpu3c_supported(*orb, *aux)) {
        // device integrals: slab k's kernels and GEMM follow each other on the stream
        GpuInts g;
        Dev dX;
        int r = rsi::gpu3c_init(&g.g, *orb, *aux, cuts, st);
        if (r) return rjk::fail(ctx, r, "rsi_build_rimatr: device integral set-up failed");
        RSI_HIP(ctx, hipMalloc(&dX.p, maxq * naux * sizeof(double)));
        std::vector<hipEvent_t> ev(2 * (cuts.size() - 1), nullptr);
        for (auto& e : ev) RSI_HIP(ctx, hipEventCreate(&e));
        for (size_t k = 0; k + 1 < cuts.size() && status == RJK_OK; ++k) {
            const size_t q0 = cuts[k] * (cuts[k] + 1) / 2, nq = cuts[k + 1] * (cuts[k + 1] + 1) / 2 - q0;
            (void)hipEventRecord(ev[2 * k], st);
            r = rsi::gpu3c_slab(g.g, k, st, (double*)dX.p, naux);
            (void)hipEventRecord(ev[2 * k + 1], st);
            if (r) {
                status = rjk::fail(ctx, r, "rsi_build_rimatr: three-centre kernel launch failed");
                break;
            }
            if (rocblas_dgemm(blas.h, rocblas_operation_none, rocblas_operation_none, (rocblas_int)nq,
                              (rocblas_int)naux, (rocblas_int)naux, &one, (const double*)dX.p, (rocblas_int)nq,
                              (const double*)dVm.p, (rocblas_int)naux, &zero, ctx->b64 + q0,
                              (rocblas_int)ldb) != rocblas_status_success)
                status = rjk::fail(ctx, RJK_ERR_HIP, "rocblas_dgemm(B slab) failed");
        }
        const double t1 = now();
        if (hipStreamSynchronize(st) != hipSuccess && status == RJK_OK)
            status = rjk::fail(ctx, RJK_ERR_HIP, "rsi_build_rimatr: device integrals / GEMM failed");
        if (status == RJK_OK)
            for (size_t k = 0; k + 1 < cuts.size(); ++k) {
                float ms = 0.f;
                if (hipEventElapsedTime(&ms, ev[2 * k], ev[2 * k + 1]) == hipSuccess) t_3c += 1e-3 * ms;
            }
        t_gemm = now() - t1;
        for (auto& e : ev) (void)hipEventDestroy(e);
    } else {
        // host integrals (orbital shells beyond d): slab k+1 on the host overlaps slab k's upload and GEMM
        Host hX[2];
        Dev dX[2];
        hipEvent_t done[2];
        for (int b = 0; b < 2; ++b) {
            RSI_HIP(ctx, hipHostMalloc(&hX[b].p, maxq * naux * sizeof(double), hipHostMallocDefault));
            RSI_HIP(ctx, hipMalloc(&dX[b].p, maxq * naux * sizeof(double)));
            RSI_HIP(ctx, hipEventCreateWithFlags(&done[b], hipEventDisableTiming));
            RSI_HIP(ctx, hipEventRecord(done[b], st));
        }
        for (size_t k = 0; k + 1 < cuts.size() && status == RJK_OK; ++k) {
            const int b = (int)(k & 1);
            const size_t j0 = cuts[k], j1 = cuts[k + 1];
            const size_t q0 = j0 * (j0 + 1) / 2, nq = j1 * (j1 + 1) / 2 - q0;
            if (hipEventSynchronize(done[b]) != hipSuccess) {   // host buffer b free again
                status = rjk::fail(ctx, RJK_ERR_HIP, "hipEventSynchronize failed");
                break;
            }
            double t1 = now();
            rsi::int3c_slab(*orb, *aux, j0, j1, (double*)hX[b].p);
            t_3c += now() - t1;
            if (hipMemcpyAsync(dX[b].p, hX[b].p, nq * naux * sizeof(double), hipMemcpyHostToDevice, st) !=
                hipSuccess) {
                status = rjk::fail(ctx, RJK_ERR_HIP, "hipMemcpyAsync(X slab) failed");
                break;
            }
            // B[P][q0+q] = Σ_Q X[Q][q] Vm[Q][P]   (column-major: Bm(nq × naux, ld ldb) = Xm(nq × naux) Vm)
            if (rocblas_dgemm(blas.h, rocblas_operation_none, rocblas_operation_none, (rocblas_int)nq,
                              (rocblas_int)naux, (rocblas_int)naux, &one, (const double*)dX[b].p, (rocblas_int)nq,
                              (const double*)dVm.p, (rocblas_int)naux, &zero, ctx->b64 + q0,
                              (rocblas_int)ldb) != rocblas_status_success) {
                status = rjk::fail(ctx, RJK_ERR_HIP, "rocblas_dgemm(B slab) failed");
                break;
            }
            if (hipEventRecord(done[b], st) != hipSuccess) {
                status = rjk::fail(ctx, RJK_ERR_HIP, "hipEventRecord failed");
                break;
            }
        }
        double t2 = now();
        (void)hipStreamSynchronize(st);
        t_gemm = now() - t2;   // tail of the device work after the last host slab
        for (int b = 0; b < 2; ++b) (void)hipEventDestroy(done[b]);
    }
    if (status != RJK_OK) {
        rsi::set_error(ctx->err);
        return status;
    }
    rc = rjk::finish_b(ctx);
    if (seconds) {
        seconds[0] = t_3c;
        seconds[1] = t_v;
        seconds[2] = t_gemm;
    }
    return rc;
}

extern "C" int rsi_int3c2e_packed_dev(rjk_ctx* ctx, const rsi_basis* orb, const rsi_basis* aux, size_t j0, size_t j1,
                                      double* out_dev) {
    if (!ctx || !orb || !aux || !out_dev) {
        rsi::set_error("rsi_int3c2e_packed_dev: NULL argument");
        return RJK_ERR_ARG;
    }
    if (j0 >= j1 || j1 > orb->nbf || !rsi::on_shell_boundary(*orb, j0) || !rsi::on_shell_boundary(*orb, j1)) {
        rsi::set_error("rsi_int3c2e_packed_dev: [j0, j1) must be a non-empty range on shell boundaries");
        return RJK_ERR_ARG;
    }
    if (!rsi::gpu3c_supported(*orb, *aux)) {
        rsi::set_error("rsi_int3c2e_packed_dev: device kernels cover orbital shells up to d and auxiliary up to g");
        return RJK_ERR_ARG;
    }
    if (hipSetDevice(ctx->device) != hipSuccess) {
        rsi::set_error("rsi_int3c2e_packed_dev: hipSetDevice failed");
        return RJK_ERR_HIP;
    }
    GpuInts g;
    int rc = rsi::gpu3c_init(&g.g, *orb, *aux, std::vector<size_t>{j0, j1}, ctx->stream);
    if (!rc) rc = rsi::gpu3c_slab(g.g, 0, ctx->stream, out_dev, aux->nbf);
    if (!rc && hipStreamSynchronize(ctx->stream) != hipSuccess) rc = RJK_ERR_HIP;
    if (rc) rsi::set_error("rsi_int3c2e_packed_dev: device integrals failed (code " + std::to_string(rc) + ")");
    return rc;
}

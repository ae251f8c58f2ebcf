// rscf.hip — device-resident closed-shell RI-HF SCF driver (include/rest_scf.h).
//
// One J/K build per iteration through the context's resident B; Fock,
// energy, DIIS and the density in HBM (FP64, rocBLAS GEMMs and small
// kernels); F C = S C ε through X = S^{-1/2} and rocSOLVER dsyevd.  Mirrors REST's scf()
// (src/scf_io/mod.rs:3190-3268) with prepare_next_input (:2947-3091),
// generate_diis_error_vector (:3094-3136), diis_solver (:3138-3188) and
// check_scf_convergence (:1875-1951).
#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <vector>

#include "../../include/rest_scf.h"
#include "rjk_internal.hpp"

namespace {

constexpr int kRedBlocks = 1024;
constexpr double kGapMin = 0.05;   // HOMO-LUMO gap below which pseudo-diagonalisation is not used (Eh)

double now() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

__device__ inline size_t pk(size_t i, size_t j) { return j >= i ? j * (j + 1) / 2 + i : i * (i + 1) / 2 + j; }

// F = H + J + sK·K  (J, K packed upper -> full symmetric)
__global__ void k_fock(const double* __restrict__ H, const double* __restrict__ vj, const double* __restrict__ vk,
                       size_t n, double sk, double* __restrict__ F) {
    size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n * n) return;
    const size_t i = t % n, j = t / n, p = pk(i, j);
    F[t] = H[t] + vj[p] + sk * vk[p];
}

__device__ inline double block_sum(double v) {
    __shared__ double sh[256 / 64];
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 0) sh[w] = v;
    __syncthreads();
    double s = 0.0;
    if (threadIdx.x == 0)
        for (int k = 0; k < 256 / 64; ++k) s += sh[k];
    return s;
}
__device__ inline double block_max(double v) {
    __shared__ double sh[256 / 64];
    for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off, 64));
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 0) sh[w] = v;
    __syncthreads();
    double s = 0.0;
    if (threadIdx.x == 0)
        for (int k = 0; k < 256 / 64; ++k) s = fmax(s, sh[k]);
    return s;
}

// partial[b] = Σ D (H + F)
__global__ void k_energy(const double* __restrict__ D, const double* __restrict__ H, const double* __restrict__ F,
                         size_t n2, double* __restrict__ partial) {
    double acc = 0.0;
    for (size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x; t < n2; t += (size_t)gridDim.x * blockDim.x)
        acc += D[t] * (H[t] + F[t]);
    const double s = block_sum(acc);
    if (threadIdx.x == 0) partial[blockIdx.x] = s;
}

// e = T − Tᵀ (T = F D S, so Tᵀ = S D F); partial[b] = max |e|
__global__ void k_commutator(const double* __restrict__ T, size_t n, double* __restrict__ e,
                             double* __restrict__ partial) {
    double m = 0.0;
    for (size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x; t < n * n; t += (size_t)gridDim.x * blockDim.x) {
        const size_t i = t % n, j = t / n;
        const double v = T[t] - T[i * n + j];
        e[t] = v;
        m = fmax(m, fabs(v));
    }
    const double s = block_max(m);
    if (threadIdx.x == 0) partial[blockIdx.x] = s;
}

// e += T − Tᵀ (n × n, column-major)
__global__ void k_commutator_acc(const double* __restrict__ T, size_t n, double* __restrict__ e) {
    for (size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x; t < n * n; t += (size_t)gridDim.x * blockDim.x) {
        const size_t i = t % n, j = t / n;
        e[t] += T[t] - T[i * n + j];
    }
}

// y += a·x
__global__ void k_axpy_inplace(double* __restrict__ y, const double* __restrict__ x, double a, size_t n2) {
    for (size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x; t < n2; t += (size_t)gridDim.x * blockDim.x)
        y[t] += a * x[t];
}

// Columns of ΔD = occ·(A Aᵀ − R Rᵀ) over the first nocc columns of A and R (ld n):
// P = A_i + R_i, M = A_i − R_i, u = a P + b M, v = a P − b M with a b = ½ and
// a²|P|² = b²|M|², so Σ_i (u uᵀ − v vᵀ) = Σ_i (P Mᵀ + M Pᵀ) = 2 (A Aᵀ − R Rᵀ) and
// |u|², |v|² <= 2|P||M|: K of [u | v] with weights ±occ/2 at a low precision errs in
// proportion to |A − R|, not to |A|.  u → UV[:, i], v → UV[:, nocc + i]; one block per column.
__global__ void k_inc_columns(const double* __restrict__ A, const double* __restrict__ R, size_t n, size_t nocc,
                              double* __restrict__ UV) {
    __shared__ double sp[256], sm[256];
    const size_t i = blockIdx.x;
    const double* a_col = A + i * n;
    const double* r_col = R + i * n;
    double pp = 0.0, mm = 0.0;
    for (size_t r = threadIdx.x; r < n; r += blockDim.x) {
        const double p = a_col[r] + r_col[r], m = a_col[r] - r_col[r];
        pp += p * p;
        mm += m * m;
    }
    sp[threadIdx.x] = pp;
    sm[threadIdx.x] = mm;
    __syncthreads();
    for (unsigned w = blockDim.x / 2; w > 0; w >>= 1) {
        if (threadIdx.x < w) {
            sp[threadIdx.x] += sp[threadIdx.x + w];
            sm[threadIdx.x] += sm[threadIdx.x + w];
        }
        __syncthreads();
    }
    const double pn = sqrt(sp[0]), mn = sqrt(sm[0]);
    double ca = 0.0, cb = 0.0;   // P = 0 or M = 0: the column adds nothing to P Mᵀ + M Pᵀ
    if (pn > 0.0 && mn > 0.0) {
        ca = sqrt(mn / (2.0 * pn));
        cb = 0.5 / ca;
    }
    for (size_t r = threadIdx.x; r < n; r += blockDim.x) {
        const double p = a_col[r] + r_col[r], m = a_col[r] - r_col[r];
        UV[r + i * n] = ca * p + cb * m;
        UV[r + (nocc + i) * n] = ca * p - cb * m;
    }
}

// partial[b] = max |x|  and  max |a − b|
__global__ void k_maxabs(const double* __restrict__ x, size_t n, double* __restrict__ partial) {
    double m = 0.0;
    for (size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (size_t)gridDim.x * blockDim.x)
        m = fmax(m, fabs(x[t]));
    const double r = block_max(m);
    if (threadIdx.x == 0) partial[blockIdx.x] = r;
}
__global__ void k_maxabs_diff(const double* __restrict__ a, const double* __restrict__ b, size_t n,
                              double* __restrict__ partial) {
    double m = 0.0;
    for (size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (size_t)gridDim.x * blockDim.x)
        m = fmax(m, fabs(a[t] - b[t]));
    const double r = block_max(m);
    if (threadIdx.x == 0) partial[blockIdx.x] = r;
}

// out = A − B
__global__ void k_sub(const double* __restrict__ A, const double* __restrict__ B, size_t n2, double* __restrict__ out) {
    for (size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x; t < n2; t += (size_t)gridDim.x * blockDim.x)
        out[t] = A[t] - B[t];
}

// U = I + X with X[a][i] = F_ai / (ε_i − ε_a), X[i][a] = −X[a][i] for i < nocc <= a
// (first-order rotation of the occupied orbitals into the virtual space); ε = diag(Fmo),
// |ε_i − ε_a| clamped below by gmin.  Column-major n × n.
__global__ void k_pseudo_rot(const double* __restrict__ Fmo, size_t n, size_t nocc, double gmin,
                             double* __restrict__ U) {
    for (size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x; t < n * n; t += (size_t)gridDim.x * blockDim.x) {
        const size_t r = t % n, c = t / n;
        double v = r == c ? 1.0 : 0.0;
        if (c < nocc && r >= nocc) {          // X[a = r][i = c]
            const double den = fmin(Fmo[c * (n + 1)] - Fmo[r * (n + 1)], -gmin);
            v = Fmo[r + c * n] / den;
        } else if (r < nocc && c >= nocc) {   // X[i = r][a = c] = −X[a][i]
            const double den = fmin(Fmo[r * (n + 1)] - Fmo[c * (n + 1)], -gmin);
            v = -Fmo[c + r * n] / den;
        }
        U[t] = v;
    }
}

__global__ void k_diag_copy(const double* __restrict__ M, size_t n, double* __restrict__ d) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        d[i] = M[i * (n + 1)];
}

// y = x − y
__global__ void k_delta_inplace(const double* __restrict__ x, double* __restrict__ y, size_t n2) {
    for (size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x; t < n2; t += (size_t)gridDim.x * blockDim.x)
        y[t] = x[t] - y[t];
}

// partial[b] = Σ (A − B)²
__global__ void k_sumsq_diff(const double* __restrict__ A, const double* __restrict__ B, size_t n2,
                             double* __restrict__ partial) {
    double acc = 0.0;
    for (size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x; t < n2; t += (size_t)gridDim.x * blockDim.x) {
        const double d = A[t] - B[t];
        acc += d * d;
    }
    const double s = block_sum(acc);
    if (threadIdx.x == 0) partial[blockIdx.x] = s;
}

struct Coefs {
    double c[16];
    const double* f[16];
};
// out = Σ_k c_k F_k
__global__ void k_lincomb(Coefs cf, int m, size_t n2, double* __restrict__ out) {
    for (size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x; t < n2; t += (size_t)gridDim.x * blockDim.x) {
        double v = 0.0;
        for (int k = 0; k < m; ++k) v += cf.c[k] * cf.f[k][t];
        out[t] = v;
    }
}

// out = A + α (B − A)
__global__ void k_mix(const double* __restrict__ A, const double* __restrict__ B, double alpha, size_t n2,
                      double* __restrict__ out) {
    size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n2) out[t] = A[t] + alpha * (B[t] - A[t]);
}

// out = prev·(1 − α) + new·α, REST's form of the energy-rise mixing (src/scf_io/mod.rs:3045-3047):
// α = 1 gives new and α = 0 gives prev exactly
__global__ void k_blend(const double* __restrict__ prev, const double* __restrict__ cur, double alpha, size_t n2,
                        double* __restrict__ out) {
    size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n2) out[t] = prev[t] * (1.0 - alpha) + cur[t] * alpha;
}

// out[:, j] = U[:, j] · w[j]   (column-major n × n)
__global__ void k_scale_columns(const double* __restrict__ U, const double* __restrict__ w, size_t n,
                                double* __restrict__ out) {
    size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n * n) out[t] = U[t] * w[t / n];
}

__global__ void k_add(const double* __restrict__ A, const double* __restrict__ B, size_t n2, double* __restrict__ out) {
    size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n2) out[t] = A[t] + B[t];
}

__global__ void k_fill(double* p, size_t n, double v) {
    size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) p[t] = v;
}

// Gauss–Jordan inverse with partial pivoting of the small DIIS matrix; false when singular.
bool small_inverse(std::vector<double> a, int n, std::vector<double>* inv) {
    inv->assign((size_t)n * n, 0.0);
    for (int i = 0; i < n; ++i) (*inv)[(size_t)i * n + i] = 1.0;
    double scale = 0.0;
    for (double v : a) scale = std::max(scale, std::fabs(v));
    if (!(scale > 0) || !std::isfinite(scale)) return false;
    for (int c = 0; c < n; ++c) {
        int piv = c;
        for (int r = c + 1; r < n; ++r)
            if (std::fabs(a[(size_t)r * n + c]) > std::fabs(a[(size_t)piv * n + c])) piv = r;
        if (std::fabs(a[(size_t)piv * n + c]) <= 1e-300 || std::fabs(a[(size_t)piv * n + c]) < 1e-15 * scale * 1e-14)
            return false;
        if (piv != c)
            for (int k = 0; k < n; ++k) {
                std::swap(a[(size_t)c * n + k], a[(size_t)piv * n + k]);
                std::swap((*inv)[(size_t)c * n + k], (*inv)[(size_t)piv * n + k]);
            }
        const double d = a[(size_t)c * n + c];
        for (int k = 0; k < n; ++k) {
            a[(size_t)c * n + k] /= d;
            (*inv)[(size_t)c * n + k] /= d;
        }
        for (int r = 0; r < n; ++r) {
            if (r == c) continue;
            const double f = a[(size_t)r * n + c];
            if (f == 0.0) continue;
            for (int k = 0; k < n; ++k) {
                a[(size_t)r * n + k] -= f * a[(size_t)c * n + k];
                (*inv)[(size_t)r * n + k] -= f * (*inv)[(size_t)c * n + k];
            }
        }
    }
    for (double v : *inv)
        if (!std::isfinite(v)) return false;
    return true;
}

// SCF buffers carved from the context's grow-only arena (ctx->scf_ws), 256-byte aligned.
// A call whose buffers do not fit allocates the rest one by one and, at its end, grows the
// arena to its whole demand, so later calls of that size allocate nothing.  Calls on one
// context are serialised (one host thread, one stream), so the arena is never shared.
struct DevMem {
    rjk_ctx* ctx;
    size_t used = 0;                   // bytes requested so far
    std::vector<void*> extra;          // buffers beyond the arena
    explicit DevMem(rjk_ctx* c) : ctx(c) {}
    ~DevMem() {
        for (void* p : extra) (void)hipFree(p);
        if (used > ctx->scf_ws_bytes) {   // grow for the next call (hipFree waits for the device)
            if (ctx->scf_ws) (void)hipFree(ctx->scf_ws);
            ctx->scf_ws = nullptr;
            ctx->scf_ws_bytes = 0;
            if (hipMalloc(&ctx->scf_ws, used) == hipSuccess) ctx->scf_ws_bytes = used;
            else (void)hipGetLastError();
        }
    }
    double* get(rjk_ctx*, size_t n, int* rc) {
        const size_t bytes = (n * sizeof(double) + 255) / 256 * 256;
        const size_t off = used;
        used += bytes;
        if (used <= ctx->scf_ws_bytes) return (double*)((char*)ctx->scf_ws + off);
        void* p = nullptr;
        if (hipMalloc(&p, bytes) != hipSuccess) {
            (void)hipGetLastError();
            *rc = rjk::fail(ctx, RJK_ERR_NOMEM, "SCF device allocation failed");
            return nullptr;
        }
        extra.push_back(p);
        return (double*)p;
    }
};

}  // namespace

#define SCF_HIP(call)                                                                            \
    do {                                                                                         \
        hipError_t e_ = (call);                                                                  \
        if (e_ != hipSuccess) return rjk::fail(ctx, RJK_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
    } while (0)
#define SCF_BLAS(call)                                                                                    \
    do {                                                                                                  \
        rocblas_status s_ = (call);                                                                       \
        if (s_ != rocblas_status_success)                                                                 \
            return rjk::fail(ctx, RJK_ERR_HIP, std::string(#call) + " failed (" + std::to_string((int)s_) + ")"); \
    } while (0)
#define SCF_RC(expr)          \
    do {                      \
        int r_ = (expr);      \
        if (r_) return r_;    \
    } while (0)

extern "C" void rscf_default_ctrl(rscf_ctrl* c) {
    if (!c) return;
    c->num_max_diis = 8;
    c->start_diis_cycle = 1;
    c->mix_param = 0.6;
    c->max_scf_cycle = 100;
    c->scf_acc_rho = 1e-5;
    c->scf_acc_eev = 1e-5;
    c->scf_acc_etot = 1e-8;
    c->jk_prec = RJK_PREC_FP64;
    c->switch_diis_err = 1e-5;
    c->print_level = 0;
    c->switch_mode = RSCF_SWITCH_RESTART;
    c->jk_prec_start = -1;
    c->start_switch_err = 1e-3;
    c->guess_energy_max = HUGE_VAL;
    c->pseudo_diag_err = 0.0;
    c->jk_prec_increment = -1;
    c->guess_err_max = HUGE_VAL;
    c->start_check_oscillation = 20;
    c->inc_refresh = 1e-5;
    c->inc_check = 0;
}

namespace {

// One SCF driver for RHF (nspin 1: nel[0] electrons, occupation 2) and UHF
// (nspin 2: nel[0] α and nel[1] β electrons, occupation 1).  With guess_only
// it performs REST's density-guess step only: F(D0) by the density-only J/K,
// E(D0), F C = S C ε and the new density, written to dm_out.
int scf_core(rjk_ctx* ctx, const char* who, const double* S_h, const double* H_h, double enuc, int nspin,
             const size_t nel[2], const double* d_guess, const rscf_ctrl* ctrl_in, bool guess_only,
             rscf_result* result, double* mo_coeff, double* mo_energy, double* dm_out, double* iter_log,
             bool energy_only = false, const double* c_guess = nullptr, double* energy_d1 = nullptr,
             const double* c0_guess = nullptr, const double* w0_guess = nullptr, const size_t* ncol0 = nullptr) {
    if (!S_h || !H_h) return rjk::fail(ctx, RJK_ERR_ARG, std::string(who) + ": S and H are required");
    rscf_ctrl ctrl;
    rscf_default_ctrl(&ctrl);
    if (ctrl_in) ctrl = *ctrl_in;
    const size_t n = ctx->nbf, n2 = n * n, npair = ctx->npair;
    size_t nocc[2] = {0, 0};
    double occ_val, sk;
    if (nspin == 1) {
        if (nel[0] == 0 || (nel[0] & 1) || nel[0] / 2 > n)
            return rjk::fail(ctx, RJK_ERR_ARG, std::string(who) + ": nelec must be even, positive and <= 2 nbf");
        nocc[0] = nel[0] / 2;
        occ_val = 2.0;
        sk = -0.5;   // REST scaling for RHF exchange (src/scf_io/mod.rs:1347-1350)
    } else {
        if (nel[0] + nel[1] == 0 || nel[0] > n || nel[1] > n)
            return rjk::fail(ctx, RJK_ERR_ARG, std::string(who) + ": need 0 < nalpha + nbeta and each <= nbf");
        nocc[0] = nel[0];
        nocc[1] = nel[1];
        occ_val = 1.0;
        sk = -1.0;
    }
    if (ctrl.num_max_diis < 1 || ctrl.num_max_diis > 16 || ctrl.max_scf_cycle < 1)
        return rjk::fail(ctx, RJK_ERR_ARG, std::string(who) + ": num_max_diis must be 1..16 and max_scf_cycle >= 1");
    if (!ctx->b64) return rjk::fail(ctx, RJK_ERR_STATE, std::string(who) + ": B is not set");
    if (ctrl.jk_prec != RJK_PREC_FP64 && ctrl.jk_prec != RJK_PREC_FP32 && !rjk::is_bf16(ctrl.jk_prec))
        return rjk::fail(ctx, RJK_ERR_ARG, std::string(who) + ": jk_prec must be FP64, FP32 or a BF16 precision");
    if (ctrl.jk_prec_start != -1 && ctrl.jk_prec_start != RJK_PREC_FP32 && !rjk::is_bf16(ctrl.jk_prec_start))
        return rjk::fail(ctx, RJK_ERR_ARG, std::string(who) + ": jk_prec_start must be -1, FP32 or a BF16 precision");
    if (ctrl.jk_prec_increment != -1 && ctrl.jk_prec_increment != RJK_PREC_FP32 && !rjk::is_bf16(ctrl.jk_prec_increment))
        return rjk::fail(ctx, RJK_ERR_ARG, std::string(who) + ": jk_prec_increment must be -1, FP32 or a BF16 precision");
    for (int pr : {ctrl.jk_prec, ctrl.jk_prec_start, ctrl.jk_prec_increment}) {
        if (pr == RJK_PREC_FP32 && !ctx->b32)
            return rjk::fail(ctx, RJK_ERR_STATE,
                             std::string(who) + ": mixed precision needs the FP32 copy of B (RJK_STORE_FP32)");
        if (rjk::is_bf16(pr) && !ctx->bpl)
            return rjk::fail(ctx, RJK_ERR_STATE,
                             std::string(who) + ": BF16 mixed precision needs the BF16 planes of B (RJK_STORE_BF16)");
    }
    if (guess_only && !d_guess) return rjk::fail(ctx, RJK_ERR_ARG, std::string(who) + ": a guess density is required");
    if (d_guess && c_guess) return rjk::fail(ctx, RJK_ERR_ARG, std::string(who) + ": give a guess density or orbitals");
    const double t_start = now();
    SCF_HIP(hipSetDevice(ctx->device));
    const hipStream_t st = ctx->stream;
    rocblas_handle h = nullptr;
    SCF_RC(rjk::solver_handle(ctx, &h));

    int rc = RJK_OK;
    DevMem mem(ctx);
    double* S = mem.get(ctx, n2, &rc);
    double* H = mem.get(ctx, n2, &rc);
    double* Bw = mem.get(ctx, n2, &rc);     // eigensolver operand
    double* T1 = mem.get(ctx, n2, &rc);
    double* T2 = mem.get(ctx, n2, &rc);
    double* Dt = mem.get(ctx, n2, &rc);     // total density for J (UHF)
    double* Ework = mem.get(ctx, n, &rc);
    double* vj = mem.get(ctx, npair, &rc);
    double* partial = mem.get(ctx, kRedBlocks, &rc);
    double* info_d = mem.get(ctx, 1, &rc);
    // per spin: Fock, extrapolated Fock, eigenvectors, densities (current, previous,
    // mixed), eigenvalues, occupations, exchange
    double *F[2], *Fx[2], *A[2], *D[2], *Dp[2], *Dm[2], *W[2], *occd[2], *vk[2], *Fl[2], *gw[2] = {nullptr, nullptr};
    for (int s = 0; s < nspin; ++s) {
        F[s] = mem.get(ctx, n2, &rc);
        Fl[s] = mem.get(ctx, n2, &rc);      // low-precision Fock at the switch
        Fx[s] = mem.get(ctx, n2, &rc);
        A[s] = mem.get(ctx, n2, &rc);
        D[s] = mem.get(ctx, n2, &rc);
        Dp[s] = mem.get(ctx, n2, &rc);
        Dm[s] = mem.get(ctx, n2, &rc);
        W[s] = mem.get(ctx, n, &rc);
        occd[s] = mem.get(ctx, n, &rc);
        vk[s] = mem.get(ctx, npair, &rc);
        if (c0_guess) gw[s] = mem.get(ctx, n, &rc);   // weights of the factored guess density
    }
    const int nd = ctrl.num_max_diis;
    const size_t ev_len = (size_t)nspin * n2;   // error vector: spins back to back (REST :3133)
    double* ering = mem.get(ctx, (size_t)nd * ev_len, &rc);
    double* fring = mem.get(ctx, (size_t)nd * ev_len, &rc);
    double* dring = mem.get(ctx, (size_t)nd * ev_len, &rc);   // the densities of the stored Fock matrices
    // incremental builds after the FP64 switch: the reference build's total density, J and
    // per-spin K and occupied orbitals; per spin the [u | v] columns of ΔD and their weights
    const bool inc_enabled = ctrl.jk_prec_increment >= 0;
    double *Dref = nullptr, *vjref = nullptr, *Dinc = nullptr;
    double *Cref[2] = {nullptr, nullptr}, *vkref[2] = {nullptr, nullptr}, *UV[2] = {nullptr, nullptr},
           *occ2[2] = {nullptr, nullptr};
    if (inc_enabled) {
        Dref = mem.get(ctx, n2, &rc);
        Dinc = mem.get(ctx, n2, &rc);
        vjref = mem.get(ctx, npair, &rc);
        for (int s = 0; s < nspin; ++s) {
            const size_t no = std::max<size_t>(nocc[s], 1);
            Cref[s] = mem.get(ctx, n * no, &rc);
            UV[s] = mem.get(ctx, 2 * n * no, &rc);
            occ2[s] = mem.get(ctx, 2 * no, &rc);
            vkref[s] = mem.get(ctx, npair, &rc);
        }
    }
    // inc_check: a full FP64 J/K of the same density beside every incremental build
    double *vjc = nullptr, *vkc[2] = {nullptr, nullptr};
    if (inc_enabled && ctrl.inc_check) {
        vjc = mem.get(ctx, npair, &rc);
        for (int s = 0; s < nspin; ++s) vkc[s] = mem.get(ctx, npair, &rc);
    }
    // REST's prev_hamiltonian (src/scf_io/mod.rs:2851): the last two Fock matrices DIIS handed
    // to the eigensolver, per spin, for the energy-rise mixing
    double* hprev[2][2] = {{nullptr, nullptr}, {nullptr, nullptr}};
    for (int k = 0; k < 2; ++k)
        for (int s = 0; s < nspin; ++s) hprev[k][s] = mem.get(ctx, n2, &rc);
    if (rc) return rc;
    if (ctrl.switch_mode != RSCF_SWITCH_RESTART && ctrl.switch_mode != RSCF_SWITCH_SHIFT)
        return rjk::fail(ctx, RJK_ERR_ARG, std::string(who) + ": switch_mode must be RSCF_SWITCH_RESTART or _SHIFT");
    SCF_HIP(hipMemcpyAsync(S, S_h, n2 * sizeof(double), hipMemcpyHostToDevice, st));
    SCF_HIP(hipMemcpyAsync(H, H_h, n2 * sizeof(double), hipMemcpyHostToDevice, st));
    for (int s = 0; s < nspin; ++s) {
        std::vector<double> occ(n, 0.0);
        for (size_t i = 0; i < nocc[s]; ++i) occ[i] = occ_val;
        SCF_HIP(hipMemcpyAsync(occd[s], occ.data(), n * sizeof(double), hipMemcpyHostToDevice, st));
        if (inc_enabled && nocc[s] > 0) {
            std::vector<double> w2(2 * nocc[s]);
            for (size_t i = 0; i < nocc[s]; ++i) {
                w2[i] = 0.5 * occ_val;
                w2[nocc[s] + i] = -0.5 * occ_val;
            }
            SCF_HIP(hipMemcpyAsync(occ2[s], w2.data(), w2.size() * sizeof(double), hipMemcpyHostToDevice, st));
        }
        SCF_HIP(hipStreamSynchronize(st));   // occ and w2 are host temporaries
    }
    const dim3 g2((unsigned)((n2 + 255) / 256)), b256(256);
    const double one = 1.0, zero = 0.0;
    std::vector<double> part(kRedBlocks);
    double t_jk = 0, t_diag = 0, t_diis = 0;

    auto reduce_sum = [&](double* out) -> int {
        SCF_HIP(hipMemcpyAsync(part.data(), partial, kRedBlocks * sizeof(double), hipMemcpyDeviceToHost, st));
        SCF_HIP(hipStreamSynchronize(st));
        double v = 0.0;
        for (double x : part) v += x;
        *out = v;
        return RJK_OK;
    };
    auto reduce_max = [&](double* out) -> int {
        SCF_HIP(hipMemcpyAsync(part.data(), partial, kRedBlocks * sizeof(double), hipMemcpyDeviceToHost, st));
        SCF_HIP(hipStreamSynchronize(st));
        double v = 0.0;
        for (double x : part) v = std::max(v, x);
        *out = v;
        return RJK_OK;
    };
    // diagonalise Fm for spin s: A[s] <- C, W[s] <- ε (host copy in eps[s])
    std::vector<double> eps[2], eps_prev[2];
    for (int s = 0; s < nspin; ++s) {
        eps[s].assign(n, 0.0);
        eps_prev[s].assign(n, 0.0);
    }
    // Orthogonaliser X = S^{-1/2} = U s^{-1/2} Uᵀ, once: every F C = S C ε then becomes
    // F' = X F X (two GEMMs), F' C' = C' ε (rocSOLVER dsyevd), C = X C' (one GEMM).
    // Same eigenpairs as REST's generalised solver (dspgvx, src/scf_io/mod.rs:1851),
    // without re-factorising S in every iteration.
    // X cached on the context for the same S (the guess step and the SCF of one molecule share it)
    double* X = nullptr;
    if (ctx->scf_x && ctx->scf_s.size() == n2 && std::memcmp(ctx->scf_s.data(), S_h, n2 * sizeof(double)) == 0) {
        X = ctx->scf_x;
    } else {
        if (ctx->scf_x) (void)hipFree(ctx->scf_x);
        ctx->scf_x = nullptr;
        ctx->scf_s.clear();
        SCF_HIP(hipMalloc(&X, n2 * sizeof(double)));
        ctx->scf_x = X;
    }
    if (ctx->scf_s.empty()) {
        SCF_HIP(hipMemcpyAsync(Bw, S, n2 * sizeof(double), hipMemcpyDeviceToDevice, st));
        SCF_BLAS(rocsolver_dsyevd(h, rocblas_evect_original, rocblas_fill_upper, (rocblas_int)n, Bw, (rocblas_int)n,
                                  W[0], Ework, (rocblas_int*)info_d));
        int info = 0;
        std::vector<double> sv(n);
        SCF_HIP(hipMemcpyAsync(&info, info_d, sizeof(int), hipMemcpyDeviceToHost, st));
        SCF_HIP(hipMemcpyAsync(sv.data(), W[0], n * sizeof(double), hipMemcpyDeviceToHost, st));
        SCF_HIP(hipStreamSynchronize(st));
        if (info != 0) return rjk::fail(ctx, RJK_ERR_NOSOLVER, "rocsolver_dsyevd(S) info = " + std::to_string(info));
        if (!(sv[0] > 0.0))
            return rjk::fail(ctx, RJK_ERR_ARG, std::string(who) + ": the overlap matrix is not positive definite");
        for (double& v : sv) v = 1.0 / std::sqrt(v);
        SCF_HIP(hipMemcpyAsync(W[0], sv.data(), n * sizeof(double), hipMemcpyHostToDevice, st));
        hipLaunchKernelGGL(k_scale_columns, g2, b256, 0, st, Bw, W[0], n, T1);   // T1 = U s^{-1/2}
        SCF_HIP(hipGetLastError());
        SCF_BLAS(rocblas_dgemm(h, rocblas_operation_none, rocblas_operation_transpose, (rocblas_int)n, (rocblas_int)n,
                               (rocblas_int)n, &one, T1, (rocblas_int)n, Bw, (rocblas_int)n, &zero, X,
                               (rocblas_int)n));
        SCF_HIP(hipStreamSynchronize(st));   // sv is a host temporary
        ctx->scf_s.assign(S_h, S_h + n2);
    }
    // F C = S C ε for spin s (FP64 throughout: eigenvectors from an FP32 solver at nbf ~1500 are
    // orthonormal only to ~1e-5, which holds the DIIS error above the precision switch).
    auto diagonalise = [&](const double* Fm, int s) -> int {
        const double t0 = now();
        SCF_BLAS(rocblas_dgemm(h, rocblas_operation_none, rocblas_operation_none, (rocblas_int)n, (rocblas_int)n,
                               (rocblas_int)n, &one, Fm, (rocblas_int)n, X, (rocblas_int)n, &zero, T1,
                               (rocblas_int)n));
        SCF_BLAS(rocblas_dgemm(h, rocblas_operation_none, rocblas_operation_none, (rocblas_int)n, (rocblas_int)n,
                               (rocblas_int)n, &one, X, (rocblas_int)n, T1, (rocblas_int)n, &zero, Bw,
                               (rocblas_int)n));
        SCF_BLAS(rocsolver_dsyevd(h, rocblas_evect_original, rocblas_fill_upper, (rocblas_int)n, Bw, (rocblas_int)n,
                                  W[s], Ework, (rocblas_int*)info_d));
        SCF_BLAS(rocblas_dgemm(h, rocblas_operation_none, rocblas_operation_none, (rocblas_int)n, (rocblas_int)n,
                               (rocblas_int)n, &one, X, (rocblas_int)n, Bw, (rocblas_int)n, &zero, A[s],
                               (rocblas_int)n));
        int info = 0;
        SCF_HIP(hipMemcpyAsync(&info, info_d, sizeof(int), hipMemcpyDeviceToHost, st));
        SCF_HIP(hipMemcpyAsync(eps[s].data(), W[s], n * sizeof(double), hipMemcpyDeviceToHost, st));
        SCF_HIP(hipStreamSynchronize(st));
        if (info != 0) return rjk::fail(ctx, RJK_ERR_NOSOLVER, "rocsolver_dsyevd info = " + std::to_string(info));
        t_diag += now() - t0;
        return RJK_OK;
    };
    // Pseudo-diagonalisation of Fm for spin s (Stewart, Császár & Pulay, J. Comput. Chem. 3, 227
    // (1982)): the current orbitals A[s] turn F into Fmo = Aᵀ F A; the occupied orbitals take the
    // first-order rotation U = I + X toward the virtual space; A U is made S-orthonormal again by
    // the Cholesky factor of (A U)ᵀ S (A U) = Uᵀ U; ε ≈ diag(Fmo).  Four GEMMs, one Cholesky and one
    // triangular solve instead of a full eigensolver; used in low-precision iterations below
    // pseudo_diag_err and in every iteration from an incremental reference build on (other
    // FP64 iterations diagonalise exactly).
    auto pseudo_diagonalise = [&](const double* Fm, int s) -> int {
        const double t0 = now();
        SCF_BLAS(rocblas_dgemm(h, rocblas_operation_none, rocblas_operation_none, (rocblas_int)n, (rocblas_int)n,
                               (rocblas_int)n, &one, Fm, (rocblas_int)n, A[s], (rocblas_int)n, &zero, T1,
                               (rocblas_int)n));
        SCF_BLAS(rocblas_dgemm(h, rocblas_operation_transpose, rocblas_operation_none, (rocblas_int)n, (rocblas_int)n,
                               (rocblas_int)n, &one, A[s], (rocblas_int)n, T1, (rocblas_int)n, &zero, T2,
                               (rocblas_int)n));
        hipLaunchKernelGGL(k_diag_copy, dim3((unsigned)((n + 255) / 256)), b256, 0, st, T2, n, W[s]);
        SCF_HIP(hipGetLastError());
        hipLaunchKernelGGL(k_pseudo_rot, dim3(kRedBlocks), b256, 0, st, T2, n, nocc[s], kGapMin, T1);
        SCF_HIP(hipGetLastError());
        SCF_BLAS(rocblas_dgemm(h, rocblas_operation_none, rocblas_operation_none, (rocblas_int)n, (rocblas_int)n,
                               (rocblas_int)n, &one, A[s], (rocblas_int)n, T1, (rocblas_int)n, &zero, Bw,
                               (rocblas_int)n));
        SCF_BLAS(rocblas_dgemm(h, rocblas_operation_transpose, rocblas_operation_none, (rocblas_int)n, (rocblas_int)n,
                               (rocblas_int)n, &one, T1, (rocblas_int)n, T1, (rocblas_int)n, &zero, T2,
                               (rocblas_int)n));
        SCF_BLAS(rocsolver_dpotrf(h, rocblas_fill_upper, (rocblas_int)n, T2, (rocblas_int)n, (rocblas_int*)info_d));
        SCF_BLAS(rocblas_dtrsm(h, rocblas_side_right, rocblas_fill_upper, rocblas_operation_none,
                               rocblas_diagonal_non_unit, (rocblas_int)n, (rocblas_int)n, &one, T2, (rocblas_int)n, Bw,
                               (rocblas_int)n));
        SCF_HIP(hipMemcpyAsync(A[s], Bw, n2 * sizeof(double), hipMemcpyDeviceToDevice, st));
        int info = 0;
        SCF_HIP(hipMemcpyAsync(&info, info_d, sizeof(int), hipMemcpyDeviceToHost, st));
        SCF_HIP(hipMemcpyAsync(eps[s].data(), W[s], n * sizeof(double), hipMemcpyDeviceToHost, st));
        SCF_HIP(hipStreamSynchronize(st));
        if (info != 0) return rjk::fail(ctx, RJK_ERR_NOSOLVER, "rocsolver_dpotrf (pseudo-diagonalisation) info = " +
                                                                   std::to_string(info));
        t_diag += now() - t0;
        return RJK_OK;
    };
    // D_s = occ · C_occ C_occᵀ from A[s] (generate_density_matrix, REST :441-478)
    auto density = [&](int s, double* Dout) -> int {
        if (nocc[s] == 0) {
            hipLaunchKernelGGL(k_fill, g2, b256, 0, st, Dout, n2, 0.0);
            SCF_HIP(hipGetLastError());
            return RJK_OK;
        }
        SCF_BLAS(rocblas_dgemm(h, rocblas_operation_none, rocblas_operation_transpose, (rocblas_int)n, (rocblas_int)n,
                               (rocblas_int)nocc[s], &occ_val, A[s], (rocblas_int)n, A[s], (rocblas_int)n, &zero,
                               Dout, (rocblas_int)n));
        return RJK_OK;
    };
    // F_s = H + J[Σ_s D_s] + sK·K[D_s]; K by the occupied route when the densities
    // are those of the current orbitals A, else by the density-only route.
    // E = E_nuc + ½ Σ_s Σ D_s (H + F_s)   (REST :1715-1782, :1673-1713)
    // Incremental form (inc_armed: after the switch to FP64 with jk_prec_increment set): the
    // first FP64 build by the occupied route becomes the reference (D_ref, C_ref, J_ref, K_ref);
    // later FP64 builds are J_ref + J(D − D_ref) and K_ref + K(ΔD) with the increments at
    // jk_prec_increment, K(ΔD) through the [u | v] columns of k_inc_columns, so their rounding
    // error scales with |D − D_ref| (Almlöf, Faegri & Korsell, J. Comput. Chem. 3, 385 (1982)).
    // A build is an increment only while max|D − D_ref| <= inc_refresh·max|D_ref|; beyond that
    // (or after an exact diagonalisation dropped the reference) it is a full FP64 build and
    // becomes the new reference.
    bool inc_armed = false, have_ref = false;
    // a factored guess density D0_s = Σ_k w_k c_k c_kᵀ: its first K by the occupied route over
    // those columns (A[s], gw[s], gcols[s]) instead of the density-only route
    bool guess_cols = false;
    size_t gcols[2] = {0, 0};
    int n_inc = 0, n_refbuild = 0;
    double dref_max = 0.0, last_dref = 0.0, last_inc_err = 0.0;
    auto fock = [&](double* const* Dsrc, bool occ_route, int prec, double* E) -> int {
        const double t0 = now();
        const double* Djk = Dsrc[0];
        if (nspin == 2) {
            hipLaunchKernelGGL(k_add, g2, b256, 0, st, Dsrc[0], Dsrc[1], n2, Dt);   // D_α + D_β
            SCF_HIP(hipGetLastError());
            Djk = Dt;
        }
        const bool inc_route = inc_armed && prec == RJK_PREC_FP64 && occ_route;
        bool use_inc = inc_route && have_ref;
        last_dref = 0.0;
        last_inc_err = 0.0;
        if (use_inc) {
            hipLaunchKernelGGL(k_sub, dim3(kRedBlocks), b256, 0, st, Djk, Dref, n2, Dinc);
            SCF_HIP(hipGetLastError());
            hipLaunchKernelGGL(k_maxabs, dim3(kRedBlocks), b256, 0, st, Dinc, n2, partial);
            SCF_HIP(hipGetLastError());
            double dd = 0.0;
            SCF_RC(reduce_max(&dd));
            last_dref = dref_max > 0.0 ? dd / dref_max : HUGE_VAL;
            if (!(last_dref <= ctrl.inc_refresh)) use_inc = false;
        }
        if (use_inc) {
            const int pinc = ctrl.jk_prec_increment;
            SCF_RC(rjk_vj_upper_dev(ctx, Dinc, 1, 1.0, pinc, 0, vj));
            hipLaunchKernelGGL(k_axpy_inplace, dim3(kRedBlocks), b256, 0, st, vj, vjref, 1.0, npair);
            SCF_HIP(hipGetLastError());
            for (int s = 0; s < nspin; ++s) {
                if (nocc[s] == 0) {
                    hipLaunchKernelGGL(k_fill, dim3((unsigned)((npair + 255) / 256)), b256, 0, st, vk[s], npair, 0.0);
                    SCF_HIP(hipGetLastError());
                    continue;
                }
                hipLaunchKernelGGL(k_inc_columns, dim3((unsigned)nocc[s]), b256, 0, st, A[s], Cref[s], n, nocc[s],
                                   UV[s]);
                SCF_HIP(hipGetLastError());
                const size_t homo[2] = {2 * nocc[s] - 1, 0};
                SCF_RC(rjk_vk_upper_occ_dev(ctx, UV[s], 2 * nocc[s], homo, occ2[s], 1, sk, pinc, vk[s]));
                hipLaunchKernelGGL(k_axpy_inplace, dim3(kRedBlocks), b256, 0, st, vk[s], vkref[s], 1.0, npair);
                SCF_HIP(hipGetLastError());
            }
            ++n_inc;
        } else {
            SCF_RC(rjk_vj_upper_dev(ctx, Djk, 1, 1.0, prec, 0, vj));
            for (int s = 0; s < nspin; ++s) {
                if (nocc[s] == 0) {
                    hipLaunchKernelGGL(k_fill, dim3((unsigned)((npair + 255) / 256)), b256, 0, st, vk[s], npair, 0.0);
                    SCF_HIP(hipGetLastError());
                } else if (guess_cols) {
                    if (gcols[s] == 0) {
                        hipLaunchKernelGGL(k_fill, dim3((unsigned)((npair + 255) / 256)), b256, 0, st, vk[s], npair,
                                           0.0);
                        SCF_HIP(hipGetLastError());
                    } else {
                        const size_t homo[2] = {gcols[s] - 1, 0};
                        SCF_RC(rjk_vk_upper_occ_dev(ctx, A[s], n, homo, gw[s], 1, sk, prec, vk[s]));
                    }
                } else if (occ_route) {
                    const size_t homo[2] = {nocc[s] - 1, 0};
                    SCF_RC(rjk_vk_upper_occ_dev(ctx, A[s], n, homo, occd[s], 1, sk, prec, vk[s]));
                } else {
                    SCF_RC(rjk_vk_upper_dm_dev(ctx, Dsrc[s], 1, sk, prec, vk[s]));
                }
            }
            if (inc_route) {   // this FP64 build is the reference of the increments that follow
                hipLaunchKernelGGL(k_maxabs, dim3(kRedBlocks), b256, 0, st, Djk, n2, partial);
                SCF_HIP(hipGetLastError());
                SCF_RC(reduce_max(&dref_max));
                ++n_refbuild;
                SCF_HIP(hipMemcpyAsync(Dref, Djk, n2 * sizeof(double), hipMemcpyDeviceToDevice, st));
                SCF_HIP(hipMemcpyAsync(vjref, vj, npair * sizeof(double), hipMemcpyDeviceToDevice, st));
                for (int s = 0; s < nspin; ++s) {
                    if (nocc[s])
                        SCF_HIP(hipMemcpyAsync(Cref[s], A[s], n * nocc[s] * sizeof(double), hipMemcpyDeviceToDevice, st));
                    SCF_HIP(hipMemcpyAsync(vkref[s], vk[s], npair * sizeof(double), hipMemcpyDeviceToDevice, st));
                }
                have_ref = true;
            }
        }
        double e = 0.0;
        for (int s = 0; s < nspin; ++s) {
            hipLaunchKernelGGL(k_fock, g2, b256, 0, st, H, vj, vk[s], n, 1.0, F[s]);
            SCF_HIP(hipGetLastError());
            hipLaunchKernelGGL(k_energy, dim3(kRedBlocks), b256, 0, st, Dsrc[s], H, F[s], n2, partial);
            SCF_HIP(hipGetLastError());
            double v = 0.0;
            SCF_RC(reduce_sum(&v));
            e += v;
        }
        *E = enuc + 0.5 * e;
        if (use_inc && ctrl.inc_check) {   // test mode: the full FP64 build of the same density
            SCF_RC(rjk_vj_upper_dev(ctx, Djk, 1, 1.0, RJK_PREC_FP64, 0, vjc));
            for (int s = 0; s < nspin; ++s) {
                if (nocc[s] == 0) {
                    hipLaunchKernelGGL(k_fill, dim3((unsigned)((npair + 255) / 256)), b256, 0, st, vkc[s], npair, 0.0);
                    SCF_HIP(hipGetLastError());
                } else {
                    const size_t homo[2] = {nocc[s] - 1, 0};
                    SCF_RC(rjk_vk_upper_occ_dev(ctx, A[s], n, homo, occd[s], 1, sk, RJK_PREC_FP64, vkc[s]));
                }
                hipLaunchKernelGGL(k_fock, g2, b256, 0, st, H, vjc, vkc[s], n, 1.0, T1);
                SCF_HIP(hipGetLastError());
                hipLaunchKernelGGL(k_maxabs_diff, dim3(kRedBlocks), b256, 0, st, F[s], T1, n2, partial);
                SCF_HIP(hipGetLastError());
                double d = 0.0, f = 0.0;
                SCF_RC(reduce_max(&d));
                hipLaunchKernelGGL(k_maxabs, dim3(kRedBlocks), b256, 0, st, T1, n2, partial);
                SCF_HIP(hipGetLastError());
                SCF_RC(reduce_max(&f));
                last_inc_err = std::max(last_inc_err, f > 0.0 ? d / f : d);
            }
        }
        t_jk += now() - t0;
        return RJK_OK;
    };

    // ---- guess --------------------------------------------------------------
    bool guess_from_orbitals = false;
    if (d_guess) {
        for (int s = 0; s < nspin; ++s)
            SCF_HIP(hipMemcpyAsync(Dp[s], d_guess + (size_t)s * n2, n2 * sizeof(double), hipMemcpyHostToDevice, st));
        if (c0_guess) {
            for (int s = 0; s < nspin; ++s) {
                gcols[s] = ncol0[s];
                SCF_HIP(hipMemcpyAsync(A[s], c0_guess + (size_t)s * n2, n2 * sizeof(double), hipMemcpyHostToDevice,
                                       st));
                SCF_HIP(hipMemcpyAsync(gw[s], w0_guess + (size_t)s * n, n * sizeof(double), hipMemcpyHostToDevice,
                                       st));
            }
            guess_cols = true;
        }
    } else if (c_guess) {   // guess orbitals: their density, first Fock by the occupied route
        for (int s = 0; s < nspin; ++s) {
            SCF_HIP(hipMemcpyAsync(A[s], c_guess + (size_t)s * n2, n2 * sizeof(double), hipMemcpyHostToDevice, st));
            SCF_RC(density(s, Dp[s]));
        }
        guess_from_orbitals = true;
    } else {   // core-Hamiltonian guess
        for (int s = 0; s < nspin; ++s) {
            SCF_RC(diagonalise(H, s));
            SCF_RC(density(s, Dp[s]));
        }
        guess_from_orbitals = true;
    }
    int prec = ctrl.jk_prec_start >= 0 ? ctrl.jk_prec_start : ctrl.jk_prec;
    int prec_low = prec;   // the precision the FP64 switch leaves
    double E = 0.0;
    SCF_RC(fock(Dp, guess_from_orbitals, prec, &E));
    guess_cols = false;
    const double E_guess = E;
    // the guess's own self-consistency error, max|F D S − S D F| over the spins
    double guess_err = 0.0;
    if (!energy_only && !guess_only) {
        for (int s = 0; s < nspin; ++s) {
            SCF_BLAS(rocblas_dgemm(h, rocblas_operation_none, rocblas_operation_none, (rocblas_int)n, (rocblas_int)n,
                                   (rocblas_int)n, &one, F[s], (rocblas_int)n, Dp[s], (rocblas_int)n, &zero, T1,
                                   (rocblas_int)n));
            SCF_BLAS(rocblas_dgemm(h, rocblas_operation_none, rocblas_operation_none, (rocblas_int)n, (rocblas_int)n,
                                   (rocblas_int)n, &one, T1, (rocblas_int)n, S, (rocblas_int)n, &zero, T2,
                                   (rocblas_int)n));
            hipLaunchKernelGGL(k_commutator, dim3(kRedBlocks), b256, 0, st, T2, n, Bw, partial);
            SCF_HIP(hipGetLastError());
            double m = 0.0;
            SCF_RC(reduce_max(&m));
            guess_err = std::max(guess_err, m);
        }
    }
    if (!energy_only && !guess_only && (E >= ctrl.guess_energy_max || guess_err >= ctrl.guess_err_max)) {
        // guess not good enough: stop here
        SCF_HIP(hipStreamSynchronize(st));
        if (result) {
            *result = rscf_result{};
            result->energy = E;
            result->guess_energy = E;
            result->guess_error = guess_err;
            result->guess_rejected = 1;
            result->wall_seconds = now() - t_start;
            result->jk_seconds = t_jk;
        }
        return RJK_OK;
    }
    if (energy_only) {
        SCF_HIP(hipStreamSynchronize(st));
        if (result) {
            *result = rscf_result{};
            result->energy = E;
            result->wall_seconds = now() - t_start;
            result->jk_seconds = t_jk;
        }
        return RJK_OK;
    }
    int prec_E = prec;
    for (int s = 0; s < nspin; ++s) {
        SCF_RC(diagonalise(F[s], s));
        SCF_RC(density(s, D[s]));
    }
    if (guess_only) {
        double e1 = 0.0;
        if (energy_d1) SCF_RC(fock(D, true, prec, &e1));   // energy of the new orbitals' density
        SCF_HIP(hipStreamSynchronize(st));
        if (energy_d1) *energy_d1 = e1;
        if (result) {
            *result = rscf_result{};
            result->energy = E;
            result->wall_seconds = now() - t_start;
            result->jk_seconds = t_jk;
            result->diag_seconds = t_diag;
        }
        for (int s = 0; s < nspin; ++s) {
            if (dm_out) SCF_HIP(hipMemcpy(dm_out + (size_t)s * n2, D[s], n2 * sizeof(double), hipMemcpyDeviceToHost));
            if (mo_coeff) SCF_HIP(hipMemcpy(mo_coeff + (size_t)s * n2, A[s], n2 * sizeof(double), hipMemcpyDeviceToHost));
            if (mo_energy) std::memcpy(mo_energy + (size_t)s * n, eps[s].data(), n * sizeof(double));
        }
        return RJK_OK;
    }
    double E_prev = E;
    int prec_prev = prec_E;
    for (int s = 0; s < nspin; ++s) eps_prev[s] = eps[s];
    int num_iter = 1, n_fp64 = 0, start_diis = ctrl.start_diis_cycle;
    bool last_pseudo = false;
    int nvec = 0, head = 0;   // DIIS ring: vectors [head, head+nvec) mod nd
    bool shift_pending = false;   // first FP64 build after the switch: shift the stored vectors
    std::vector<double> Bmat((size_t)nd * nd, 0.0);
    bool converged = false;
    // REST's energy_records / prev_hamiltonian (src/scf_io/mod.rs:2850-2851, :3002-3057):
    // energies recorded by the DIIS steps, and a list of at most two slots of Fock matrices
    // that starts as one empty entry (-1 here)
    int n_erec = 0, prec_erec = -1;
    double E_erec = 0.0;
    int hl[2] = {-1, -1}, hl_len = 1;
    if (ctrl.print_level > 0) std::printf("rscf: guess energy %.10f\n", E);

    while (true) {
        const double t_it = now();
        double diis_max = 0.0;
        bool osc = false;
        const double jk0 = t_jk, dg0 = t_diag;
        const int prec_used = prec;
        auto linear_mix = [&]() -> int {   // n1[in] = n0[in] + α (n0[out] − n0[in])
            for (int s = 0; s < nspin; ++s) {
                hipLaunchKernelGGL(k_mix, g2, b256, 0, st, Dp[s], D[s], ctrl.mix_param, n2, Dm[s]);
                SCF_HIP(hipGetLastError());
            }
            SCF_RC(fock(Dm, false, prec, &E));
            for (int s = 0; s < nspin; ++s)
                SCF_HIP(hipMemcpyAsync(Fx[s], F[s], n2 * sizeof(double), hipMemcpyDeviceToDevice, st));
            return RJK_OK;
        };
        if (num_iter < start_diis) {
            shift_pending = false;
            SCF_RC(linear_mix());
        } else {
            if (shift_pending) {   // Flow = F at the low precision, same density
                double e_low = 0.0;
                SCF_RC(fock(D, true, prec_low, &e_low));
                for (int s = 0; s < nspin; ++s)
                    SCF_HIP(hipMemcpyAsync(Fl[s], F[s], n2 * sizeof(double), hipMemcpyDeviceToDevice, st));
            }
            SCF_RC(fock(D, true, prec, &E));
            // energy record and oscillation flag (REST :3002-3010); energies of different J/K
            // precisions are not compared
            ++n_erec;
            osc = n_erec >= 2 && n_erec > ctrl.start_check_oscillation && prec_used == prec_erec && E - E_erec > 0.0;
            E_erec = E;
            prec_erec = prec_used;
            const double t0 = now();
            if (shift_pending && nvec > 0) {
                // δ_s = F_s(FP64) − F_s(low) at the current density.  The low-precision
                // error of F is a smooth function of D, so every stored F_k(low) + δ is
                // F_k(FP64) up to (low-precision error) × |D_k − D|, and its error vector
                // becomes e_k + δ D_k S − S D_k δ.
                for (int s = 0; s < nspin; ++s) {
                    hipLaunchKernelGGL(k_delta_inplace, dim3(kRedBlocks), b256, 0, st, F[s], Fl[s], n2);
                    SCF_HIP(hipGetLastError());
                }
                for (int k = 0; k < nvec; ++k) {
                    const size_t slot_k = (size_t)((head + k) % nd) * ev_len;
                    for (int s = 0; s < nspin; ++s) {
                        hipLaunchKernelGGL(k_axpy_inplace, dim3(kRedBlocks), b256, 0, st, fring + slot_k + s * n2,
                                           Fl[s], 1.0, n2);
                        SCF_HIP(hipGetLastError());
                        SCF_BLAS(rocblas_dgemm(h, rocblas_operation_none, rocblas_operation_none, (rocblas_int)n,
                                               (rocblas_int)n, (rocblas_int)n, &one, Fl[s], (rocblas_int)n,
                                               dring + slot_k + s * n2, (rocblas_int)n, &zero, T1, (rocblas_int)n));
                        SCF_BLAS(rocblas_dgemm(h, rocblas_operation_none, rocblas_operation_none, (rocblas_int)n,
                                               (rocblas_int)n, (rocblas_int)n, &one, T1, (rocblas_int)n, S,
                                               (rocblas_int)n, &zero, T2, (rocblas_int)n));
                        hipLaunchKernelGGL(k_commutator_acc, dim3(kRedBlocks), b256, 0, st, T2, n,
                                           ering + slot_k + s * n2);
                        SCF_HIP(hipGetLastError());
                    }
                }
                for (int i = 0; i < nvec; ++i)
                    for (int j = 0; j <= i; ++j) {
                        double dot = 0.0;
                        SCF_BLAS(rocblas_ddot(h, (rocblas_int)ev_len, ering + (size_t)((head + i) % nd) * ev_len, 1,
                                              ering + (size_t)((head + j) % nd) * ev_len, 1, &dot));
                        Bmat[(size_t)i * nd + j] = Bmat[(size_t)j * nd + i] = dot;
                    }
            }
            shift_pending = false;
            // error vector e_s = F_s D_s S − S D_s F_s, spins back to back, into the ring slot
            if (nvec == nd) {   // drop the oldest
                head = (head + 1) % nd;
                --nvec;
                for (int i = 0; i < nd - 1; ++i)
                    for (int j = 0; j < nd - 1; ++j) Bmat[(size_t)i * nd + j] = Bmat[(size_t)(i + 1) * nd + (j + 1)];
            }
            const int slot = (head + nvec) % nd;
            double* e_new = ering + (size_t)slot * ev_len;
            for (int s = 0; s < nspin; ++s) {
                SCF_BLAS(rocblas_dgemm(h, rocblas_operation_none, rocblas_operation_none, (rocblas_int)n,
                                       (rocblas_int)n, (rocblas_int)n, &one, F[s], (rocblas_int)n, D[s],
                                       (rocblas_int)n, &zero, T1, (rocblas_int)n));
                SCF_BLAS(rocblas_dgemm(h, rocblas_operation_none, rocblas_operation_none, (rocblas_int)n,
                                       (rocblas_int)n, (rocblas_int)n, &one, T1, (rocblas_int)n, S, (rocblas_int)n,
                                       &zero, T2, (rocblas_int)n));
                hipLaunchKernelGGL(k_commutator, dim3(kRedBlocks), b256, 0, st, T2, n, e_new + (size_t)s * n2, partial);
                SCF_HIP(hipGetLastError());
                double m = 0.0;
                SCF_RC(reduce_max(&m));
                diis_max = std::max(diis_max, m);
                SCF_HIP(hipMemcpyAsync(fring + (size_t)slot * ev_len + (size_t)s * n2, F[s], n2 * sizeof(double),
                                       hipMemcpyDeviceToDevice, st));
                SCF_HIP(hipMemcpyAsync(dring + (size_t)slot * ev_len + (size_t)s * n2, D[s], n2 * sizeof(double),
                                       hipMemcpyDeviceToDevice, st));
            }
            ++nvec;
            for (int k = 0; k < nvec; ++k) {
                double dot = 0.0;
                SCF_BLAS(rocblas_ddot(h, (rocblas_int)ev_len, e_new, 1, ering + (size_t)((head + k) % nd) * ev_len, 1,
                                      &dot));
                Bmat[(size_t)(nvec - 1) * nd + k] = Bmat[(size_t)k * nd + (nvec - 1)] = dot;
            }
            std::vector<double> bm((size_t)nvec * nvec), inv;
            for (int i = 0; i < nvec; ++i)
                for (int j = 0; j < nvec; ++j) bm[(size_t)i * nvec + j] = Bmat[(size_t)i * nd + j];
            if (small_inverse(bm, nvec, &inv)) {
                double tot = 0.0;
                for (double v : inv) tot += v;
                for (int s = 0; s < nspin; ++s) {
                    Coefs cf;
                    for (int k = 0; k < nvec; ++k) {
                        double r = 0.0;
                        for (int j = 0; j < nvec; ++j) r += inv[(size_t)k * nvec + j];
                        cf.c[k] = r / tot;
                        cf.f[k] = fring + (size_t)((head + k) % nd) * ev_len + (size_t)s * n2;
                    }
                    hipLaunchKernelGGL(k_lincomb, dim3(kRedBlocks), b256, 0, st, cf, nvec, n2, Fx[s]);
                    SCF_HIP(hipGetLastError());
                }
                // energy rise: (1 − α)·H[DIIS, i−1] + α·H[DIIS, i+1], H[DIIS, i−1] being the front of
                // prev_hamiltonian; its empty first entry leaves the un-extrapolated F (REST :3039-3048)
                if (osc) {
                    for (int s = 0; s < nspin; ++s) {
                        if (hl[0] >= 0)
                            hipLaunchKernelGGL(k_blend, g2, b256, 0, st, hprev[hl[0]][s], Fx[s], ctrl.mix_param, n2,
                                               Fx[s]);
                        else
                            SCF_HIP(hipMemcpyAsync(Fx[s], F[s], n2 * sizeof(double), hipMemcpyDeviceToDevice, st));
                        SCF_HIP(hipGetLastError());
                    }
                }
                // prev_hamiltonian keeps the last two (REST :3056-3057)
                if (hl_len == 2) {
                    hl[0] = hl[1];
                    hl_len = 1;
                }
                const int slot_h = hl[0] == 0 ? 1 : 0;
                for (int s = 0; s < nspin; ++s)
                    SCF_HIP(hipMemcpyAsync(hprev[slot_h][s], Fx[s], n2 * sizeof(double), hipMemcpyDeviceToDevice,
                                           st));
                hl[hl_len++] = slot_h;
            } else {   // DIIS failed: linear mixing now, DIIS again 8 cycles later (REST :3061-3078)
                SCF_RC(linear_mix());
                start_diis = num_iter + 8;
                nvec = 0;
                head = 0;
            }
            t_diis += now() - t0;
            // precision switch: FP64 J/K from the next build on.  The stored Fock
            // matrices and error vectors carry the low-precision J/K error (about the
            // size of the switch threshold): either DIIS restarts with FP64 vectors
            // only, or the next iteration shifts them by F(FP64) − F(low) measured at
            // its density (one extra low-precision build).
            // The start precision gives way to jk_prec below start_switch_err without a
            // DIIS restart (its error is far below that threshold).
            const bool to64 = prec != RJK_PREC_FP64 &&
                              (diis_max < ctrl.switch_diis_err ||
                               (ctrl.jk_prec == RJK_PREC_FP64 && diis_max < ctrl.start_switch_err));
            if (to64) {
                prec_low = prec;
                prec = RJK_PREC_FP64;
                inc_armed = inc_enabled;
                if (ctrl.switch_mode == RSCF_SWITCH_SHIFT) {
                    shift_pending = nvec > 0;
                } else {
                    nvec = 0;
                    head = 0;
                }
            } else if (ctrl.jk_prec_start >= 0 && prec == ctrl.jk_prec_start && diis_max < ctrl.start_switch_err) {
                prec = ctrl.jk_prec;   // one step up only: FP64 never gives way again
            }
        }
        prec_E = prec_used;
        if (prec_used == RJK_PREC_FP64) ++n_fp64;
        // diagonalise the extrapolated Fock matrices, new densities; low-precision iterations
        // below pseudo_diag_err rotate the previous orbitals instead
        // (after an incremental reference every iteration does, so that the occupied orbitals
        // stay aligned with C_ref)
        // The first-order rotation needs a HOMO-LUMO gap: below kGapMin (the denominator clamp of
        // k_pseudo_rot) the eigensolver runs, and an incremental reference is dropped, since an
        // exact diagonalisation may rotate the occupied orbitals away from C_ref.
        bool gap_ok = true;
        for (int s = 0; s < nspin; ++s)
            if (nocc[s] > 0 && nocc[s] < n && eps[s][nocc[s]] - eps[s][nocc[s] - 1] < kGapMin) gap_ok = false;
        bool pseudo = have_ref || (prec_used != RJK_PREC_FP64 && num_iter >= start_diis && diis_max > 0.0 &&
                                   diis_max < ctrl.pseudo_diag_err);
        if (pseudo && !gap_ok) {
            pseudo = false;
            have_ref = false;
        }
        last_pseudo = pseudo;
        double dm_err[2] = {0.0, 0.0}, eev = 0.0;
        for (int s = 0; s < nspin; ++s) {
            SCF_HIP(hipMemcpyAsync(Dp[s], D[s], n2 * sizeof(double), hipMemcpyDeviceToDevice, st));
            if (pseudo) SCF_RC(pseudo_diagonalise(Fx[s], s));
            else SCF_RC(diagonalise(Fx[s], s));
            SCF_RC(density(s, D[s]));
            hipLaunchKernelGGL(k_sumsq_diff, dim3(kRedBlocks), b256, 0, st, D[s], Dp[s], n2, partial);
            SCF_HIP(hipGetLastError());
            double ss = 0.0;
            SCF_RC(reduce_sum(&ss));
            dm_err[s] = std::sqrt(ss) / (double)n;
            for (size_t i = 0; i < n; ++i) eev += (eps[s][i] - eps_prev[s][i]) * (eps[s][i] - eps_prev[s][i]);
        }
        eev = std::sqrt(eev);
        const double dm_max = std::max(dm_err[0], dm_err[1]);
        const double dE = E - E_prev;
        const bool both_fp64 = prec_E == RJK_PREC_FP64 && prec_prev == RJK_PREC_FP64;
        // REST checks nothing in the first iteration (num_iter < 2, :1877-1879) and stops at
        // num_iter >= max_scf_cycle (:1947)
        converged = num_iter >= 2 && both_fp64 && std::fabs(dE) <= ctrl.scf_acc_etot && dm_max <= ctrl.scf_acc_rho &&
                    eev <= ctrl.scf_acc_eev;
        const bool at_limit = num_iter >= 2 && num_iter >= ctrl.max_scf_cycle;
        if (iter_log && num_iter - 1 < std::max(ctrl.max_scf_cycle, 2)) {
            double* row = iter_log + (size_t)(num_iter - 1) * RSCF_LOG_FIELDS;
            row[RSCF_LOG_E] = E;
            row[RSCF_LOG_DE] = dE;
            row[RSCF_LOG_DIIS_MAX] = diis_max;
            row[RSCF_LOG_DM] = dm_max;
            row[RSCF_LOG_EEV] = eev;
            row[RSCF_LOG_PREC] = prec_used;
            row[RSCF_LOG_JK_MS] = (t_jk - jk0) * 1e3;
            row[RSCF_LOG_DIAG_MS] = (t_diag - dg0) * 1e3;
            row[RSCF_LOG_ITER_MS] = (now() - t_it) * 1e3;
            row[RSCF_LOG_DREF] = prec_used == RJK_PREC_FP64 ? last_dref : 0.0;
            row[RSCF_LOG_INC_ERR] = prec_used == RJK_PREC_FP64 ? last_inc_err : 0.0;
            row[RSCF_LOG_OSC] = osc ? 1.0 : 0.0;
            row[RSCF_LOG_EXACT] = pseudo ? 0.0 : 1.0;
        }
        if (ctrl.print_level > 0)
            std::printf("rscf: iter %3d  E %.10f  dE %.3e  diis %.3e  dD %.3e  eev %.3e  prec %s\n", num_iter, E, dE,
                        diis_max, dm_max, eev,
                        prec_used == RJK_PREC_FP64 ? "fp64" : (prec_used == RJK_PREC_FP32 ? "fp32" : "bf16"));
        E_prev = E;
        prec_prev = prec_E;
        for (int s = 0; s < nspin; ++s) eps_prev[s] = eps[s];
        ++num_iter;
        if (converged || at_limit) break;
    }
    if (last_pseudo && (mo_coeff || mo_energy || dm_out)) {   // canonical orbitals for the outputs
        for (int s = 0; s < nspin; ++s) {
            SCF_RC(diagonalise(Fx[s], s));
            SCF_RC(density(s, D[s]));
        }
    }
    SCF_HIP(hipStreamSynchronize(st));
    if (result) {
        result->energy = E;
        result->converged = converged ? 1 : 0;
        result->iterations = num_iter - 1;
        result->fp64_iterations = n_fp64;
        result->wall_seconds = now() - t_start;
        result->jk_seconds = t_jk;
        result->diag_seconds = t_diag;
        result->diis_seconds = t_diis;
        result->guess_energy = E_guess;
        result->guess_error = guess_err;
        result->guess_rejected = 0;
        result->incremental_iterations = n_inc;
        result->reference_builds = n_refbuild;
    }
    for (int s = 0; s < nspin; ++s) {
        if (mo_coeff) SCF_HIP(hipMemcpy(mo_coeff + (size_t)s * n2, A[s], n2 * sizeof(double), hipMemcpyDeviceToHost));
        if (mo_energy) std::memcpy(mo_energy + (size_t)s * n, eps[s].data(), n * sizeof(double));
        if (dm_out) SCF_HIP(hipMemcpy(dm_out + (size_t)s * n2, D[s], n2 * sizeof(double), hipMemcpyDeviceToHost));
    }
    return RJK_OK;
}

}  // namespace

extern "C" int rscf_rhf(rjk_ctx* ctx, const double* S, const double* H, double enuc, size_t nelec,
                        const double* d_guess, const rscf_ctrl* ctrl, rscf_result* result, double* mo_coeff,
                        double* mo_energy, double* dm, double* iter_log) {
    if (!ctx) return RJK_ERR_ARG;
    const size_t nel[2] = {nelec, 0};
    return scf_core(ctx, "rscf_rhf", S, H, enuc, 1, nel, d_guess, ctrl, false, result, mo_coeff, mo_energy, dm,
                    iter_log);
}

extern "C" int rscf_uhf(rjk_ctx* ctx, const double* S, const double* H, double enuc, size_t nalpha, size_t nbeta,
                        const double* d_guess, const rscf_ctrl* ctrl, rscf_result* result, double* mo_coeff,
                        double* mo_energy, double* dm, double* iter_log) {
    if (!ctx) return RJK_ERR_ARG;
    const size_t nel[2] = {nalpha, nbeta};
    return scf_core(ctx, "rscf_uhf", S, H, enuc, 2, nel, d_guess, ctrl, false, result, mo_coeff, mo_energy, dm,
                    iter_log);
}

extern "C" int rscf_rhf_orbitals(rjk_ctx* ctx, const double* S, const double* H, double enuc, size_t nelec,
                                 const double* c_guess, const rscf_ctrl* ctrl, rscf_result* result, double* mo_coeff,
                                 double* mo_energy, double* dm, double* iter_log) {
    if (!ctx) return RJK_ERR_ARG;
    if (!c_guess) return rjk::fail(ctx, RJK_ERR_ARG, "rscf_rhf_orbitals: guess orbitals are required");
    const size_t nel[2] = {nelec, 0};
    return scf_core(ctx, "rscf_rhf_orbitals", S, H, enuc, 1, nel, nullptr, ctrl, false, result, mo_coeff, mo_energy,
                    dm, iter_log, false, c_guess);
}

extern "C" int rscf_guess_step(rjk_ctx* ctx, const double* S, const double* H, double enuc, size_t nspin,
                               const size_t* nelec, const double* d0, int prec, double* energy, double* d1,
                               double* mo_coeff, double* mo_energy, double* energy_d1) {
    if (!ctx) return RJK_ERR_ARG;
    if (!nelec || !d0 || (nspin != 1 && nspin != 2))
        return rjk::fail(ctx, RJK_ERR_ARG, "rscf_guess_step: nspin must be 1 or 2, nelec and d0 non-NULL");
    const size_t nel[2] = {nelec[0], nspin == 2 ? nelec[1] : 0};
    rscf_ctrl c;
    rscf_default_ctrl(&c);
    c.jk_prec = prec;
    rscf_result r{};
    const int rc = scf_core(ctx, "rscf_guess_step", S, H, enuc, (int)nspin, nel, d0, &c, true, &r, mo_coeff,
                            mo_energy, d1, nullptr, false, nullptr, energy_d1);
    if (!rc && energy) *energy = r.energy;
    return rc;
}

extern "C" int rscf_guess_step_factored(rjk_ctx* ctx, const double* S, const double* H, double enuc, size_t nspin,
                                        const size_t* nelec, const double* d0, const double* c0, const double* w0,
                                        const size_t* ncol, int prec, double* energy, double* d1, double* mo_coeff,
                                        double* mo_energy, double* energy_d1) {
    if (!ctx) return RJK_ERR_ARG;
    if (!nelec || !d0 || !c0 || !w0 || !ncol || (nspin != 1 && nspin != 2))
        return rjk::fail(ctx, RJK_ERR_ARG,
                         "rscf_guess_step_factored: nspin must be 1 or 2, nelec, d0, c0, w0 and ncol non-NULL");
    for (size_t s = 0; s < nspin; ++s)
        if (ncol[s] > ctx->nbf) return rjk::fail(ctx, RJK_ERR_ARG, "rscf_guess_step_factored: ncol exceeds nbf");
    const size_t nel[2] = {nelec[0], nspin == 2 ? nelec[1] : 0};
    rscf_ctrl c;
    rscf_default_ctrl(&c);
    c.jk_prec = prec;
    rscf_result r{};
    const int rc = scf_core(ctx, "rscf_guess_step_factored", S, H, enuc, (int)nspin, nel, d0, &c, true, &r, mo_coeff,
                            mo_energy, d1, nullptr, false, nullptr, energy_d1, c0, w0, ncol);
    if (!rc && energy) *energy = r.energy;
    return rc;
}

extern "C" int rscf_energy(rjk_ctx* ctx, const double* S, const double* H, double enuc, size_t nspin,
                           const size_t* nelec, const double* dm, int prec, double* energy) {
    if (!ctx) return RJK_ERR_ARG;
    if (!nelec || !dm || !energy || (nspin != 1 && nspin != 2))
        return rjk::fail(ctx, RJK_ERR_ARG, "rscf_energy: nspin must be 1 or 2, nelec, dm and energy non-NULL");
    const size_t nel[2] = {nelec[0], nspin == 2 ? nelec[1] : 0};
    rscf_ctrl c;
    rscf_default_ctrl(&c);
    c.jk_prec = prec;
    rscf_result r{};
    const int rc = scf_core(ctx, "rscf_energy", S, H, enuc, (int)nspin, nel, dm, &c, true, &r, nullptr, nullptr,
                            nullptr, nullptr, true);
    if (!rc) *energy = r.energy;
    return rc;
}

namespace {
// out = a·X + b·Y
__global__ void k_axpby(const double* __restrict__ X, const double* __restrict__ Y, double a, double b, size_t n2,
                        double* __restrict__ out) {
    size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n2) out[t] = a * X[t] + b * Y[t];
}
// partial[b] = Σ_i M[i][i]
__global__ void k_trace(const double* __restrict__ M, size_t n, double* __restrict__ partial) {
    double acc = 0.0;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        acc += M[i * n + i];
    const double s = block_sum(acc);
    if (threadIdx.x == 0) partial[blockIdx.x] = s;
}
}  // namespace

extern "C" int rscf_mcweeny(rjk_ctx* ctx, const double* S_h, const double* D_h, int iterations, double* D_out,
                            double* electrons) {
    if (!ctx) return RJK_ERR_ARG;
    if (!S_h || !D_h || !D_out || iterations < 0 || iterations > 64)
        return rjk::fail(ctx, RJK_ERR_ARG, "rscf_mcweeny: S, D, D_out required and 0 <= iterations <= 64");
    const size_t n = ctx->nbf, n2 = n * n;
    SCF_HIP(hipSetDevice(ctx->device));
    const hipStream_t st = ctx->stream;
    rocblas_handle h = nullptr;
    SCF_RC(rjk::solver_handle(ctx, &h));
    int rc = RJK_OK;
    DevMem mem(ctx);
    double* S = mem.get(ctx, n2, &rc);
    double* P = mem.get(ctx, n2, &rc);
    double* T1 = mem.get(ctx, n2, &rc);
    double* T2 = mem.get(ctx, n2, &rc);
    double* T3 = mem.get(ctx, n2, &rc);
    double* partial = mem.get(ctx, kRedBlocks, &rc);
    if (rc) return rc;
    SCF_HIP(hipMemcpyAsync(S, S_h, n2 * sizeof(double), hipMemcpyHostToDevice, st));
    SCF_HIP(hipMemcpyAsync(T1, D_h, n2 * sizeof(double), hipMemcpyHostToDevice, st));
    const dim3 g2((unsigned)((n2 + 255) / 256)), b256(256);
    hipLaunchKernelGGL(k_axpby, g2, b256, 0, st, T1, T1, 0.5, 0.0, n2, P);   // P = D / 2
    SCF_HIP(hipGetLastError());
    const double one = 1.0, zero = 0.0;
    const rocblas_int ni = (rocblas_int)n;
    auto gemm = [&](const double* A, const double* B, double* C) -> int {
        SCF_BLAS(rocblas_dgemm(h, rocblas_operation_none, rocblas_operation_none, ni, ni, ni, &one, A, ni, B, ni,
                               &zero, C, ni));
        return RJK_OK;
    };
    for (int it = 0; it < iterations; ++it) {   // P <- 3 PSP − 2 PSPSP
        SCF_RC(gemm(P, S, T1));
        SCF_RC(gemm(T1, P, T2));
        SCF_RC(gemm(T1, T2, T3));
        hipLaunchKernelGGL(k_axpby, g2, b256, 0, st, T2, T3, 3.0, -2.0, n2, P);
        SCF_HIP(hipGetLastError());
    }
    if (electrons) {   // 2 tr(P S)
        SCF_RC(gemm(P, S, T1));
        hipLaunchKernelGGL(k_trace, dim3(kRedBlocks), b256, 0, st, T1, n, partial);
        SCF_HIP(hipGetLastError());
        std::vector<double> part(kRedBlocks);
        SCF_HIP(hipMemcpyAsync(part.data(), partial, kRedBlocks * sizeof(double), hipMemcpyDeviceToHost, st));
        SCF_HIP(hipStreamSynchronize(st));
        double tr = 0.0;
        for (double v : part) tr += v;
        *electrons = 2.0 * tr;
    }
    hipLaunchKernelGGL(k_axpby, g2, b256, 0, st, P, P, 2.0, 0.0, n2, T1);   // D = 2 P
    SCF_HIP(hipGetLastError());
    SCF_HIP(hipMemcpyAsync(D_out, T1, n2 * sizeof(double), hipMemcpyDeviceToHost, st));
    SCF_HIP(hipStreamSynchronize(st));
    return RJK_OK;
}

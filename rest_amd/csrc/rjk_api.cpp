// rjk_api.cpp — the C ABI declared in include/rest_jk.h: contexts, device B
// storage, workspace, the J/K entry points and phase timing.
#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "rjk_internal.hpp"

namespace {
thread_local std::string g_create_error;

constexpr double kEigRelCut = 1e-13;     // density-only route: |λ| <= cut·max|λ| is dropped
constexpr double kAsymTol = 1e-10;       // density-only route: D symmetry tolerance

struct Carve {
    std::vector<size_t> off;
    size_t total = 0;
    size_t add(size_t bytes) {
        off.push_back(total);
        total += rjk::round_up(bytes, 256);
        return off.size() - 1;
    }
};
}  // namespace

namespace rjk {

int fail(rjk_ctx* ctx, int code, const std::string& msg) {
    if (ctx) ctx->err = msg;
    else g_create_error = msg;
    return code;
}

int buf_get(rjk_ctx* ctx, void** buf, size_t* size, size_t bytes, void** out) {
    if (bytes > *size) {
        RJK_HIP(ctx, hipStreamSynchronize(ctx->stream));
        if (*buf) (void)hipFree(*buf);
        *buf = nullptr;
        *size = 0;
        size_t sz = round_up(bytes, size_t(1) << 20);
        if (hipMalloc(buf, sz) != hipSuccess) {
            (void)hipGetLastError();
            *buf = nullptr;
            return fail(ctx, RJK_ERR_NOMEM, "device allocation of " + std::to_string(sz) + " bytes failed");
        }
        *size = sz;
    }
    *out = *buf;
    return RJK_OK;
}

int ws_get(rjk_ctx* ctx, size_t bytes, void** out) { return buf_get(ctx, &ctx->ws, &ctx->ws_size, bytes, out); }

PhaseScope::PhaseScope(rjk_ctx* c, int phase, int prec, double work, bool allow) : ctx(c), on(allow && c->prof) {
    if (!on) return;
    if (!ctx->ev_free.empty()) {
        ev = ctx->ev_free.back();
        ctx->ev_free.pop_back();
    } else {
        (void)hipEventCreate(&ev.a);
        (void)hipEventCreate(&ev.b);
    }
    ev.phase = phase;
    ev.prec = prof_row(prec);
    ev.work = work;
    (void)hipEventRecord(ev.a, ctx->stream);
}
PhaseScope::~PhaseScope() {
    if (!on) return;
    (void)hipEventRecord(ev.b, ctx->stream);
    ctx->ev_pending.push_back(ev);
}

int solver_handle(rjk_ctx* ctx, rocblas_handle* h) {
    if (!ctx->solver) {
        rocblas_handle hh;
        if (rocblas_create_handle(&hh) != rocblas_status_success)
            return fail(ctx, RJK_ERR_NOSOLVER, "rocblas_create_handle failed");
        ctx->solver = hh;
    }
    rocblas_set_stream((rocblas_handle)ctx->solver, ctx->stream);
    *h = (rocblas_handle)ctx->solver;
    return RJK_OK;
}

}  // namespace rjk

using namespace rjk;

// ---------------------------------------------------------------------------
// context
// ---------------------------------------------------------------------------
extern "C" int rjk_create(int device, size_t num_basis, size_t num_auxbas, rjk_ctx** out) {
    if (!out) return fail(nullptr, RJK_ERR_ARG, "rjk_create: out is NULL");
    *out = nullptr;
    if (num_basis == 0 || num_auxbas == 0)
        return fail(nullptr, RJK_ERR_ARG, "rjk_create: num_basis and num_auxbas must be positive");
    if (num_basis > 46340)  // nbf^2 must fit the int tile arithmetic of the K kernels
        return fail(nullptr, RJK_ERR_ARG, "rjk_create: num_basis too large");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || device < 0 || device >= ndev)
        return fail(nullptr, RJK_ERR_HIP, "rjk_create: HIP device " + std::to_string(device) + " not available");
    if (hipSetDevice(device) != hipSuccess) return fail(nullptr, RJK_ERR_HIP, "rjk_create: hipSetDevice failed");
    rjk_ctx* c = new rjk_ctx();
    c->device = device;
    c->nbf = num_basis;
    c->naux = num_auxbas;
    c->npair = num_basis * (num_basis + 1) / 2;
    c->ldb = round_up(c->npair, kLdAlign);
    if (hipStreamCreateWithFlags(&c->own_stream, hipStreamNonBlocking) != hipSuccess) {
        delete c;
        return fail(nullptr, RJK_ERR_HIP, "rjk_create: stream creation failed");
    }
    c->stream = c->own_stream;
    // The allocation-free J entry point's scratch: its failure only disables that entry point.
    if (scratch_init(c) != RJK_OK) c->scratch = nullptr;
    c->err.clear();
    *out = c;
    return RJK_OK;
}

extern "C" int rjk_destroy(rjk_ctx* ctx) {
    if (!ctx) return RJK_OK;
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    if (ctx->b64) (void)hipFree(ctx->b64);
    if (ctx->b32) (void)hipFree(ctx->b32);
    if (ctx->bpl) (void)hipFree(ctx->bpl);
    if (ctx->scf_x) (void)hipFree(ctx->scf_x);
    if (ctx->scf_ws) (void)hipFree(ctx->scf_ws);
    if (ctx->ws) (void)hipFree(ctx->ws);
    if (ctx->io) (void)hipFree(ctx->io);
    if (ctx->aux) (void)hipFree(ctx->aux);
    if (ctx->solver) rocblas_destroy_handle((rocblas_handle)ctx->solver);
    for (auto& e : ctx->ev_pending) { (void)hipEventDestroy(e.a); (void)hipEventDestroy(e.b); }
    for (auto& e : ctx->ev_free) { (void)hipEventDestroy(e.a); (void)hipEventDestroy(e.b); }
    (void)hipStreamDestroy(ctx->own_stream);
    delete ctx;
    return RJK_OK;
}

extern "C" const char* rjk_last_error(const rjk_ctx* ctx) {
    return ctx ? ctx->err.c_str() : g_create_error.c_str();
}
extern "C" size_t rjk_num_basis(const rjk_ctx* ctx) { return ctx ? ctx->nbf : 0; }
extern "C" size_t rjk_num_auxbas(const rjk_ctx* ctx) { return ctx ? ctx->naux : 0; }
extern "C" size_t rjk_num_baspar(const rjk_ctx* ctx) { return ctx ? ctx->npair : 0; }
extern "C" size_t rjk_ld_baspar(const rjk_ctx* ctx) { return ctx ? ctx->ldb : 0; }

extern "C" int rjk_set_stream(rjk_ctx* ctx, void* stream) {
    if (!ctx) return RJK_ERR_ARG;
    RJK_HIP(ctx, hipStreamSynchronize(ctx->stream));
    ctx->stream = stream ? (hipStream_t)stream : ctx->own_stream;
    if (ctx->solver) rocblas_set_stream((rocblas_handle)ctx->solver, ctx->stream);
    return RJK_OK;
}

extern "C" int rjk_synchronize(rjk_ctx* ctx) {
    if (!ctx) return RJK_ERR_ARG;
    RJK_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return RJK_OK;
}

extern "C" int rjk_set_workspace_limit(rjk_ctx* ctx, size_t bytes) {
    if (!ctx) return RJK_ERR_ARG;
    ctx->ws_limit = bytes ? bytes : (size_t(4) << 30);
    return RJK_OK;
}

// ---------------------------------------------------------------------------
// B storage
// ---------------------------------------------------------------------------
namespace rjk {

int alloc_b(rjk_ctx* ctx, unsigned mask) {
    if (!(mask & RJK_STORE_FP64))
        return fail(ctx, RJK_ERR_ARG, "store_mask must include RJK_STORE_FP64 (the FP64 copy is the reference)");
    if (mask & ~unsigned(RJK_STORE_FP64 | RJK_STORE_FP32 | RJK_STORE_BF16))
        return fail(ctx, RJK_ERR_ARG, "store_mask holds unknown bits");
    RJK_HIP(ctx, hipSetDevice(ctx->device));
    RJK_HIP(ctx, hipStreamSynchronize(ctx->stream));
    const size_t n = ctx->naux * ctx->ldb;
    if (!ctx->b64) {
        if (hipMalloc(&ctx->b64, n * sizeof(double)) != hipSuccess) {
            (void)hipGetLastError();
            ctx->b64 = nullptr;
            return fail(ctx, RJK_ERR_NOMEM, "B (FP64) allocation of " + std::to_string(n * 8) + " bytes failed");
        }
    }
    if ((mask & RJK_STORE_FP32) && !ctx->b32) {
        if (hipMalloc(&ctx->b32, n * sizeof(float)) != hipSuccess) {
            (void)hipGetLastError();
            ctx->b32 = nullptr;
            return fail(ctx, RJK_ERR_NOMEM, "B (FP32) allocation failed");
        }
    }
    if (!(mask & RJK_STORE_FP32) && ctx->b32) {
        (void)hipFree(ctx->b32);
        ctx->b32 = nullptr;
    }
    if ((mask & RJK_STORE_BF16) && !ctx->bpl) {
        const size_t pad = round_up(ctx->nbf, kTileMN);
        const size_t bytes = (size_t)kBfStoredPlanes * ctx->naux * pad * pad * 2;
        if (hipMalloc(&ctx->bpl, bytes) != hipSuccess) {
            (void)hipGetLastError();
            ctx->bpl = nullptr;
            return fail(ctx, RJK_ERR_NOMEM, "B (BF16 planes) allocation of " + std::to_string(bytes) + " bytes failed");
        }
        ctx->bpl_pad = pad;
    }
    if (!(mask & RJK_STORE_BF16) && ctx->bpl) {
        (void)hipFree(ctx->bpl);
        ctx->bpl = nullptr;
        ctx->bpl_pad = 0;
    }
    return RJK_OK;
}

int finish_b(rjk_ctx* ctx) {
    if (ctx->b32) RJK_HIP(ctx, launch_to_fp32(ctx->stream, ctx->b64, ctx->naux * ctx->ldb, ctx->b32));
    if (ctx->bpl)
        RJK_HIP(ctx, launch_bf_planes_b(ctx->stream, ctx->b64, ctx->ldb, ctx->nbf, ctx->bpl_pad, ctx->naux, ctx->bpl));
    RJK_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return RJK_OK;
}

}  // namespace rjk

extern "C" int rjk_set_rimatr(rjk_ctx* ctx, const double* rimatr, unsigned store_mask) {
    if (!ctx || !rimatr) return ctx ? fail(ctx, RJK_ERR_ARG, "rimatr is NULL") : RJK_ERR_ARG;
    int rc = alloc_b(ctx, store_mask);
    if (rc) return rc;
    RJK_HIP(ctx, hipMemsetAsync(ctx->b64, 0, ctx->naux * ctx->ldb * sizeof(double), ctx->stream));
    RJK_HIP(ctx, hipMemcpy2DAsync(ctx->b64, ctx->ldb * sizeof(double), rimatr, ctx->npair * sizeof(double),
                                  ctx->npair * sizeof(double), ctx->naux, hipMemcpyHostToDevice, ctx->stream));
    return finish_b(ctx);
}

extern "C" int rjk_set_rimatr_dev(rjk_ctx* ctx, const double* rimatr_dev, unsigned store_mask) {
    if (!ctx || !rimatr_dev) return ctx ? fail(ctx, RJK_ERR_ARG, "rimatr_dev is NULL") : RJK_ERR_ARG;
    int rc = alloc_b(ctx, store_mask);
    if (rc) return rc;
    RJK_HIP(ctx, launch_pad_rimatr(ctx->stream, rimatr_dev, ctx->npair, ctx->naux, ctx->ldb, ctx->b64));
    return finish_b(ctx);
}

extern "C" int rjk_set_ri3fn(rjk_ctx* ctx, const double* ri3fn, unsigned store_mask) {
    if (!ctx || !ri3fn) return ctx ? fail(ctx, RJK_ERR_ARG, "ri3fn is NULL") : RJK_ERR_ARG;
    int rc = alloc_b(ctx, store_mask);
    if (rc) return rc;
    const size_t n2 = ctx->nbf * ctx->nbf;
    // stream full columns through a staging slab of at most 256 MiB
    size_t slab = std::max<size_t>(1, (size_t(256) << 20) / (n2 * sizeof(double)));
    slab = std::min(slab, ctx->naux);
    void* ws = nullptr;
    rc = ws_get(ctx, slab * n2 * sizeof(double), &ws);
    if (rc) return rc;
    for (size_t P0 = 0; P0 < ctx->naux; P0 += slab) {
        size_t nP = std::min(slab, ctx->naux - P0);
        RJK_HIP(ctx, hipMemcpyAsync(ws, ri3fn + P0 * n2, nP * n2 * sizeof(double), hipMemcpyHostToDevice,
                                    ctx->stream));
        RJK_HIP(ctx, launch_pack_ri3fn(ctx->stream, (const double*)ws, ctx->nbf, nP, ctx->ldb,
                                       ctx->b64 + P0 * ctx->ldb));
        RJK_HIP(ctx, hipStreamSynchronize(ctx->stream));  // pageable source: keep the slab alive
    }
    return finish_b(ctx);
}

extern "C" int rjk_fill_rimatr_synthetic(rjk_ctx* ctx, uint64_t seed, unsigned store_mask) {
    if (!ctx) return RJK_ERR_ARG;
    int rc = alloc_b(ctx, store_mask);
    if (rc) return rc;
    RJK_HIP(ctx, launch_fill_synth(ctx->stream, seed, ctx->npair, ctx->naux, ctx->ldb, ctx->b64));
    return finish_b(ctx);
}

extern "C" const double* rjk_rimatr_dev(const rjk_ctx* ctx) { return ctx ? ctx->b64 : nullptr; }

extern "C" int rjk_get_rimatr(rjk_ctx* ctx, double* rimatr) {
    if (!ctx) return RJK_ERR_ARG;
    if (!rimatr) return fail(ctx, RJK_ERR_ARG, "rjk_get_rimatr: NULL output");
    if (!ctx->b64) return fail(ctx, RJK_ERR_STATE, "B is not set");
    RJK_HIP(ctx, hipSetDevice(ctx->device));
    RJK_HIP(ctx, hipMemcpy2DAsync(rimatr, ctx->npair * sizeof(double), ctx->b64, ctx->ldb * sizeof(double),
                                  ctx->npair * sizeof(double), ctx->naux, hipMemcpyDeviceToHost, ctx->stream));
    RJK_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return RJK_OK;
}

// ---------------------------------------------------------------------------
// J
// ---------------------------------------------------------------------------
namespace rjk {

// The B copy a call reads.  J: the BF16 precisions run on the FP32 copy when it is
// stored, else on the FP64 copy (*jprec receives the precision J runs at).  K: the
// BF16 precisions need the planes.
int check_common(rjk_ctx* ctx, size_t spin_channel, int prec, bool for_k, const void** bptr, int* jprec = nullptr) {
    if (spin_channel != 1 && spin_channel != 2)
        return fail(ctx, RJK_ERR_ARG, "spin_channel must be 1 or 2, got " + std::to_string(spin_channel));
    if (!ctx->b64) return fail(ctx, RJK_ERR_STATE, "B is not set (rjk_set_rimatr / rjk_set_ri3fn first)");
    if (!is_bf16(prec) && prec != RJK_PREC_FP64 && prec != RJK_PREC_FP32)
        return fail(ctx, RJK_ERR_ARG, "unsupported precision " + std::to_string(prec));
    int eff = prec;
    if (is_bf16(prec)) {
        if (for_k) {
            if (!ctx->bpl) return fail(ctx, RJK_ERR_STATE, "BF16 planes of B not stored (RJK_STORE_BF16)");
            *bptr = ctx->bpl;
            return RJK_OK;
        }
        eff = ctx->b32 ? RJK_PREC_FP32 : RJK_PREC_FP64;
    }
    if (jprec) *jprec = eff;
    if (eff == RJK_PREC_FP64) *bptr = ctx->b64;
    else {
        if (!ctx->b32) return fail(ctx, RJK_ERR_STATE, "FP32 copy of B not stored (RJK_STORE_FP32)");
        *bptr = ctx->b32;
    }
    return RJK_OK;
}

int j_source(rjk_ctx* ctx, size_t spin_channel, int prec, const void** bptr, int* jprec) {
    return check_common(ctx, spin_channel, prec, false, bptr, jprec);
}

namespace {
// [w | γ partials | γ | J partials] of one J call
struct JLayout {
    size_t nchunk = 0, nsplit = 0;
    Carve cv;
    size_t iw = 0, ig = 0, iG = 0, ij = 0;
};
JLayout j_layout(const rjk_ctx* ctx, int ns, int prec) {
    JLayout L;
    j_plan(ctx->ldb, ctx->naux, prec, &L.nchunk, &L.nsplit);
    L.iw = L.cv.add(ns * ctx->ldb * sizeof(double));
    L.ig = L.cv.add(L.nchunk * ns * ctx->naux * sizeof(double));
    L.iG = L.cv.add(ns * ctx->naux * sizeof(double));
    L.ij = L.cv.add(L.nsplit * ns * ctx->ldb * sizeof(double));
    return L;
}
}  // namespace

size_t j_ws_bytes(const rjk_ctx* ctx, int nspin, int prec) { return j_layout(ctx, nspin, prec).cv.total; }

// Every workspace element a J call reads is written earlier in the same call, so the
// workspace needs no clearing (the allocation-free entry point relies on this).
int j_run(rjk_ctx* ctx, const void* b, int prec, const double* dm_dev, int ns, double scale, int full_dm,
          double* vj_dev, void* ws, bool timed) {
    const JLayout L = j_layout(ctx, ns, prec);
    char* base = (char*)ws;
    double* w = (double*)(base + L.cv.off[L.iw]);
    double* gpart = (double*)(base + L.cv.off[L.ig]);
    double* gamma = (double*)(base + L.cv.off[L.iG]);
    double* jpart = (double*)(base + L.cv.off[L.ij]);
    hipStream_t st = ctx->stream;
    {
        PhaseScope ps(ctx, 4, 0, 0.0, timed);
        RJK_HIP(ctx, launch_dm_weights(st, dm_dev, ctx->nbf, ctx->npair, ctx->ldb, ns, full_dm, w));
    }
    const double sb = prec == RJK_PREC_FP64 ? 8.0 : 4.0;
    const double jpass = (double)ctx->npair * ctx->naux * sb + 8.0 * ns * ctx->npair;   // one pass over B
    {
        PhaseScope ps(ctx, 0, prec, jpass, timed);
        RJK_HIP(ctx, launch_jgamma(st, prec, b, ctx->ldb, ctx->naux, ns, w, L.nchunk, gpart));
    }
    {
        PhaseScope ps(ctx, 4, 0, 0.0, timed);
        RJK_HIP(ctx, launch_gamma_reduce(st, gpart, L.nchunk, ns, ctx->naux, gamma));
    }
    {
        PhaseScope ps(ctx, 1, prec, jpass, timed);
        RJK_HIP(ctx, launch_jacc(st, prec, b, ctx->ldb, ctx->naux, ns, gamma, L.nsplit, jpart));
    }
    {
        PhaseScope ps(ctx, 4, 0, 0.0, timed);
        RJK_HIP(ctx, launch_jreduce(st, jpart, L.nsplit, ns, ctx->ldb, ctx->npair, scale, vj_dev));
    }
    return RJK_OK;
}

}  // namespace rjk

extern "C" int rjk_vj_upper_dev(rjk_ctx* ctx, const double* dm_dev, size_t spin_channel, double scaling_factor,
                                int prec, int full_dm, double* vj_dev) {
    if (!ctx) return RJK_ERR_ARG;
    if (!dm_dev || !vj_dev) return fail(ctx, RJK_ERR_ARG, "vj: NULL device pointer");
    const void* b = nullptr;
    int rc = check_common(ctx, spin_channel, prec, false, &b, &prec);
    if (rc) return rc;
    RJK_HIP(ctx, hipSetDevice(ctx->device));
    const int ns = (int)spin_channel;
    void* ws = nullptr;
    rc = ws_get(ctx, j_ws_bytes(ctx, ns, prec), &ws);
    if (rc) return rc;
    return j_run(ctx, b, prec, dm_dev, ns, scaling_factor, full_dm, vj_dev, ws, true);
}

namespace {

// staging of the host-pointer entry points: the context's grow-only io buffer
int io_get(rjk_ctx* ctx, size_t bytes, void** out) { return buf_get(ctx, &ctx->io, &ctx->io_size, bytes, out); }

int check_host_finite(rjk_ctx* ctx, const double* x, size_t n, const char* what) {
    for (size_t i = 0; i < n; ++i)
        if (!std::isfinite(x[i])) return fail(ctx, RJK_ERR_NONFINITE, std::string(what) + " holds NaN/Inf");
    return RJK_OK;
}

int vj_host(rjk_ctx* ctx, const double* dm, size_t spin_channel, double scaling_factor, int prec,
            double* vj_upper, int full_dm) {
    if (!ctx) return RJK_ERR_ARG;
    if (!dm || !vj_upper) return fail(ctx, RJK_ERR_ARG, "vj: NULL host pointer");
    if (spin_channel != 1 && spin_channel != 2)
        return fail(ctx, RJK_ERR_ARG, "spin_channel must be 1 or 2, got " + std::to_string(spin_channel));
    const size_t n2 = ctx->nbf * ctx->nbf;
    int rc = check_host_finite(ctx, dm, spin_channel * n2, "dm");
    if (rc) return rc;
    RJK_HIP(ctx, hipSetDevice(ctx->device));
    void* io = nullptr;
    const size_t out_n = spin_channel * ctx->npair;
    rc = io_get(ctx, (spin_channel * n2 + out_n) * sizeof(double), &io);
    if (rc) return rc;
    double* d_dm = (double*)io;
    double* d_out = d_dm + spin_channel * n2;
    RJK_HIP(ctx, hipMemcpyAsync(d_dm, dm, spin_channel * n2 * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
    rc = rjk_vj_upper_dev(ctx, d_dm, spin_channel, scaling_factor, prec, full_dm, d_out);
    if (rc) return rc;
    RJK_HIP(ctx, hipMemcpyAsync(vj_upper, d_out, out_n * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    RJK_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return RJK_OK;
}

}  // namespace

extern "C" int rjk_vj_upper_with_rimatr_v_sync(rjk_ctx* ctx, const double* dm, size_t spin_channel,
                                               double scaling_factor, int prec, double* vj_upper) {
    return vj_host(ctx, dm, spin_channel, scaling_factor, prec, vj_upper, 0);
}

extern "C" int rjk_vj_upper_with_ri_v_sync(rjk_ctx* ctx, const double* dm, size_t spin_channel,
                                           double scaling_factor, int prec, double* vj_upper) {
    return vj_host(ctx, dm, spin_channel, scaling_factor, prec, vj_upper, 1);
}

// ---------------------------------------------------------------------------
// K
// ---------------------------------------------------------------------------
namespace {

void k_plan_cols(KPlan& kp, size_t nbf, const size_t* ncol, int ns, int prec) {
    const size_t gran = is_bf16(prec) ? kBfColGran : kTileK;
    kp.nbf_pad = round_up(nbf, kTileMN);
    kp.col_cnt[0] = round_up(ncol[0], gran);
    kp.col_off[0] = 0;
    kp.col_cnt[1] = ns == 2 ? round_up(ncol[1], gran) : 0;
    kp.col_off[1] = kp.col_cnt[0];
    kp.nk = round_up(kp.col_cnt[0] + kp.col_cnt[1], kColPad);
}

// chunking and split-K; carves [cocc | sign | kpart | M'] out of the workspace.
// C', the signs and M' use the element type of the K precision; the split-K
// partial tiles are FP64.
int k_plan_ws(rjk_ctx* ctx, KPlan& kp, int ns, int prec, void** cocc, void** sign, double** kpart, void** m) {
    const size_t es = k_elem_size(prec);
    const size_t nt = kp.nbf_pad / kTileMN;
    const size_t ntiles = nt * (nt + 1) / 2;
    const size_t target = 1024;  // SYRK work-groups
    const size_t slot = kp.nbf_pad * kp.nbf_pad;
    const size_t slot_bytes = (size_t)ns * slot * sizeof(double);
    const size_t perP = kp.nk * kp.nbf_pad * es;
    const size_t small = round_up(kp.nbf_pad * kp.nk * es, 256) + round_up(kp.nk * es, 256);
    // split-K slots: a multiple of 8 (one XCD per split), about `target`
    // work-groups, at most a quarter of the workspace
    size_t nsplit = round_up(std::max<size_t>(1, target / (ntiles * ns)), 8);
    const size_t cap = std::max<size_t>(8, ((ctx->ws_limit / 4) / slot_bytes) / 8 * 8);
    nsplit = std::min(nsplit, cap);
    size_t fixed = small + nsplit * slot_bytes;
    size_t avail = ctx->ws_limit > fixed ? ctx->ws_limit - fixed : 0;
    size_t chunk = perP ? std::min(ctx->naux, std::max<size_t>(1, avail / perP)) : ctx->naux;
    if (prec != RJK_PREC_FP64 && kp.nk > 0) {
        // FP32 accumulators: each split sums at most kFp32Chain products (pps auxiliary
        // columns × nk) before its partial joins the FP64 sum, so rounding does not grow
        // with naux.  Up to half the workspace goes to the partials.
        const size_t pps = std::max<size_t>(1, kFp32Chain / kp.nk);
        const size_t smax = std::max<size_t>(8, ((ctx->ws_limit / 2) / slot_bytes) / 8 * 8);
        fixed = small + smax * slot_bytes;
        avail = ctx->ws_limit > fixed ? ctx->ws_limit - fixed : 0;
        chunk = std::min({ctx->naux, std::max<size_t>(1, avail / perP), smax * pps});
        nsplit = std::min(smax, std::max(nsplit, round_up((chunk + pps - 1) / pps, 8)));
    }
    kp.nsplit = nsplit;
    kp.chunk = chunk;
    Carve cv;
    size_t ic = cv.add(kp.nbf_pad * kp.nk * es);
    size_t is = cv.add(kp.nk * es);
    size_t ik = cv.add(kp.nsplit * ns * slot * sizeof(double));
    size_t im = cv.add(chunk * perP);
    void* ws = nullptr;
    int rc = ws_get(ctx, cv.total, &ws);
    if (rc) return rc;
    char* base = (char*)ws;
    *cocc = base + cv.off[ic];
    *sign = base + cv.off[is];
    *kpart = (double*)(base + cv.off[ik]);
    *m = base + cv.off[im];
    return RJK_OK;
}

int k_run(rjk_ctx* ctx, const KPlan& kp, int ns, int prec, const void* b, const void* cocc, const void* sign,
          double* kpart, void* m, double scale, double* vk_dev) {
    hipStream_t st = ctx->stream;
    if (kp.nk == 0) {  // no occupied column (or D = 0): K = 0
        RJK_HIP(ctx, hipMemsetAsync(vk_dev, 0, (size_t)ns * ctx->npair * sizeof(double), st));
        return RJK_OK;
    }
    const bool bf = is_bf16(prec);
    for (size_t P0 = 0; P0 < ctx->naux; P0 += kp.chunk) {
        size_t Pc = std::min(kp.chunk, ctx->naux - P0);
        const double nb = (double)ctx->nbf;
        {
            PhaseScope ps(ctx, 2, prec, 2.0 * nb * nb * (double)kp.nk * (double)Pc);
            if (bf) RJK_HIP(ctx, launch_bf_khalf(st, prec, b, ctx->naux, ctx->nbf, P0, Pc, kp, cocc, m));
            else RJK_HIP(ctx, launch_khalf(st, prec, b, ctx->ldb, ctx->nbf, P0, Pc, kp, cocc, m));
        }
        {
            PhaseScope ps(ctx, 3, prec, nb * (nb + 1.0) * (double)kp.nk * (double)Pc);
            if (bf) RJK_HIP(ctx, launch_bf_ksyrk(st, prec, m, Pc, kp, ns, sign, P0 > 0 ? 1 : 0, kpart));
            else RJK_HIP(ctx, launch_ksyrk(st, prec, m, Pc, kp, ns, sign, P0 > 0 ? 1 : 0, kpart));
        }
    }
    PhaseScope ps(ctx, 4);
    RJK_HIP(ctx, launch_kreduce(st, kpart, kp, ns, ctx->nbf, scale, vk_dev));
    return RJK_OK;
}

}  // namespace

extern "C" int rjk_vk_upper_occ_dev(rjk_ctx* ctx, const double* eigv_dev, size_t num_state, const size_t* homo,
                                    const double* occ_dev, size_t spin_channel, double scaling_factor, int prec,
                                    double* vk_dev) {
    if (!ctx) return RJK_ERR_ARG;
    if (!eigv_dev || !homo || !occ_dev || !vk_dev) return fail(ctx, RJK_ERR_ARG, "vk: NULL pointer");
    const void* b = nullptr;
    int rc = check_common(ctx, spin_channel, prec, true, &b);
    if (rc) return rc;
    const int ns = (int)spin_channel;
    size_t nw[2] = {0, 0};
    for (int s = 0; s < ns; ++s) {
        nw[s] = homo[s] + 1;
        if (homo[s] >= num_state)
            return fail(ctx, RJK_ERR_ARG, "vk: homo[" + std::to_string(s) + "] >= num_state");
    }
    RJK_HIP(ctx, hipSetDevice(ctx->device));
    KPlan kp;
    k_plan_cols(kp, ctx->nbf, nw, ns, prec);
    void *cocc, *sign, *m;
    double* kpart;
    rc = k_plan_ws(ctx, kp, ns, prec, &cocc, &sign, &kpart, &m);
    if (rc) return rc;
    {
        PhaseScope ps(ctx, 4);
        if (is_bf16(prec))
            RJK_HIP(ctx, launch_bf_prep_c(ctx->stream, prec, eigv_dev, ctx->nbf, num_state, nw, occ_dev, ns, kp,
                                          cocc, sign));
        else
            RJK_HIP(ctx, launch_prep_cocc(ctx->stream, prec, eigv_dev, ctx->nbf, num_state, nw, occ_dev, ns, kp,
                                          cocc, sign));
    }
    return k_run(ctx, kp, ns, prec, b, cocc, sign, kpart, m, scaling_factor, vk_dev);
}

extern "C" int rjk_vk_upper_with_ri_v_sync(rjk_ctx* ctx, const double* eigv, size_t num_state, const size_t* homo,
                                           const double* occupation, size_t spin_channel, double scaling_factor,
                                           int prec, double* vk_upper) {
    if (!ctx) return RJK_ERR_ARG;
    if (!eigv || !homo || !occupation || !vk_upper) return fail(ctx, RJK_ERR_ARG, "vk: NULL host pointer");
    if (spin_channel != 1 && spin_channel != 2)
        return fail(ctx, RJK_ERR_ARG, "spin_channel must be 1 or 2, got " + std::to_string(spin_channel));
    const size_t ne = spin_channel * ctx->nbf * num_state, no = spin_channel * num_state;
    int rc = check_host_finite(ctx, eigv, ne, "eigv");
    if (rc) return rc;
    rc = check_host_finite(ctx, occupation, no, "occupation");
    if (rc) return rc;
    RJK_HIP(ctx, hipSetDevice(ctx->device));
    void* io = nullptr;
    const size_t out_n = spin_channel * ctx->npair;
    rc = io_get(ctx, (ne + no + out_n) * sizeof(double), &io);
    if (rc) return rc;
    double* d_e = (double*)io;
    double* d_o = d_e + ne;
    double* d_out = d_o + no;
    RJK_HIP(ctx, hipMemcpyAsync(d_e, eigv, ne * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
    RJK_HIP(ctx, hipMemcpyAsync(d_o, occupation, no * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
    rc = rjk_vk_upper_occ_dev(ctx, d_e, num_state, homo, d_o, spin_channel, scaling_factor, prec, d_out);
    if (rc) return rc;
    RJK_HIP(ctx, hipMemcpyAsync(vk_upper, d_out, out_n * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    RJK_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return RJK_OK;
}

namespace {


// Density-only K of a D that is not symmetric, all spins: K_s = scale · upper(Σ_P B_P D_s B_P)
// written as Σ_P (B_P D_s)(B_P I)ᵀ.  Both column sets, D's columns and the identity's, go
// through the FP64 half-transform (M_X = B_P D, M_Y = B_P exactly), and one DGEMM per P-chunk
// with inner dimension P·nk accumulates G = M_X M_Yᵀ over the chunk.  This is REST's
// T = B_P D, K_P = T B_P (src/scf_io/mod.rs:2730-2732) with the P-sum moved into the GEMM;
// it runs at FP64 whatever the requested precision (no SCF density takes it).
int vk_dm_general(rjk_ctx* ctx, const double* dm_dev, int ns, double scale, double* vk_dev) {
    const size_t nbf = ctx->nbf, n2 = nbf * nbf;
    hipStream_t st = ctx->stream;
    KPlan kp;
    kp.nbf_pad = round_up(nbf, kTileMN);
    kp.nk = round_up(nbf, kColPad);
    kp.col_cnt[0] = kp.nk;
    const size_t ccol = kp.nbf_pad * kp.nk * sizeof(double);
    const size_t gbytes = n2 * sizeof(double);
    const size_t perM = kp.nk * kp.nbf_pad * sizeof(double);   // one M' column block per P
    const size_t fixed = 2 * round_up(ccol, 256) + round_up(gbytes, 256) + 512;
    const size_t avail = ctx->ws_limit > fixed ? ctx->ws_limit - fixed : 0;
    const size_t kmax = (size_t)0x7fffffff / kp.nk;              // rocBLAS inner dimension is an int
    const size_t chunk = std::min({ctx->naux, std::max<size_t>(1, avail / (2 * perM)), kmax});
    Carve cv;
    const size_t ix = cv.add(ccol), iy = cv.add(ccol), ig = cv.add(gbytes), imx = cv.add(chunk * perM),
                 imy = cv.add(chunk * perM);
    void* ws = nullptr;
    int rc = ws_get(ctx, cv.total, &ws);
    if (rc) return rc;
    char* base = (char*)ws;
    double* cx = (double*)(base + cv.off[ix]);
    double* cy = (double*)(base + cv.off[iy]);
    double* g = (double*)(base + cv.off[ig]);
    double* mx = (double*)(base + cv.off[imx]);
    double* my = (double*)(base + cv.off[imy]);
    rocblas_handle h;
    rc = solver_handle(ctx, &h);
    if (rc) return rc;
    const double one = 1.0, zero = 0.0;
    PhaseScope ps(ctx, 4);
    for (int s = 0; s < ns; ++s) {
        RJK_HIP(ctx, launch_prep_general(st, dm_dev + s * n2, nbf, kp.nk, kp.nbf_pad, cx, cy));
        for (size_t P0 = 0; P0 < ctx->naux; P0 += chunk) {
            const size_t Pc = std::min(chunk, ctx->naux - P0);
            RJK_HIP(ctx, launch_khalf(st, RJK_PREC_FP64, ctx->b64, ctx->ldb, nbf, P0, Pc, kp, cx, mx));
            RJK_HIP(ctx, launch_khalf(st, RJK_PREC_FP64, ctx->b64, ctx->ldb, nbf, P0, Pc, kp, cy, my));
            if (rocblas_dgemm(h, rocblas_operation_none, rocblas_operation_transpose, (rocblas_int)nbf,
                              (rocblas_int)nbf, (rocblas_int)(Pc * kp.nk), &one, mx, (rocblas_int)kp.nbf_pad, my,
                              (rocblas_int)kp.nbf_pad, P0 ? &one : &zero, g, (rocblas_int)nbf) !=
                rocblas_status_success)
                return fail(ctx, RJK_ERR_HIP, "vk (density-only, general D): rocblas_dgemm failed");
        }
        RJK_HIP(ctx, launch_pack_upper(st, g, nbf, nbf, scale, vk_dev + s * ctx->npair));
    }
    return RJK_OK;
}

// Density-only K on device memory, K_s = scaling_factor · upper(Σ_P B_P D_s B_P) into vk_dev
// (spin_channel × npair).  A device reduction measures max|D − Dᵀ| against max|D| per spin.
// When every D is symmetric to kAsymTol, (D + Dᵀ)/2 is eigen-decomposed (rocSOLVER dsyevd) and
// the eigenpairs with |λ| > kEigRelCut·max|λ| go through the occupied route with signed
// weights λ; otherwise the general route above reads every D in full.
int vk_dm_device(rjk_ctx* ctx, const void* b, const double* dm_dev, size_t spin_channel, double scaling_factor,
                 int prec, double* vk_dev) {
    const size_t nbf = ctx->nbf, n2 = nbf * nbf;
    const int ns = (int)spin_channel;
    RJK_HIP(ctx, hipSetDevice(ctx->device));
    rocblas_handle h;
    int rc = solver_handle(ctx, &h);
    if (rc) return rc;
    hipStream_t st = ctx->stream;
    Carve cv;
    const size_t i_st = cv.add(2 * kAsymBlocks * spin_channel * sizeof(double));
    const size_t i_vec = cv.add(spin_channel * n2 * sizeof(double));
    const size_t i_lam = cv.add(spin_channel * nbf * sizeof(double));
    const size_t i_E = cv.add(nbf * sizeof(double));
    const size_t i_sel = cv.add(spin_channel * nbf * sizeof(double));
    const size_t i_idx = cv.add(spin_channel * nbf * sizeof(int));
    const size_t i_info = cv.add(sizeof(int));
    void* aux = nullptr;
    rc = buf_get(ctx, &ctx->aux, &ctx->aux_size, cv.total, &aux);
    if (rc) return rc;
    char* base = (char*)aux;
    double* d_stats = (double*)(base + cv.off[i_st]);
    double* d_vec = (double*)(base + cv.off[i_vec]);
    double* d_lam = (double*)(base + cv.off[i_lam]);
    double* d_E = (double*)(base + cv.off[i_E]);
    double* d_sel = (double*)(base + cv.off[i_sel]);
    int* d_idx = (int*)(base + cv.off[i_idx]);
    int* d_info = (int*)(base + cv.off[i_info]);
    // symmetry of every D
    const unsigned nblk = asym_stats_blocks(nbf);
    std::vector<double> stats(2 * nblk * spin_channel);
    for (int s = 0; s < ns; ++s) {
        RJK_HIP(ctx, launch_asym_stats(st, dm_dev + s * n2, nbf, d_stats + 2 * kAsymBlocks * s));
        RJK_HIP(ctx, hipMemcpyAsync(stats.data() + 2 * nblk * s, d_stats + 2 * kAsymBlocks * s,
                                    2 * nblk * sizeof(double), hipMemcpyDeviceToHost, st));
    }
    RJK_HIP(ctx, hipStreamSynchronize(st));
    bool general = false;
    for (int s = 0; s < ns; ++s) {
        double asym = 0, amax = 0;
        for (unsigned q = 0; q < nblk; ++q) {
            asym = std::max(asym, stats[2 * nblk * s + 2 * q]);
            amax = std::max(amax, stats[2 * nblk * s + 2 * q + 1]);
        }
        if (!std::isfinite(amax)) return fail(ctx, RJK_ERR_NONFINITE, "dm holds NaN/Inf");
        if (asym > kAsymTol * std::max(amax, 1e-300)) general = true;
    }
    if (general) return vk_dm_general(ctx, dm_dev, ns, scaling_factor, vk_dev);
    std::vector<double> lam(nbf);
    std::vector<int> sel_idx[2];
    std::vector<double> sel_lam[2];
    size_t ncol[2] = {0, 0};
    for (int s = 0; s < ns; ++s) {
        RJK_HIP(ctx, launch_symmetrize(st, dm_dev + s * n2, nbf, d_vec + s * n2));
        rocblas_status rs = rocsolver_dsyevd(h, rocblas_evect_original, rocblas_fill_upper, (rocblas_int)nbf,
                                             d_vec + s * n2, (rocblas_int)nbf, d_lam + s * nbf, d_E, d_info);
        if (rs != rocblas_status_success) return fail(ctx, RJK_ERR_NOSOLVER, "rocsolver_dsyevd failed");
        int info = 0;
        RJK_HIP(ctx, hipMemcpyAsync(lam.data(), d_lam + s * nbf, nbf * sizeof(double), hipMemcpyDeviceToHost, st));
        RJK_HIP(ctx, hipMemcpyAsync(&info, d_info, sizeof(int), hipMemcpyDeviceToHost, st));
        RJK_HIP(ctx, hipStreamSynchronize(st));
        if (info != 0) return fail(ctx, RJK_ERR_NOSOLVER, "rocsolver_dsyevd info = " + std::to_string(info));
        double lmax = 0;
        for (double v : lam) {
            if (!std::isfinite(v)) return fail(ctx, RJK_ERR_NOSOLVER, "vk (density-only): non-finite eigenvalue");
            lmax = std::max(lmax, std::fabs(v));
        }
        for (size_t q = 0; q < nbf; ++q)
            if (std::fabs(lam[q]) > kEigRelCut * lmax) {
                sel_idx[s].push_back((int)q);
                sel_lam[s].push_back(lam[q]);
            }
        ncol[s] = sel_idx[s].size();
    }
    KPlan kp;
    k_plan_cols(kp, nbf, ncol, ns, prec);
    void *cocc, *sign, *m;
    double* kpart;
    rc = k_plan_ws(ctx, kp, ns, prec, &cocc, &sign, &kpart, &m);
    if (rc) return rc;
    const size_t es = k_elem_size(prec);
    RJK_HIP(ctx, hipMemsetAsync(cocc, 0, kp.nbf_pad * kp.nk * es, st));
    RJK_HIP(ctx, hipMemsetAsync(sign, 0, kp.nk * es, st));
    for (int s = 0; s < ns; ++s) {
        if (!ncol[s]) continue;
        RJK_HIP(ctx, hipMemcpyAsync(d_idx + s * nbf, sel_idx[s].data(), ncol[s] * sizeof(int), hipMemcpyHostToDevice,
                                    st));
        RJK_HIP(ctx, hipMemcpyAsync(d_sel + s * nbf, sel_lam[s].data(), ncol[s] * sizeof(double),
                                    hipMemcpyHostToDevice, st));
        if (is_bf16(prec))
            RJK_HIP(ctx, launch_bf_prep_c_idx(st, prec, d_vec + s * n2, nbf, d_idx + s * nbf, d_sel + s * nbf,
                                              ncol[s], kp.col_off[s], kp, cocc, sign));
        else
            RJK_HIP(ctx, launch_prep_cocc_idx(st, prec, d_vec + s * n2, nbf, d_idx + s * nbf, d_sel + s * nbf,
                                              ncol[s], kp.col_off[s], kp, cocc, sign));
    }
    rc = k_run(ctx, kp, ns, prec, b, cocc, sign, kpart, m, scaling_factor, vk_dev);
    if (rc) return rc;
    // sel_idx / sel_lam are host sources of the async copies above
    RJK_HIP(ctx, hipStreamSynchronize(st));
    return RJK_OK;
}

int vk_dm_host(rjk_ctx* ctx, const double* dm, size_t spin_channel, double scaling_factor, int prec,
               double* vk_upper) {
    if (!ctx) return RJK_ERR_ARG;
    if (!dm || !vk_upper) return fail(ctx, RJK_ERR_ARG, "vk: NULL host pointer");
    const void* b = nullptr;
    int rc = check_common(ctx, spin_channel, prec, true, &b);
    if (rc) return rc;
    const size_t nbf = ctx->nbf, n2 = nbf * nbf;
    rc = check_host_finite(ctx, dm, spin_channel * n2, "dm");
    if (rc) return rc;
    RJK_HIP(ctx, hipSetDevice(ctx->device));
    hipStream_t st = ctx->stream;
    void* io = nullptr;
    const size_t out_n = spin_channel * ctx->npair;
    rc = io_get(ctx, (spin_channel * n2 + out_n) * sizeof(double), &io);
    if (rc) return rc;
    double* d_dm = (double*)io;
    double* d_out = d_dm + spin_channel * n2;
    RJK_HIP(ctx, hipMemcpyAsync(d_dm, dm, spin_channel * n2 * sizeof(double), hipMemcpyHostToDevice, st));
    rc = vk_dm_device(ctx, b, d_dm, spin_channel, scaling_factor, prec, d_out);
    if (rc) return rc;
    RJK_HIP(ctx, hipMemcpyAsync(vk_upper, d_out, out_n * sizeof(double), hipMemcpyDeviceToHost, st));
    RJK_HIP(ctx, hipStreamSynchronize(st));
    return RJK_OK;
}

}  // namespace

extern "C" int rjk_vk_upper_dm_dev(rjk_ctx* ctx, const double* dm_dev, size_t spin_channel, double scaling_factor,
                                  int prec, double* vk_dev) {
    if (!ctx) return RJK_ERR_ARG;
    if (!dm_dev || !vk_dev) return fail(ctx, RJK_ERR_ARG, "vk: NULL device pointer");
    const void* b = nullptr;
    int rc = check_common(ctx, spin_channel, prec, true, &b);
    if (rc) return rc;
    return vk_dm_device(ctx, b, dm_dev, spin_channel, scaling_factor, prec, vk_dev);
}

extern "C" int rjk_vk_upper_with_rimatr_use_dm_only_sync(rjk_ctx* ctx, const double* dm, size_t spin_channel,
                                                         double scaling_factor, int prec, double* vk_upper) {
    return vk_dm_host(ctx, dm, spin_channel, scaling_factor, prec, vk_upper);
}

extern "C" int rjk_vk_upper_with_ri_v_use_dm_only_sync(rjk_ctx* ctx, const double* dm, size_t spin_channel,
                                                       double scaling_factor, int prec, double* vk_upper) {
    return vk_dm_host(ctx, dm, spin_channel, scaling_factor, prec, vk_upper);
}

// ---------------------------------------------------------------------------
// measurement
// ---------------------------------------------------------------------------
extern "C" int rjk_profile_enable(rjk_ctx* ctx, int on) {
    if (!ctx) return RJK_ERR_ARG;
    ctx->prof = on != 0;
    return RJK_OK;
}

namespace {
int profile_collect(rjk_ctx* ctx) {
    RJK_HIP(ctx, hipStreamSynchronize(ctx->stream));
    for (auto& e : ctx->ev_pending) {
        float t = 0;
        if (hipEventElapsedTime(&t, e.a, e.b) == hipSuccess) {
            ctx->prof_ms[e.prec][e.phase] += t;
            ctx->prof_n[e.prec][e.phase] += 1;
            ctx->prof_work[e.prec][e.phase] += e.work;
        }
        ctx->ev_free.push_back(e);
    }
    ctx->ev_pending.clear();
    return RJK_OK;
}
void profile_reset(rjk_ctx* ctx) {
    for (int p = 0; p < kNumPrec; ++p)
        for (int i = 0; i < kNumPhases; ++i) {
            ctx->prof_ms[p][i] = 0;
            ctx->prof_n[p][i] = 0;
            ctx->prof_work[p][i] = 0;
        }
}
}  // namespace

extern "C" int rjk_profile_read(rjk_ctx* ctx, double* ms, uint64_t* launches) {
    if (!ctx) return RJK_ERR_ARG;
    int rc = profile_collect(ctx);
    if (rc) return rc;
    for (int i = 0; i < kNumPhases; ++i) {
        if (ms) ms[i] = 0;
        if (launches) launches[i] = 0;
        for (int p = 0; p < kNumPrec; ++p) {
            if (ms) ms[i] += ctx->prof_ms[p][i];
            if (launches) launches[i] += ctx->prof_n[p][i];
        }
    }
    profile_reset(ctx);
    return RJK_OK;
}

extern "C" int rjk_profile_read_detail(rjk_ctx* ctx, double* ms, uint64_t* launches, double* work) {
    if (!ctx) return RJK_ERR_ARG;
    int rc = profile_collect(ctx);
    if (rc) return rc;
    for (int p = 0; p < kNumPrec; ++p)
        for (int i = 0; i < kNumPhases; ++i) {
            if (ms) ms[p * kNumPhases + i] = ctx->prof_ms[p][i];
            if (launches) launches[p * kNumPhases + i] = ctx->prof_n[p][i];
            if (work) work[p * kNumPhases + i] = ctx->prof_work[p][i];
        }
    profile_reset(ctx);
    return RJK_OK;
}

extern "C" int rjk_copy_dev(rjk_ctx* ctx, void* dst_dev, const void* src_dev, size_t bytes) {
    if (!ctx || !dst_dev || !src_dev) return ctx ? fail(ctx, RJK_ERR_ARG, "copy: NULL pointer") : RJK_ERR_ARG;
    RJK_HIP(ctx, hipSetDevice(ctx->device));
    RJK_HIP(ctx, launch_copy(ctx->stream, dst_dev, src_dev, bytes));
    return RJK_OK;
}

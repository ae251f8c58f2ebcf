// rjk_internal.hpp — context, workspace and launcher declarations shared by
// the C-ABI layer (rjk_api.cpp) and the kernel files (rjk_j.hip, rjk_k.hip,
// rjk_util.hip).
#pragma once
#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <string>
#include <vector>

#include "../../include/rest_jk.h"

struct _rocblas_handle;   // rocblas_handle is _rocblas_handle*

namespace rjk {

constexpr size_t kFp32Chain = 4096;   // products per FP32 accumulator in the K SYRK before FP64
constexpr size_t kLdAlign = 64;      // elements; device B columns start 512-B aligned (FP64)
constexpr int kNumPhases = 5;        // J-gamma, J-acc, K-half, K-syrk, K-other
constexpr int kTileMN = 128;         // nbf padding for the K tiles
constexpr int kTileK = 16;           // k-block of the K tiles
constexpr int kColPad = 64;          // occupied-column padding of M (all spins)
constexpr int kBfColGran = 64;       // BF16 K: per-spin column granule (the deepest SYRK k-block)
constexpr int kBfStoredPlanes = 3;   // BF16 planes kept per B element (RJK_STORE_BF16)
constexpr int kNumPrec = RJK_NUM_PREC; // profile rows: one per RJK_PREC_* value

inline bool is_bf16(int prec) {
    return prec == RJK_PREC_BF16X1 || prec == RJK_PREC_BF16X2 || prec == RJK_PREC_BF16X3;
}
inline int prof_row(int prec) { return prec >= 0 && prec < kNumPrec ? prec : 0; }

inline size_t round_up(size_t x, size_t a) { return (x + a - 1) / a * a; }

struct Status {
    int code = RJK_OK;
    std::string msg;
};

}  // namespace rjk

struct rjk_ctx {
    int device = 0;
    size_t nbf = 0, naux = 0, npair = 0, ldb = 0;
    hipStream_t own_stream = nullptr;
    hipStream_t stream = nullptr;
    double* b64 = nullptr;    // [naux][ldb]
    float* b32 = nullptr;     // [naux][ldb]
    void* bpl = nullptr;      // BF16 planes [kBfStoredPlanes][naux][nbf_pad][nbf_pad]
    size_t bpl_pad = 0;       // nbf_pad of the planes
    size_t ws_limit = size_t(4) << 30;
    void* ws = nullptr;       // grow-only workspace
    size_t ws_size = 0;
    void* io = nullptr;       // grow-only staging of the host-pointer entry points (inputs, outputs)
    size_t io_size = 0;
    void* aux = nullptr;      // grow-only scratch of the density-only route (eigenpairs, statistics)
    size_t aux_size = 0;
    void* scratch = nullptr;  // this device's allocation-free J scratch (address from hipGetSymbolAddress)
    void* solver = nullptr;   // rocblas_handle, created on first use (density-only K, SCF driver)
    void* scf_ws = nullptr;   // SCF driver buffers, grow-only, reused by every driver call
    size_t scf_ws_bytes = 0;
    double* scf_x = nullptr;          // SCF driver: orthogonaliser S^-1/2 of the last S seen
    std::vector<double> scf_s;        // that S (host copy, compared before reuse)
    std::string err;
    // phase timing
    bool prof = false;
    double prof_ms[rjk::kNumPrec][rjk::kNumPhases] = {};     // [RJK_PREC_*][phase]
    uint64_t prof_n[rjk::kNumPrec][rjk::kNumPhases] = {};
    double prof_work[rjk::kNumPrec][rjk::kNumPhases] = {};   // algorithmic bytes (J) or FLOP (K)
    struct EvPair { hipEvent_t a, b; int phase; int prec; double work; };
    std::vector<EvPair> ev_pending;
    std::vector<EvPair> ev_free;
};

namespace rjk {

// ---- error helpers ------------------------------------------------------
int fail(rjk_ctx* ctx, int code, const std::string& msg);
// the context's rocBLAS/rocSOLVER handle (rocblas_handle), bound to ctx->stream
int solver_handle(rjk_ctx* ctx, _rocblas_handle** h);
#define RJK_HIP(ctx, call)                                                        \
    do {                                                                          \
        hipError_t e_ = (call);                                                   \
        if (e_ != hipSuccess)                                                     \
            return ::rjk::fail((ctx), RJK_ERR_HIP,                               \
                               std::string(#call) + ": " + hipGetErrorString(e_)); \
    } while (0)

// workspace: returns a device buffer of at least `bytes` (grow-only)
int ws_get(rjk_ctx* ctx, size_t bytes, void** out);
// any grow-only device buffer of the context (*buf, *size): at least `bytes` in *out
int buf_get(rjk_ctx* ctx, void** buf, size_t* size, size_t bytes, void** out);
// the allocation-free J scratch of ctx->device: looked up once per device (rjk_create)
int scratch_init(rjk_ctx* ctx);
size_t scratch_bytes();
// J on a given workspace (at least j_ws_bytes); `timed` records phase events
size_t j_ws_bytes(const rjk_ctx* ctx, int nspin, int prec);
int j_run(rjk_ctx* ctx, const void* b, int prec, const double* dm_dev, int nspin, double scale, int full_dm,
          double* vj_dev, void* ws, bool timed);
// the B copy and J precision of a J call (checks spin_channel, B state and prec)
int j_source(rjk_ctx* ctx, size_t spin_channel, int prec, const void** bptr, int* jprec);
// device B storage: allocate the copies `mask` asks for; then refresh the FP32 copy and synchronise
int alloc_b(rjk_ctx* ctx, unsigned mask);
int finish_b(rjk_ctx* ctx);

// phase timing around a launch
struct PhaseScope {
    rjk_ctx* ctx;
    rjk_ctx::EvPair ev;
    bool on;
    PhaseScope(rjk_ctx* c, int phase, int prec = 0, double work = 0.0, bool allow = true);
    ~PhaseScope();
};

// ---- launchers (return hipError_t of the launch) ------------------------
// J
hipError_t launch_dm_weights(hipStream_t st, const double* dm, size_t nbf, size_t npair, size_t ldb,
                             int nspin, int full_dm, double* w);
hipError_t launch_jgamma(hipStream_t st, int prec, const void* b, size_t ldb, size_t naux, int nspin,
                         const double* w, size_t nchunk, double* gpart);
hipError_t launch_gamma_reduce(hipStream_t st, const double* gpart, size_t nchunk, int nspin, size_t naux,
                               double* gamma);
hipError_t launch_jacc(hipStream_t st, int prec, const void* b, size_t ldb, size_t naux, int nspin,
                       const double* gamma, size_t nsplit, double* jpart);
hipError_t launch_jreduce(hipStream_t st, const double* jpart, size_t nsplit, int nspin, size_t ldb,
                          size_t npair, double scale, double* vj);
void j_plan(size_t ldb, size_t naux, int prec, size_t* nchunk, size_t* nsplit);

// K
struct KPlan {
    size_t nbf_pad = 0;      // multiple of kTileMN
    size_t nk = 0;           // padded occupied columns, all spins (multiple of kColPad)
    size_t col_off[2] = {0, 0}, col_cnt[2] = {0, 0};   // per-spin column ranges (multiples of 16)
    size_t chunk = 0;        // P per chunk
    size_t nsplit = 0;       // split-K slots of the SYRK
};
size_t k_elem_size(int prec);   // bytes per element of C', M' and signs (all planes) for a K precision
hipError_t launch_prep_cocc(hipStream_t st, int prec, const double* eigv, size_t nbf, size_t nstate,
                            const size_t* nw, const double* occ, int nspin, const KPlan& kp, void* cocc,
                            void* sign);
hipError_t launch_prep_cocc_idx(hipStream_t st, int prec, const double* vecs, size_t nbf, const int* idx,
                                const double* lam, size_t ncol, size_t col_off, const KPlan& kp, void* cocc,
                                void* sign);
hipError_t launch_khalf(hipStream_t st, int prec, const void* b, size_t ldb, size_t nbf, size_t P0, size_t Pc,
                        const KPlan& kp, const void* cocc, void* m);
hipError_t launch_ksyrk(hipStream_t st, int prec, const void* m, size_t Pc, const KPlan& kp, int nspin,
                        const void* sign, int beta, double* kpart);
hipError_t launch_kreduce(hipStream_t st, const double* kpart, const KPlan& kp, int nspin, size_t nbf,
                          double scale, double* vk);

// K on BF16 planes (rjk_kbf.hip): C' planes [S][nk][nbf_pad], sign masks [nk] (uint16),
// M' planes [S][Pc][nbf_pad][nk]
int bf_planes(int prec);
hipError_t launch_bf_planes_b(hipStream_t st, const double* b64, size_t ldb, size_t nbf, size_t nbf_pad,
                              size_t naux, void* planes);
hipError_t launch_bf_prep_c(hipStream_t st, int prec, const double* eigv, size_t nbf, size_t nstate,
                            const size_t* nw, const double* occ, int nspin, const KPlan& kp, void* ct, void* smask);
hipError_t launch_bf_prep_c_idx(hipStream_t st, int prec, const double* vecs, size_t nbf, const int* idx,
                                const double* lam, size_t ncol, size_t col_off, const KPlan& kp, void* ct,
                                void* smask);
hipError_t launch_bf_khalf(hipStream_t st, int prec, const void* bpl, size_t naux, size_t nbf, size_t P0, size_t Pc,
                           const KPlan& kp, const void* ct, void* mpl);
hipError_t launch_bf_ksyrk(hipStream_t st, int prec, const void* mpl, size_t Pc, const KPlan& kp, int nspin,
                           const void* smask, int beta, double* kpart);

// util
hipError_t launch_fill_synth(hipStream_t st, uint64_t seed, size_t npair, size_t naux, size_t ldb, double* b);
hipError_t launch_pack_ri3fn(hipStream_t st, const double* b3, size_t nbf, size_t nP, size_t ldb, double* b);
hipError_t launch_pad_rimatr(hipStream_t st, const double* src, size_t npair, size_t nP, size_t ldb, double* b);
hipError_t launch_to_fp32(hipStream_t st, const double* b, size_t n, float* out);
hipError_t launch_symmetrize(hipStream_t st, const double* dm, size_t nbf, double* out);
hipError_t launch_copy(hipStream_t st, void* dst, const void* src, size_t bytes);
hipError_t launch_check_finite(hipStream_t st, const double* x, size_t n, int* flag);
// density-only K on a general D: asymmetry statistics (2 doubles per block, at most
// kAsymBlocks blocks), the column sets [D | I] and the packed upper triangle of G
constexpr size_t kAsymBlocks = 256;
unsigned asym_stats_blocks(size_t nbf);
hipError_t launch_asym_stats(hipStream_t st, const double* dm, size_t nbf, double* out);
hipError_t launch_prep_general(hipStream_t st, const double* dm, size_t nbf, size_t nk, size_t nrow, double* cx,
                               double* cy);
hipError_t launch_pack_upper(hipStream_t st, const double* g, size_t nbf, size_t ldg, double scale, double* vk);

}  // namespace rjk

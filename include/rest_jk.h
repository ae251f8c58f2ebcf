/*
 * rest_jk.h — C ABI of the MI355X RI-V Coulomb/exchange (J/K) build.
 *
 * This library is a drop-in for the vj/vk builders of REST's SCF
 * (src/scf_io/mod.rs:2386-2834).  The three-centre tensor B[μν,P] stays
 * resident in HBM in the packed pair layout of REST's `rimatr`
 * (src/molecule_io/mod.rs:1563-1721): element (p,P) with p = j(j+1)/2 + i,
 * i <= j, every auxiliary column P contiguous.  Device columns are padded to
 * a 256-byte aligned leading dimension; the host never sees the padding.
 *
 * Conventions (all entry points):
 *   - return 0 (RJK_OK) on success, a positive RJK_ERR_* code otherwise;
 *     rjk_last_error() gives the message.  REST panics on a size mismatch
 *     (unwrap); this ABI reports it as RJK_ERR_ARG instead.
 *   - matrices are column-major (REST's MatrixFull); packed-upper outputs
 *     follow REST's MatrixUpper order: p = j(j+1)/2 + i for row i <= col j.
 *   - `spin_channel` is 1 (RHF) or 2 (UHF).  Outputs hold spin_channel
 *     packed vectors of length npair = nbf(nbf+1)/2 back to back.  REST
 *     returns a 1-element dummy for the unused spin slot when
 *     spin_channel == 1 (src/scf_io/mod.rs:2494); the host wrapper, not this
 *     ABI, recreates that slot.
 *   - `scaling_factor` is applied last, as in REST.
 *   - host-pointer entry points copy in, compute and copy out synchronously.
 *     `_dev` entry points take device pointers, are ordered on the context's
 *     stream and return without synchronising.
 *   - a context is used by one host thread at a time (REST's builders are
 *     not re-entrant either: they flip the global OpenBLAS thread count,
 *     src/scf_io/mod.rs:2709-2711).
 */
#ifndef REST_JK_H
#define REST_JK_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct rjk_ctx rjk_ctx;

enum {
    RJK_OK = 0,
    RJK_ERR_ARG = 1,       /* bad argument or size mismatch            */
    RJK_ERR_HIP = 2,       /* a HIP runtime / library call failed       */
    RJK_ERR_NOMEM = 3,     /* device allocation failed                  */
    RJK_ERR_STATE = 4,     /* B (or the requested precision copy) unset */
    RJK_ERR_NONFINITE = 5, /* NaN/Inf found in an input                 */
    RJK_ERR_NOSOLVER = 6   /* eigensolver failed (density-only K)       */
};

/* Arithmetic used by a J/K call. */
enum {
    RJK_PREC_FP64 = 0,   /* FP64 loads, FP64 MFMA / FMA, FP64 accumulation          */
    RJK_PREC_FP32 = 1,   /* FP32 copy of B; J: FP64 accumulation                    */
    RJK_PREC_BF16X2 = 2, /* K: two BF16 planes per operand, 3 products (~2^-16)     */
    RJK_PREC_BF16X3 = 3, /* K: three BF16 planes per operand, 6 products (~2^-24)   */
    RJK_PREC_BF16X1 = 4, /* K: one BF16 plane per operand, 1 product (~2^-8): cheap
                          * estimates (the energy rule of a guess), not SCF steps   */
    RJK_NUM_PREC = 5     /* rows of rjk_profile_read_detail                         */
    /* The BF16 precisions run K on BF16 MFMA with FP32 accumulation and FP64
     * split sums; J then runs on the FP32 copy of B when it is stored, else on
     * the FP64 copy. */
};

/* Which device copies of B to keep (bit mask for rjk_set_rimatr*). */
enum {
    RJK_STORE_FP64 = 1,
    RJK_STORE_FP32 = 2,
    RJK_STORE_BF16 = 4   /* three BF16 planes of every B_P as a full padded
                          * square: 6·naux·nbf_pad² bytes, nbf_pad = nbf
                          * rounded up to 128 (serves both BF16 precisions) */
};

/* ---- context ------------------------------------------------------------ */

/* Create a context on HIP device `device` for num_basis orbitals and
 * num_auxbas auxiliary functions.  Replaces the shape fields REST reads from
 * `rimatr`/`ri3fn` (src/scf_io/mod.rs:2500-2502, :2441-2443). */
int rjk_create(int device, size_t num_basis, size_t num_auxbas, rjk_ctx** out);
int rjk_destroy(rjk_ctx* ctx);
/* Message of the last error on `ctx`, or of the calling thread's last failed
 * rjk_create when ctx is NULL.  Never NULL. */
const char* rjk_last_error(const rjk_ctx* ctx);
/* Shape queries. */
size_t rjk_num_basis(const rjk_ctx* ctx);
size_t rjk_num_auxbas(const rjk_ctx* ctx);
size_t rjk_num_baspar(const rjk_ctx* ctx);
/* Leading dimension (elements) of one device B column. */
size_t rjk_ld_baspar(const rjk_ctx* ctx);
/* Run on an external HIP stream (hipStream_t passed as void*); NULL restores
 * the context's own stream.  The stream must have default priority. */
int rjk_set_stream(rjk_ctx* ctx, void* stream);
int rjk_synchronize(rjk_ctx* ctx);
/* Upper bound (bytes) for the K workspace; 0 = default (4 GiB). */
int rjk_set_workspace_limit(rjk_ctx* ctx, size_t bytes);

/* ---- B storage (replaces the SCF.rimatr / SCF.ri3fn fields,
 *      src/scf_io/mod.rs:50-51, filled at :264-291) -------------------------- */

/* Upload packed B, host layout [naux][npair] (REST rimatr, MatrixFull
 * [npair,naux] column-major).  store_mask selects the device copies. */
int rjk_set_rimatr(rjk_ctx* ctx, const double* rimatr, unsigned store_mask);
/* Same from a device pointer (same compact layout). */
int rjk_set_rimatr_dev(rjk_ctx* ctx, const double* rimatr_dev, unsigned store_mask);
/* Upload full B, host layout [naux][nbf][nbf] (REST RIFull, element (i,j,P)
 * at i + nbf*j + nbf*nbf*P, src/scf_io/mod.rs:2400-2408).  Every B_P is
 * stored through its upper triangle (row <= col); B_P is symmetric for every
 * RI producer in REST (src/molecule_io/mod.rs:1275-1393, :1725-1744). */
int rjk_set_ri3fn(rjk_ctx* ctx, const double* ri3fn, unsigned store_mask);
/* Fill B on the device with the seeded synthetic generator
 * b = (2u-1)*sqrt(3/naux), u = splitmix64(seed + (P*npair+p+1)*φ64)/2^64;
 * used for bandwidth/throughput runs. */
int rjk_fill_rimatr_synthetic(rjk_ctx* ctx, uint64_t seed, unsigned store_mask);
/* Device pointer of the stored FP64 B (for read-only inspection). */
const double* rjk_rimatr_dev(const rjk_ctx* ctx);
/* Copy the stored FP64 B back to the host in REST's rimatr layout
 * [naux][npair] (the SCF.rimatr field, src/scf_io/mod.rs:50-51). */
int rjk_get_rimatr(rjk_ctx* ctx, double* rimatr);

/* ---- J (Coulomb) -------------------------------------------------------- */

/* vj_upper_with_rimatr_v_sync (src/scf_io/mod.rs:2490-2561):
 *   γ_P = 2 Σ_{i<=j} B_{ij,P} D_{ij} − Σ_i B_{ii,P} D_{ii}   (upper of D only)
 *   J_p = scaling_factor · Σ_P B_{p,P} γ_P
 * dm: spin_channel matrices nbf×nbf column-major; vj_upper: spin_channel×npair. */
int rjk_vj_upper_with_rimatr_v_sync(rjk_ctx* ctx, const double* dm, size_t spin_channel,
                                    double scaling_factor, int prec, double* vj_upper);
/* vj_upper_with_ri_v_sync (src/scf_io/mod.rs:2428-2488): γ_P = Σ_{k,l} B_{kl,P} D_{kl}
 * over the full D (for symmetric B_P this is the packed form applied to
 * (D+Dᵀ)/2); J = upper triangle of Σ_P B_P γ_P. */
int rjk_vj_upper_with_ri_v_sync(rjk_ctx* ctx, const double* dm, size_t spin_channel,
                                double scaling_factor, int prec, double* vj_upper);
/* Device form of both: dm_dev/vj_dev device pointers, `full_dm` selects the
 * ri3fn semantics (1) or the rimatr upper-triangle semantics (0). */
int rjk_vj_upper_dev(rjk_ctx* ctx, const double* dm_dev, size_t spin_channel,
                     double scaling_factor, int prec, int full_dm, double* vj_dev);
/* Allocation-free form of rjk_vj_upper_dev.  It has no counterpart in REST: it
 * is for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  Same
 * arguments, kernels and results as rjk_vj_upper_dev, but the workspace is one
 * fixed 128 MiB device array per GPU (held until the process ends) and the call
 * allocates nothing.  A shape whose J workspace exceeds 128 MiB (about
 * 8·spin_channel·npair·(1 + 4) bytes; nbf 1536 fits with two spins) returns
 * RJK_ERR_ARG.  Every call on one GPU shares that array and leaves its last
 * contents there: calls that overlap on one GPU race. */
int rjk_vj_upper_dev_noalloc(rjk_ctx* ctx, const double* dm_dev, size_t spin_channel,
                             double scaling_factor, int prec, int full_dm, double* vj_dev);

/* ---- K (exchange) ------------------------------------------------------- */

/* vk_upper_with_ri_v_sync (src/scf_io/mod.rs:2758-2834), occupied-orbital route:
 *   M_P = B_P C[:,0..homo_s], K_s = scaling_factor · upper(Σ_P M_P diag(occ_s) M_Pᵀ).
 * eigv: spin_channel matrices nbf×num_state column-major; homo[s] is the
 * index of the highest occupied state (nw = homo+1 columns are used);
 * occupation: spin_channel vectors of num_state entries. */
int rjk_vk_upper_with_ri_v_sync(rjk_ctx* ctx, const double* eigv, size_t num_state,
                                const size_t* homo, const double* occupation,
                                size_t spin_channel, double scaling_factor, int prec,
                                double* vk_upper);
/* Device form: eigv_dev (spin_channel × nbf × num_state) and occ_dev
 * (spin_channel × num_state) in HBM; homo on the host. */
int rjk_vk_upper_occ_dev(rjk_ctx* ctx, const double* eigv_dev, size_t num_state,
                         const size_t* homo, const double* occ_dev, size_t spin_channel,
                         double scaling_factor, int prec, double* vk_dev);
/* vk_upper_with_rimatr_use_dm_only_sync (src/scf_io/mod.rs:2703-2756) and
 * vk_upper_with_ri_v_use_dm_only_sync (:2640-2701):
 *   K_s = scaling_factor · upper(Σ_P B_P D_s B_P)
 * for any D, read in full (REST's _dsymm(B_P, D) then _dsymm(·, B_P), :2730-2732,
 * :2675-2677).  A symmetric D (max|D−Dᵀ| <= 1e-10·max|D| for every spin) is
 * factorised on the device as D = U Λ Uᵀ and the occupied route runs with
 * signed weights Λ at `prec`; any other D takes the general route at FP64,
 * K = upper(Σ_P (B_P D)(B_P)ᵀ) through the FP64 half-transform and one DGEMM
 * per block of auxiliary columns (about 6·nbf³ FLOP per column). */
int rjk_vk_upper_with_rimatr_use_dm_only_sync(rjk_ctx* ctx, const double* dm,
                                              size_t spin_channel, double scaling_factor,
                                              int prec, double* vk_upper);
int rjk_vk_upper_with_ri_v_use_dm_only_sync(rjk_ctx* ctx, const double* dm,
                                            size_t spin_channel, double scaling_factor,
                                            int prec, double* vk_upper);
/* Device form of the density-only route: dm_dev (spin_channel × nbf × nbf, column-major) and
 * vk_dev (spin_channel × npair) in HBM, the same semantics and routes as above; a D holding
 * NaN/Inf gives RJK_ERR_NONFINITE.  Synchronises the stream (the route is chosen on the host). */
int rjk_vk_upper_dm_dev(rjk_ctx* ctx, const double* dm_dev, size_t spin_channel,
                        double scaling_factor, int prec, double* vk_dev);

/* ---- measurement -------------------------------------------------------- */

/* Per-phase HIP-event timing of the kernels launched by this context.
 * Phases: 0 J-γ pass, 1 J-accumulate pass, 2 K half-transform, 3 K SYRK,
 * 4 K reduce/other.  rjk_profile_read fills ms[5] (summed milliseconds) and
 * launches[5] and resets the counters. */
int rjk_profile_enable(rjk_ctx* ctx, int on);
int rjk_profile_read(rjk_ctx* ctx, double* ms, uint64_t* launches);
/* The same counters split by arithmetic, [RJK_NUM_PREC][5] row-major (row = the
 * RJK_PREC_* value: FP64, FP32, BF16X2, BF16X3, BF16X1), plus the algorithmic work of those launches: bytes for the J
 * phases (npair·naux·s_B + 8·nspin·npair per pass), FLOP for the K phases
 * (2·nbf²·nk·P per half-transform, nbf(nbf+1)·nk·P per SYRK), 0 for phase 4.
 * Resets the counters like rjk_profile_read. */
int rjk_profile_read_detail(rjk_ctx* ctx, double* ms, uint64_t* launches, double* work);
/* A plain device copy of `bytes` bytes (HBM-traffic calibration for counters). */
int rjk_copy_dev(rjk_ctx* ctx, void* dst_dev, const void* src_dev, size_t bytes);

#ifdef __cplusplus
}
#endif
#endif /* REST_JK_H */

/*
 * rest_batch.h — J and K of many independent RI problems (molecules of
 * uneven size) in one grouped, persistent launch per stage, replayed from a
 * hipGraph.
 *
 * Per problem i (closed shell, nbf_i orbitals, naux_i auxiliary functions,
 * nocc_i occupied orbitals with weights occ_i):
 *   J_i = scale_j · vj_upper_with_rimatr_v_sync(B_i, D_i)   (src/scf_io/mod.rs:2490-2561)
 *   K_i = scale_k · vk_upper_with_ri_v_sync(B_i, C_i, homo = nocc_i − 1, occ_i)  (:2758-2834)
 * exactly as rest_jk.h computes them for one problem; the batch adds no REST
 * counterpart (REST runs one molecule per process).  It serves throughput
 * over many small molecules, e.g. generating training densities.
 *
 * Layouts as in rest_jk.h: B [naux][npair] packed (host), D nbf×nbf and C
 * nbf×nocc column-major, outputs packed upper (npair each).  `prec`:
 *   RJK_PREC_FP64;
 *   RJK_PREC_FP32 (B stored in FP32, J accumulated in FP64, K on FP32 MFMA
 *     with FP64 partial sums every 4096 products, as the single-problem K);
 *   RJK_PREC_BF16X1 / X2 / X3 (B stored once, in FP32; J as at FP32; K on the
 *     BF16 matrix cores with every operand split into 1 / 2 / 3 BF16 planes
 *     and the plane products of rest_jk.h's BF16 precisions, the planes of B
 *     made from the FP32 copy inside the half-transform).
 *
 * Inputs and outputs live in device slots owned by the batch, so the
 * captured graph replays on new data: set inputs, run, read outputs.
 * rjk_batch_autotune times the launch configurations (the half-transform
 * tile per problem size and the persistent work-groups per CU) on the
 * batch's own data and keeps the fastest; the next run re-captures.
 */
#ifndef REST_BATCH_H
#define REST_BATCH_H

#include <stddef.h>
#include <stdint.h>

#include "rest_jk.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct rjk_batch rjk_batch;

/* Stages timed by rjk_batch_profile: 0 D weights, 1 γ, 2 J accumulation, 3 C' preparation,
 * 4 / 5 K half-transform (narrow / wide tile class), 6 / 7 K SYRK (narrow or, at FP64/FP32,
 * every problem / wide), 8 split reduction. */
#define RJK_BATCH_NUM_STAGES 9

/* Device bytes rjk_batch_create allocates for these problems at precision prec (B, inputs,
 * outputs and intermediates; 0 for invalid arguments).  No device work. */
size_t rjk_batch_device_bytes(size_t nprob, const size_t* nbf, const size_t* naux, const size_t* nocc, int prec);
/* Free and total bytes of device memory now (hipMemGetInfo on `device`). */
int rjk_batch_device_memory(int device, size_t* free_bytes, size_t* total_bytes);
/* Waves under a device-memory budget: problems 0..nprob-1, in the given order, split into
 * consecutive groups whose rjk_batch_device_bytes stays within budget_bytes (a problem larger than
 * the budget gets a wave of its own).  wave_start (nprob + 1 entries) receives the first problem
 * of every wave and, after the last, nprob; nwaves the number of waves.  budget_bytes = 0 takes
 * 90% of the device's free memory now. */
int rjk_batch_plan_waves(size_t nprob, const size_t* nbf, const size_t* naux, const size_t* nocc, int prec,
                         size_t budget_bytes, size_t* wave_start, size_t* nwaves);

int rjk_batch_create(int device, size_t nprob, const size_t* nbf, const size_t* naux, const size_t* nocc, int prec,
                     rjk_batch** out);
int rjk_batch_destroy(rjk_batch* batch);
const char* rjk_batch_last_error(const rjk_batch* batch);
size_t rjk_batch_size(const rjk_batch* batch);
/* Device bytes of the stored B of all problems (one J pass reads them once). */
size_t rjk_batch_bytes_b(const rjk_batch* batch);
/* Run on an external HIP stream (default priority); NULL = own stream. */
int rjk_batch_set_stream(rjk_batch* batch, void* stream);
/* B of problem i from the host, REST's rimatr layout [naux][npair]. */
int rjk_batch_set_rimatr(rjk_batch* batch, size_t i, const double* rimatr);
/* B of problem i from a device array in the padded rimatr layout [naux][ld], ld >= npair, holding
 * at least naux columns of problem i (for example rjk_rimatr_dev of an rjk_ctx of the same nbf and
 * naux, with ld = rjk_ld_baspar); converted to the batch precision.  Synchronises. */
int rjk_batch_set_rimatr_dev(rjk_batch* batch, size_t i, const double* rimatr_dev, size_t ld);
/* Every problem's B from the seeded synthetic generator of rjk_fill_rimatr_synthetic (seed + i). */
int rjk_batch_fill_synthetic(rjk_batch* batch, uint64_t seed);
/* Problem i takes part in later runs (active != 0, the default) or is skipped by every kernel
 * (its outputs keep their last values); the graph is not captured again.  Synchronises. */
int rjk_batch_set_active(rjk_batch* batch, size_t i, int active);
/* D (nbf×nbf), orbitals (nbf×nocc) and occupation weights (nocc) of problem i, host pointers. */
int rjk_batch_set_inputs(rjk_batch* batch, size_t i, const double* dm, const double* eigv, const double* occ);
/* J and K of problem i to the host (either may be NULL); call after rjk_batch_run. */
int rjk_batch_get_outputs(rjk_batch* batch, size_t i, double* vj, double* vk);
/* J and K of every problem: grouped kernels captured into a hipGraph on the first call (or after a
 * configuration or scale change) and replayed; stream-ordered, returns without synchronising. */
int rjk_batch_run(rjk_batch* batch, double scale_j, double scale_k);
int rjk_batch_synchronize(rjk_batch* batch);
/* One J/K of every problem launched stage by stage outside the graph, with HIP events between the
 * stages on the batch's stream; stage_ms[RJK_BATCH_NUM_STAGES] receives each stage's time.
 * Synchronises; the next rjk_batch_run captures the graph again. */
int rjk_batch_profile(rjk_batch* batch, double scale_j, double scale_k, double* stage_ms);
/* Time every candidate configuration `reps` times, keep the fastest; best_ms (may be NULL) receives its time. */
int rjk_batch_autotune(rjk_batch* batch, int reps, double* best_ms);
/* Current configuration: problems with nbf_pad < khalf_cut use the narrow tiles (FP64/FP32: 64×64
 * half-transform tiles, else 128×64; BF16: 64-wide half-transform and SYRK tiles, else 128-wide
 * where nbf_pad is a multiple of 128); waves_per_cu persistent work-groups per CU for the tile kernels. */
int rjk_batch_config(const rjk_batch* batch, int* khalf_cut, int* waves_per_cu);
int rjk_batch_set_config(rjk_batch* batch, int khalf_cut, int waves_per_cu);

#ifdef __cplusplus
}
#endif
#endif /* REST_BATCH_H */

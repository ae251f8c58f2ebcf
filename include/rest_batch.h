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
 * RJK_PREC_FP64 or RJK_PREC_FP32 (B stored in FP32, J accumulated in FP64,
 * K on FP32 MFMA with FP64 partial sums every 16384 products).
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
/* Every problem's B from the seeded synthetic generator of rjk_fill_rimatr_synthetic (seed + i). */
int rjk_batch_fill_synthetic(rjk_batch* batch, uint64_t seed);
/* D (nbf×nbf), orbitals (nbf×nocc) and occupation weights (nocc) of problem i, host pointers. */
int rjk_batch_set_inputs(rjk_batch* batch, size_t i, const double* dm, const double* eigv, const double* occ);
/* J and K of problem i to the host (either may be NULL); call after rjk_batch_run. */
int rjk_batch_get_outputs(rjk_batch* batch, size_t i, double* vj, double* vk);
/* J and K of every problem: grouped kernels captured into a hipGraph on the first call (or after a
 * configuration or scale change) and replayed; stream-ordered, returns without synchronising. */
int rjk_batch_run(rjk_batch* batch, double scale_j, double scale_k);
int rjk_batch_synchronize(rjk_batch* batch);
/* Time every candidate configuration `reps` times, keep the fastest; best_ms (may be NULL) receives its time. */
int rjk_batch_autotune(rjk_batch* batch, int reps, double* best_ms);
/* Current configuration: problems with nbf_pad < khalf_cut use 64×64 half-transform tiles, others
 * 128×64; waves_per_cu persistent work-groups per CU for the tile kernels. */
int rjk_batch_config(const rjk_batch* batch, int* khalf_cut, int* waves_per_cu);
int rjk_batch_set_config(rjk_batch* batch, int khalf_cut, int waves_per_cu);

#ifdef __cplusplus
}
#endif
#endif /* REST_BATCH_H */

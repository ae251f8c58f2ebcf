/*
 * rest_ri.h — C ABI of the RI-V three-centre producer (the B[μν,P] that the
 * J/K build in rest_jk.h consumes) and of the one-electron matrices the SCF
 * around it needs.
 *
 * REST computes these with libcint through rest_libcint:
 *   - basis normalisation       src/basis_io/mod.rs:242-341 (gto_norm, then unit
 *                               self-overlap of every contracted function)
 *   - one-electron matrices     Molecule::int_ij_matrixupper, src/molecule_io/mod.rs:700-760
 *   - auxiliary Coulomb metric  Molecule::int_ij_aux_columb, :1072-1103
 *   - packed B                  Molecule::prepare_rimatr_for_ri_v_rayon_v02, :1563-1721
 *                               B = (μν|Q) · V^{-1/2}, V^{-1/2} = lapack_power(-0.5, 1e-8)
 * This library computes them with its own McMurchie–Davidson engine on real
 * solid-harmonic (spherical) shells up to l = 4, and writes B straight into
 * the HBM-resident storage of an rjk_ctx.
 *
 * Conventions: 0 = success, RJK_ERR_* codes of rest_jk.h otherwise;
 * coordinates in bohr; matrices column-major; packed-upper order
 * p = j(j+1)/2 + i (i <= j).  Spherical function order inside a shell:
 * m = -l..l for l >= 2, and (x, y, z) for p shells (libcint order).
 */
#ifndef REST_RI_H
#define REST_RI_H

#include <stddef.h>
#include <stdint.h>

#include "rest_jk.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct rsi_basis rsi_basis;

/* A basis: nshell shells, shell s on atom shell_atom[s] (centre atom_xyz[3*a..]),
 * angular momentum shell_l[s] (0..4), shell_nprim[s] primitives whose
 * exponents and contraction coefficients (as printed in basis-set files, i.e.
 * for normalised primitives) follow back to back in exps/coefs.  One
 * contraction per shell.  Shells are kept in the given order. */
int rsi_basis_create(size_t nshell, const int* shell_atom, const int* shell_l, const int* shell_nprim,
                     const double* exps, const double* coefs, size_t natm, const double* atom_xyz,
                     rsi_basis** out);
int rsi_basis_destroy(rsi_basis* b);
size_t rsi_basis_nbf(const rsi_basis* b);
const char* rsi_last_error(void);   /* thread-local message of the last failure */

/* The engine's Cartesian -> real solid harmonic rows for angular momentum l
 * (0..4): (2l+1) × (l+1)(l+2)/2 row-major, m = -l..l (p: x, y, z), Cartesian
 * components ordered lx = l..0, then ly = l-lx..0.  Rows are not normalised
 * (each shell rescales them to unit-norm functions).  No REST counterpart;
 * it lets callers rotate density blocks in the engine's own convention. */
int rsi_cart2sph(int l, double* out);

/* One-electron matrices, full nbf×nbf column-major (NULL outputs are skipped):
 * overlap S, kinetic T, nuclear attraction V = -Σ_C Z_C <μ|1/|r-C||ν> over the
 * point charges (charge[c], xyz[3c..]). */
int rsi_int1e(const rsi_basis* orb, size_t ncharge, const double* charge, const double* charge_xyz,
              double* S, double* T, double* V);

/* Auxiliary Coulomb metric (P|Q), naux×naux column-major. */
int rsi_int2c2e(const rsi_basis* aux, double* V);

/* Three-centre integrals (μν|P) for the pair columns j in [j0, j1) of the
 * orbital basis (j0, j1 on shell boundaries), all P: out[P][q] with
 * q = p - j0(j0+1)/2 the packed pair index relative to the slab. */
int rsi_int3c2e_packed(const rsi_basis* orb, const rsi_basis* aux, size_t j0, size_t j1, double* out);
/* The same slab computed by the device kernels on ctx's device and stream
 * into the device array out_dev[naux][nq]; synchronises before returning.
 * Orbital shells up to d and auxiliary shells up to g (RJK_ERR_ARG beyond).
 * REST has no device counterpart; it replaces the libcint int3c2e calls of
 * prepare_rimatr_for_ri_v_rayon_v02 (src/molecule_io/mod.rs:1647). */
int rsi_int3c2e_packed_dev(rjk_ctx* ctx, const rsi_basis* orb, const rsi_basis* aux, size_t j0, size_t j1,
                           double* out_dev);

/* Build packed B = (μν|Q)·V^{-1/2} into the FP64 storage of ctx (and its FP32
 * copy when store_mask asks for it): V^{-1/2} = U diag(λ^{-1/2}) Uᵀ over the
 * eigenpairs with λ > lin_dep_threshold (1e-8 in REST); the three-centre
 * integrals are computed in pair slabs by the device kernels (on the host
 * when an orbital shell is beyond d) and multiplied on the device.  nkept (may be NULL) receives the number of kept eigenvalues;
 * seconds (may be NULL) receives {3c integrals, V^-1/2, GEMM tail} times (device
 * kernel time for device integrals). */
int rsi_build_rimatr(rjk_ctx* ctx, const rsi_basis* orb, const rsi_basis* aux, double lin_dep_threshold,
                     unsigned store_mask, size_t* nkept, double* seconds);

#ifdef __cplusplus
}
#endif
#endif /* REST_RI_H */

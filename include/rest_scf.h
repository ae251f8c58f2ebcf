/*
 * rest_scf.h — C ABI of the device-resident RI-HF SCF driver around the
 * J/K build (closed shell RHF and open shell UHF).
 *
 * Follows REST's `scf()` (src/scf_io/mod.rs:3190-3268) for RHF with RI-V:
 * guess density → Fock (generate_hf_hamiltonian_ri_v, :1339-1381; energy
 * :1715-1802) → diagonalise F C = S C ε (diagonalize_hamiltonian, :1843) →
 * density (generate_density_matrix, :441-478) → loop with linear mixing
 * before start_diis_cycle and Pulay DIIS on e = F D S − S D F afterwards
 * (prepare_next_input, :2947-3091; diis_solver, :3138-3188), until |ΔE|,
 * ‖ΔD‖_F/nbf and ‖Δε‖₂ meet the thresholds (check_scf_convergence, :1875-1951).
 * As in REST, the first iteration is never taken as converged (the check
 * returns early while num_iter < 2, :1877-1879), the loop stops after
 * iteration max(max_scf_cycle, 2) (:1947), and after start_check_oscillation
 * DIIS steps an energy rise switches to mixing with an earlier Fock matrix
 * (:3002-3010, :3039-3048).
 *
 * All matrices stay in HBM; J/K use the context's resident B.  With
 * jk_prec = RJK_PREC_FP32 (or RJK_PREC_BF16X3 / RJK_PREC_BF16X2: K on the
 * BF16 planes of B, J on its FP32 copy) the J/K builds run at that precision
 * until the DIIS error max|FDS − SDF| falls below switch_diis_err, then switch to FP64
 * for the remaining iterations (the FP64 builds act as iterative
 * refinement) and restart DIIS from the first FP64 Fock matrix; convergence
 * is only declared on two consecutive FP64 energies.
 * Diagonalisation, DIIS, the density and the energy are FP64 throughout.
 */
#ifndef REST_SCF_H
#define REST_SCF_H

#include <stddef.h>
#include <stdint.h>

#include "rest_jk.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct rscf_ctrl {
    int num_max_diis;        /* REST num_max_diis (8 in the examples)      */
    int start_diis_cycle;    /* REST start_diis_cycle (1)                  */
    double mix_param;        /* REST mix_param (0.6)                       */
    int max_scf_cycle;       /* REST max_scf_cycle (100)                   */
    double scf_acc_rho;      /* REST scf_acc_rho (1e-5)                    */
    double scf_acc_eev;      /* REST scf_acc_eev (1e-5)                    */
    double scf_acc_etot;     /* REST scf_acc_etot (1e-8)                   */
    int jk_prec;             /* RJK_PREC_FP64 (plain) or FP32 / BF16X3 / BF16X2 (mixed) */
    double switch_diis_err;  /* mixed mode: switch to FP64 below this max|e| (1e-5) */
    int print_level;         /* 0 silent, 1 one line per iteration on stdout */
    int switch_mode;         /* mixed mode, at the switch: RSCF_SWITCH_RESTART (default) or _SHIFT */
    int jk_prec_start;       /* builds run at this precision until max|e| < start_switch_err, then
                              * at jk_prec (no DIIS restart); -1 (default): jk_prec throughout */
    double start_switch_err; /* 1e-3 */
    double guess_energy_max; /* return before the first iteration when the guess Fock build gives
                              * E(guess) >= this (result.guess_rejected = 1); HUGE_VAL (default): never */
    double pseudo_diag_err;  /* iterations with low-precision J/K and max|e| below this rotate the
                              * previous orbitals (pseudo-diagonalisation) instead of calling the
                              * eigensolver; full FP64 iterations diagonalise exactly; 0 (default): never */
    int jk_prec_increment;   /* -1 (default): every FP64 build is a full one.  Else, after the switch
                              * from low-precision J/K to FP64, the first FP64 build is a reference
                              * and later ones add J/K of D − D_ref at this precision (RJK_PREC_FP32,
                              * _BF16X2 or _BF16X3); from the reference on every iteration
                              * pseudo-diagonalises, and the outputs come from one exact
                              * diagonalisation at the end */
    double guess_err_max;    /* return before the first iteration when the guess Fock build's
                              * max|F D S − S D F| >= this (result.guess_rejected = 1); HUGE_VAL
                              * (default): never */
    int start_check_oscillation; /* REST start_check_oscillation (20, src/ctrl_io/mod.rs:220): once more
                              * than this many DIIS steps have recorded their energy, an energy rise
                              * replaces the DIIS Fock matrix by (1 − mix_param)·H[i−1] + mix_param·H[i+1]
                              * (src/scf_io/mod.rs:3002-3010, :3039-3048); energies of different J/K
                              * precisions are not compared */
    double inc_refresh;      /* incremental tail: when max|D − D_ref| > inc_refresh·max|D_ref|, the
                              * FP64 build is a full one and becomes the new reference (1e-5) */
    int inc_check;           /* 1: every incremental build is followed by a full FP64 build of the same
                              * density, and max|F_inc − F_full| / max|F_full| goes to the log
                              * (RSCF_LOG_INC_ERR); a test mode, 0 by default */
} rscf_ctrl;

/* What DIIS keeps at the precision switch.  RESTART: the stored low-precision
 * vectors are dropped.  SHIFT: the first FP64 iteration also builds F at the
 * low precision for the same density and adds δ = F(FP64) − F(low) to every
 * stored Fock matrix (error vectors: + δ D_k S − S D_k δ), so the history
 * continues at FP64 quality for the cost of one low-precision build. */
enum { RSCF_SWITCH_RESTART = 0, RSCF_SWITCH_SHIFT = 1 };

/* Per-iteration record fields (iter_log rows of RSCF_LOG_FIELDS doubles). */
enum { RSCF_LOG_E = 0, RSCF_LOG_DE, RSCF_LOG_DIIS_MAX, RSCF_LOG_DM, RSCF_LOG_EEV, RSCF_LOG_PREC,
       RSCF_LOG_JK_MS, RSCF_LOG_DIAG_MS, RSCF_LOG_ITER_MS,
       RSCF_LOG_DREF,      /* incremental build: max|D − D_ref| / max|D_ref|; else 0            */
       RSCF_LOG_INC_ERR,   /* inc_check: max|F_inc − F_full| / max|F_full|; else 0              */
       RSCF_LOG_OSC,       /* 1: the energy-rise mixing replaced the DIIS Fock matrix            */
       RSCF_LOG_EXACT,     /* 1: the orbitals came from the eigensolver, 0: pseudo-diagonalised  */
       RSCF_LOG_FIELDS };

typedef struct rscf_result {
    double energy;           /* final total energy (Eh)                    */
    int converged;
    int iterations;          /* SCF iterations after the guess step        */
    int fp64_iterations;     /* of which with FP64 J/K                     */
    double wall_seconds;     /* whole driver, guess Fock build included    */
    double jk_seconds, diag_seconds, diis_seconds;
    double guess_energy;     /* energy of the guess Fock build (the given density or orbitals) */
    int guess_rejected;      /* 1: guess_energy >= ctrl.guess_energy_max or guess_error >=
                              * ctrl.guess_err_max, and no iteration ran */
    int incremental_iterations; /* FP64 iterations whose J/K were increments on the reference build */
    double guess_error;      /* max|F D S − S D F| of the guess build */
    int reference_builds;    /* full FP64 reference builds of the incremental tail */
} rscf_result;

void rscf_default_ctrl(rscf_ctrl* ctrl);

/* Closed-shell RI-HF.  S, H: nbf×nbf column-major host matrices (overlap and
 * core Hamiltonian); enuc: nuclear repulsion; nelec: electron count (even);
 * d_guess: host nbf×nbf guess density or NULL for the core-Hamiltonian guess.
 * Outputs (each may be NULL): mo_coeff nbf×nbf, mo_energy nbf, dm nbf×nbf,
 * iter_log max(max_scf_cycle, 2) × RSCF_LOG_FIELDS (row-major). */
int rscf_rhf(rjk_ctx* ctx, const double* S, const double* H, double enuc, size_t nelec, const double* d_guess,
             const rscf_ctrl* ctrl, rscf_result* result, double* mo_coeff, double* mo_energy, double* dm,
             double* iter_log);

/* Unrestricted RI-HF (REST's UHF branch of the same functions): nalpha and
 * nbeta electrons with occupation 1, F_s = H + J[D_α + D_β] − K[D_s]
 * (src/scf_io/mod.rs:1347-1381), E = E_nuc + ½ Σ_s Σ D_s (H + F_s)
 * (:1673-1713), one DIIS over the concatenated α and β error vectors
 * (:3094-3136) and convergence on both spin densities (:1875-1951).
 * d_guess: 2 × nbf² (α then β) or NULL for the core-Hamiltonian guess.
 * Outputs hold α then β: mo_coeff 2×nbf², mo_energy 2×nbf, dm 2×nbf². */
int rscf_uhf(rjk_ctx* ctx, const double* S, const double* H, double enuc, size_t nalpha, size_t nbeta,
             const double* d_guess, const rscf_ctrl* ctrl, rscf_result* result, double* mo_coeff,
             double* mo_energy, double* dm, double* iter_log);

/* REST's density-guess step (initial_guess for sad / hdf5 densities,
 * src/initial_guess/mod.rs:18-26, :54-64): F(D0) with the density-only J/K
 * (generate_hf_hamiltonian_for_guess, src/scf_io/mod.rs:1657-1671), the
 * guess energy E(D0) (evaluate_hf_total_energy, :1673-1713, printed by REST
 * as "Initial guess HF energy"), F C = S C ε and the density D1 of those
 * orbitals, which is where REST's scf() starts.  nspin 1: nelec[0] is the
 * electron count, d0 and d1 are nbf²; nspin 2: nelec = {nα, nβ}, d0 and d1
 * are 2×nbf².  prec: J/K arithmetic (RJK_PREC_FP64 as in REST, or a lower
 * precision to rank guesses cheaply).  mo_coeff / mo_energy may be NULL.
 * energy_d1 (may be NULL) receives E(D1), the energy REST prints as "The
 * total energy ... by the initial guess" at the start of scf() (:3195-3199),
 * from the occupied-orbital J/K of the new orbitals. */
int rscf_guess_step(rjk_ctx* ctx, const double* S, const double* H, double enuc, size_t nspin, const size_t* nelec,
                    const double* d0, int prec, double* energy, double* d1, double* mo_coeff, double* mo_energy,
                    double* energy_d1);

/* rscf_guess_step with the guess density given also in factored form,
 * D0_s = Σ_k w0_s[k] c0_s[:,k] c0_s[:,k]ᵀ over the first ncol[s] columns of c0_s
 * (c0: nspin × nbf × nbf column-major; w0: nspin × nbf): the first K runs by the
 * occupied-orbital route over those columns instead of factorising D0 on the
 * device.  A SAD density is block-diagonal by atom, so its factors come from the
 * atomic blocks (no REST counterpart: REST's guess step takes the density,
 * src/initial_guess/mod.rs:54-64).  d0 must equal the factored sum; it serves J
 * and E(D0). */
int rscf_guess_step_factored(rjk_ctx* ctx, const double* S, const double* H, double enuc, size_t nspin,
                             const size_t* nelec, const double* d0, const double* c0, const double* w0,
                             const size_t* ncol, int prec, double* energy, double* d1, double* mo_coeff,
                             double* mo_energy, double* energy_d1);

/* rscf_rhf started from guess orbitals instead of a density: c_guess is
 * nbf×nbf column-major, its first nelec/2 columns occupied.  The first Fock
 * matrix uses the occupied-orbital K, as REST's scf() does after its guess
 * step (generate_hf_hamiltonian with eigenvectors present, :3193). */
int rscf_rhf_orbitals(rjk_ctx* ctx, const double* S, const double* H, double enuc, size_t nelec,
                      const double* c_guess, const rscf_ctrl* ctrl, rscf_result* result, double* mo_coeff,
                      double* mo_energy, double* dm, double* iter_log);

/* The Hartree-Fock energy of a given density, E = E_nuc + ½ Σ_s Σ D_s (H + F_s)
 * with F_s from the density-only J/K (evaluate_hf_total_energy after
 * generate_hf_hamiltonian_for_guess, src/scf_io/mod.rs:1657-1713), without
 * diagonalising.  nspin / nelec / dm as in rscf_guess_step. */
int rscf_energy(rjk_ctx* ctx, const double* S, const double* H, double enuc, size_t nspin, const size_t* nelec,
                const double* dm, int prec, double* energy);

/* McWeeny purification of a closed-shell density in the S metric (for a
 * predicted guess density; no REST counterpart): P = D/2,
 * P <- 3 P S P − 2 P S P S P `iterations` times, D_out = 2P; electrons
 * (may be NULL) receives 2 tr(P S).  nbf×nbf column-major host matrices;
 * GEMMs in FP64 on the device. */
int rscf_mcweeny(rjk_ctx* ctx, const double* S, const double* D, int iterations, double* D_out, double* electrons);

#ifdef __cplusplus
}
#endif
#endif /* REST_SCF_H */

"""CPU restatement of REST's RI-HF SCF loop (test infrastructure only).

Follows src/scf_io/mod.rs:
  - scf()                      :3190-3268  (guess → F, E → diagonalise → loop), RHF and UHF
  - ScfTraceRecord             :2846-2934  (num_iter from 0, density records [previous, current])
  - prepare_next_input         :2947-3091  (linear mixing before start_diis_cycle, Pulay DIIS after,
                                            energy records and the energy-rise mixing with
                                            prev_hamiltonian :3002-3010, :3039-3057, the DIIS-failure
                                            fallback :3061-3078)
  - generate_diis_error_vector :3094-3136  (e = F D S − S D F, spins concatenated)
  - diis_solver                :3138-3188  (c = B⁻¹1 / 1ᵀB⁻¹1 over at most num_max_diis vectors)
  - check_scf_convergence      :1875-1951  (nothing before num_iter 2; |ΔE|, ‖ΔD‖_F/nbf, ‖Δε‖₂;
                                            stop at num_iter >= max_scf_cycle)
  - generate_hf_hamiltonian_ri_v :1339-1381, energy :1715-1802, density :441-478
J and K come from the C oracle (oracle/rjk_oracle.c); linear algebra from numpy/scipy.  The
generalised eigensolver is a parameter (`diag`): REST calls LAPACK dspgvx through rest_tensors
(:1851), whose choice inside exactly degenerate levels is a property of that call.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np
import scipy.linalg as sla

from oracle import oracle as orc


@dataclass
class ScfCtrl:
    num_max_diis: int = 8
    start_diis_cycle: int = 1
    mix_param: float = 0.6
    max_scf_cycle: int = 100
    scf_acc_rho: float = 1e-5
    scf_acc_eev: float = 1e-5
    scf_acc_etot: float = 1e-8
    start_check_oscillation: int = 20


@dataclass
class ScfResult:
    energy: float
    converged: bool
    iterations: int
    energies: list = field(default_factory=list)
    eigenvalues: np.ndarray = None
    eigenvectors: np.ndarray = None
    density: np.ndarray = None
    fock: np.ndarray = None
    osc: list = field(default_factory=list)          # per iteration: the energy-rise mixing ran
    diis_failed: list = field(default_factory=list)  # per iteration: DIIS fell back to linear mixing
    densities: list = field(default_factory=list)    # per iteration: the new densities (per spin)


def _upper_idx(n):
    j = np.concatenate([np.full(k + 1, k) for k in range(n)])
    i = np.concatenate([np.arange(k + 1) for k in range(n)])
    return i, j


def _unpack(v, n, idx):
    m = np.zeros((n, n))
    m[idx] = v
    m[(idx[1], idx[0])] = v
    return m


def oracle_jk(rimatr):
    """(J, K) of a density from the C oracle (packed B), full matrices, unscaled."""
    def jk(D):
        n = D.shape[0]
        idx = _upper_idx(n)
        vj = orc.vj_upper_rimatr(rimatr, [D], 1, 1.0)[0]
        vk = orc.vk_upper_dm(rimatr, True, [D], 1, 1.0)[0]
        return _unpack(vj, n, idx), _unpack(vk, n, idx)
    return jk


def default_diag(F, S):
    return sla.eigh(F, S)


def _diis_coefficients(errs):
    """diis_solver (:3138-3188): c = B⁻¹1 / 1ᵀB⁻¹1, None when the LU inverse fails (_dinverse)."""
    Bm = np.array([[np.dot(a, b) for b in errs] for a in errs])
    try:
        inv = np.linalg.inv(Bm)
    except np.linalg.LinAlgError:
        return None
    if not np.all(np.isfinite(inv)):
        return None
    return inv.sum(axis=1) / inv.sum()


def _scf(S, fock_energy, densities, D_guess, ctrl, diag):
    """REST's scf() over spin lists: fock_energy(Ds) -> (Fs, E); densities(eigs) -> Ds."""
    ns = len(D_guess)
    n = S.shape[0]
    a = ctrl.mix_param
    mix = lambda Din, Dout: [Din[s] + a * (Dout[s] - Din[s]) for s in range(ns)]
    D_in = [np.array(d) for d in D_guess]
    Fs, E = fock_energy(D_in)                       # generate_hf_hamiltonian on the guess
    eig = [diag(F, S) for F in Fs]
    Ds = densities(eig)
    # ScfTraceRecord::initialize + update: energy and eigenvalues of the guess build, density
    # records [guess, first density], num_iter 1
    rec_E, rec_eps, rec_D = E, [e for e, _ in eig], [D_in, Ds]
    num_iter = 1
    energies, osc_log, fail_log, dens_log = [E], [], [], []
    errs, targets = [], []
    erec = []                       # energy_records
    prev_h = [None]                 # prev_hamiltonian: one empty entry first
    start_diis = ctrl.start_diis_cycle
    converged = False
    while True:
        osc, failed = False, False
        if num_iter < start_diis:
            Fs, E = fock_energy(mix(rec_D[0], rec_D[1]))
        else:
            Fs, E = fock_energy(Ds)
            erec.append(E)
            osc = len(erec) >= 2 and len(erec) > ctrl.start_check_oscillation and erec[-1] - erec[-2] > 0.0
            if len(targets) == ctrl.num_max_diis:
                targets.pop(0)
                errs.pop(0)
            errs.append(np.concatenate([(Fs[s] @ rec_D[1][s] @ S - S @ rec_D[1][s] @ Fs[s]).ravel()
                                        for s in range(ns)]))
            targets.append([F.copy() for F in Fs])
            c = _diis_coefficients(errs)
            if c is not None:
                Fd = [sum(ci * t[s] for ci, t in zip(c, targets)) for s in range(ns)]
                if osc:   # (1 − α)·H[DIIS, i−1] + α·H[DIIS, i+1]; the empty entry leaves F as built
                    Fs = [(1.0 - a) * prev_h[0][s] + a * Fd[s] for s in range(ns)] if prev_h[0] is not None else Fs
                else:
                    Fs = Fd
                if len(prev_h) == 2:
                    prev_h.pop(0)
                prev_h.append([F.copy() for F in Fs])
            else:
                failed = True
                Fs, E = fock_energy(mix(rec_D[0], rec_D[1]))
                start_diis = num_iter + 8
                errs, targets = [], []
        energies.append(E)
        osc_log.append(osc)
        fail_log.append(failed)
        eig = [diag(F, S) for F in Fs]
        eps = [e for e, _ in eig]
        Ds = densities(eig)
        dens_log.append([d.copy() for d in Ds])
        # check_scf_convergence (:1875-1951): nothing before num_iter 2
        dE = E - rec_E
        dm_err = max(np.sqrt(np.sum((Ds[s] - rec_D[1][s]) ** 2)) / n for s in range(ns))
        eev_err = np.sqrt(sum(np.sum((eps[s] - rec_eps[s]) ** 2) for s in range(ns)))
        converged = (num_iter >= 2 and abs(dE) <= ctrl.scf_acc_etot and dm_err <= ctrl.scf_acc_rho
                     and eev_err <= ctrl.scf_acc_eev)
        at_limit = num_iter >= 2 and num_iter >= ctrl.max_scf_cycle
        rec_E, rec_eps, rec_D = E, eps, [rec_D[1], Ds]
        num_iter += 1
        if converged or at_limit:
            break
    return dict(E=E, converged=converged, iterations=num_iter - 1, energies=energies, eig=eig, Ds=Ds, Fs=Fs,
                osc=osc_log, failed=fail_log, densities=dens_log)


def rhf(S, H, rimatr, nelec: int, enuc: float, ctrl: ScfCtrl = ScfCtrl(), d_guess=None, jk=None,
        diag=default_diag) -> ScfResult:
    """Closed-shell RI-HF with packed B (rimatr [p, P]); D = 2 C_occ C_occᵀ.
    `jk(D) -> (J, K)` defaults to the C oracle."""
    nocc = nelec // 2
    jk = jk or oracle_jk(rimatr)

    def fock_energy(Ds):
        J, K = jk(Ds[0])
        F = H + J - 0.5 * K          # s_K = -0.5 for RHF (src/scf_io/mod.rs:1347-1350)
        return [F], enuc + 0.5 * np.sum(Ds[0] * (H + F))

    def densities(eig):
        C = eig[0][1]
        return [2.0 * C[:, :nocc] @ C[:, :nocc].T]

    if d_guess is None:   # core-Hamiltonian guess
        d_guess = densities([diag(H, S)])[0]
    r = _scf(S, fock_energy, densities, [d_guess], ctrl, diag)
    return ScfResult(energy=r["E"], converged=r["converged"], iterations=r["iterations"], energies=r["energies"],
                     eigenvalues=r["eig"][0][0], eigenvectors=r["eig"][0][1], density=r["Ds"][0], fock=r["Fs"][0],
                     osc=r["osc"], diis_failed=r["failed"], densities=[d[0] for d in r["densities"]])


def uhf(S, H, rimatr, nalpha: int, nbeta: int, enuc: float, ctrl: ScfCtrl = ScfCtrl(), d_guess=None,
        jk=None, diag=default_diag) -> ScfResult:
    """Unrestricted RI-HF (REST's UHF branch): occupation 1 per spin,
    F_s = H + J[D_α + D_β] − K[D_s] (s_K = −1, src/scf_io/mod.rs:1347-1381),
    E = E_nuc + ½ Σ_s Σ D_s (H + F_s) (:1673-1713), DIIS error vectors of the two
    spins concatenated (:3094-3136), convergence on both densities and on the
    joint eigenvalue change (:1875-1951).  Densities and Fock matrices are
    returned as [2, n, n]."""
    n = S.shape[0]
    nocc = (nalpha, nbeta)
    jk = jk or oracle_jk(rimatr)

    def fock_energy(Ds):
        J, _ = jk(Ds[0] + Ds[1])
        Fs, E = [], enuc
        for s in range(2):
            K = jk(Ds[s])[1] if nocc[s] else np.zeros((n, n))
            Fs.append(H + J - K)
            E += 0.5 * np.sum(Ds[s] * (H + Fs[s]))
        return Fs, E

    def densities(eig):
        return [eig[s][1][:, :nocc[s]] @ eig[s][1][:, :nocc[s]].T for s in range(2)]

    if d_guess is None:   # core-Hamiltonian guess
        e0 = diag(H, S)
        d_guess = densities([e0, e0])
    r = _scf(S, fock_energy, densities, list(d_guess), ctrl, diag)
    return ScfResult(energy=r["E"], converged=r["converged"], iterations=r["iterations"], energies=r["energies"],
                     eigenvalues=np.array([e for e, _ in r["eig"]]), eigenvectors=np.array([c for _, c in r["eig"]]),
                     density=np.array(r["Ds"]), fock=np.array(r["Fs"]), osc=r["osc"], diis_failed=r["failed"],
                     densities=r["densities"])


def rimatr_from_integrals(X, Vmetric, threshold: float = 1e-8):
    """B = X · V^{-1/2} (lapack_power(-0.5, threshold): eigenvalues <= threshold dropped)."""
    lam, U = np.linalg.eigh(Vmetric)
    keep = lam > threshold
    W = U[:, keep] * lam[keep] ** -0.25
    return X @ (W @ W.T), int(keep.sum())

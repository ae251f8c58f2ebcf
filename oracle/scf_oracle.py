"""CPU restatement of REST's RI-HF SCF loop (test infrastructure only).

Follows src/scf_io/mod.rs:
  - scf()                      :3190-3268  (guess → F, E → diagonalise → loop), RHF and UHF
  - prepare_next_input         :2947-3091  (linear mixing before start_diis_cycle, Pulay DIIS after)
  - generate_diis_error_vector :3094-3136  (e = F D S − S D F, spins concatenated)
  - diis_solver                :3138-3188  (c = B⁻¹1 / 1ᵀB⁻¹1 over at most num_max_diis vectors)
  - check_scf_convergence      :1875-1951  (|ΔE|, ‖ΔD‖_F/nbf, ‖Δε‖₂)
  - generate_hf_hamiltonian_ri_v :1339-1381, energy :1715-1802, density :441-478
J and K come from the C oracle (oracle/rjk_oracle.c); linear algebra from numpy/scipy.
The energy-increase damping branch (start_check_oscillation) is not restated:
it changes the path, not the converged energy.
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


def rhf(S, H, rimatr, nelec: int, enuc: float, ctrl: ScfCtrl = ScfCtrl(), d_guess=None, jk=None) -> ScfResult:
    """Closed-shell RI-HF with packed B (rimatr [p, P]); D = 2 C_occ C_occᵀ.
    `jk(D) -> (J, K)` defaults to the C oracle."""
    n = S.shape[0]
    nocc = nelec // 2
    occ = np.zeros(n)
    occ[:nocc] = 2.0
    jk = jk or oracle_jk(rimatr)

    def fock_energy(D):
        J, K = jk(D)
        F = H + J - 0.5 * K          # s_K = -0.5 for RHF (src/scf_io/mod.rs:1347-1350)
        E = enuc + 0.5 * np.sum(D * (H + F))
        return F, E

    def diag(F):
        e, C = sla.eigh(F, S)
        return e, C

    def density(C):
        return (C[:, :nocc] * occ[:nocc]) @ C[:, :nocc].T

    if d_guess is None:   # core-Hamiltonian guess
        _, C0 = diag(H)
        d_guess = density(C0)
    D_in = d_guess
    F, E = fock_energy(D_in)
    rec_E = E
    eps, C = diag(F)
    rec_eps = eps
    D = density(C)
    rec_D = [D_in, D]
    num_iter = 1
    energies = [E]
    errs, targets = [], []
    start_diis = ctrl.start_diis_cycle
    converged = False
    while True:
        if num_iter < start_diis:
            D_mix = rec_D[0] + ctrl.mix_param * (rec_D[1] - rec_D[0])
            F, E = fock_energy(D_mix)
        else:
            F, E = fock_energy(D)
            if len(targets) == ctrl.num_max_diis:
                targets.pop(0)
                errs.pop(0)
            e = F @ rec_D[1] @ S - S @ rec_D[1] @ F
            errs.append(e.ravel())
            targets.append(F.copy())
            Bm = np.array([[np.dot(a, b) for b in errs] for a in errs])
            try:
                inv = np.linalg.inv(Bm)
                if not np.all(np.isfinite(inv)):
                    raise np.linalg.LinAlgError
                c = inv.sum(axis=1) / inv.sum()
                F = sum(ci * Fi for ci, Fi in zip(c, targets))
            except np.linalg.LinAlgError:
                D_mix = rec_D[0] + ctrl.mix_param * (rec_D[1] - rec_D[0])
                F, E = fock_energy(D_mix)
                start_diis = num_iter + 8
                errs, targets = [], []
        energies.append(E)
        eps, C = diag(F)
        D = density(C)
        dE = E - rec_E
        dm_err = np.sqrt(np.sum((D - rec_D[1]) ** 2)) / n
        eev_err = np.sqrt(np.sum((eps - rec_eps) ** 2))
        converged = abs(dE) <= ctrl.scf_acc_etot and dm_err <= ctrl.scf_acc_rho and eev_err <= ctrl.scf_acc_eev
        rec_E, rec_eps, rec_D = E, eps, [rec_D[1], D]
        num_iter += 1
        if converged or num_iter >= ctrl.max_scf_cycle:
            break
    return ScfResult(energy=E, converged=converged, iterations=num_iter - 1, energies=energies, eigenvalues=eps,
                     eigenvectors=C, density=D, fock=F)


def uhf(S, H, rimatr, nalpha: int, nbeta: int, enuc: float, ctrl: ScfCtrl = ScfCtrl(), d_guess=None,
        jk=None) -> ScfResult:
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

    def diag(F):
        return sla.eigh(F, S)

    def density(C, s):
        return C[:, :nocc[s]] @ C[:, :nocc[s]].T

    if d_guess is None:   # core-Hamiltonian guess
        _, C0 = diag(H)
        d_guess = [density(C0, 0), density(C0, 1)]
    D_in = [np.array(d) for d in d_guess]
    Fs, E = fock_energy(D_in)
    rec_E = E
    eig = [diag(F) for F in Fs]
    rec_eps = [e for e, _ in eig]
    Ds = [density(eig[s][1], s) for s in range(2)]
    rec_D = [D_in, Ds]
    num_iter, energies = 1, [E]
    errs, targets = [], []
    start_diis = ctrl.start_diis_cycle
    converged = False
    mix = lambda: [rec_D[0][s] + ctrl.mix_param * (rec_D[1][s] - rec_D[0][s]) for s in range(2)]
    while True:
        if num_iter < start_diis:
            Fs, E = fock_energy(mix())
        else:
            Fs, E = fock_energy(Ds)
            if len(targets) == ctrl.num_max_diis:
                targets.pop(0)
                errs.pop(0)
            e = np.concatenate([(Fs[s] @ rec_D[1][s] @ S - S @ rec_D[1][s] @ Fs[s]).ravel() for s in range(2)])
            errs.append(e)
            targets.append([F.copy() for F in Fs])
            Bm = np.array([[np.dot(a, b) for b in errs] for a in errs])
            try:
                inv = np.linalg.inv(Bm)
                if not np.all(np.isfinite(inv)):
                    raise np.linalg.LinAlgError
                c = inv.sum(axis=1) / inv.sum()
                Fs = [sum(ci * t[s] for ci, t in zip(c, targets)) for s in range(2)]
            except np.linalg.LinAlgError:
                Fs, E = fock_energy(mix())
                start_diis = num_iter + 8
                errs, targets = [], []
        energies.append(E)
        eig = [diag(F) for F in Fs]
        eps = [e for e, _ in eig]
        Ds = [density(eig[s][1], s) for s in range(2)]
        dE = E - rec_E
        dm_err = max(np.sqrt(np.sum((Ds[s] - rec_D[1][s]) ** 2)) / n for s in range(2))
        eev_err = np.sqrt(sum(np.sum((eps[s] - rec_eps[s]) ** 2) for s in range(2)))
        converged = abs(dE) <= ctrl.scf_acc_etot and dm_err <= ctrl.scf_acc_rho and eev_err <= ctrl.scf_acc_eev
        rec_E, rec_eps, rec_D = E, eps, [rec_D[1], Ds]
        num_iter += 1
        if converged or num_iter >= ctrl.max_scf_cycle:
            break
    return ScfResult(energy=E, converged=converged, iterations=num_iter - 1, energies=energies,
                     eigenvalues=np.array(eps), eigenvectors=np.array([c for _, c in eig]), density=np.array(Ds),
                     fock=np.array(Fs))


def rimatr_from_integrals(X, Vmetric, threshold: float = 1e-8):
    """B = X · V^{-1/2} (lapack_power(-0.5, threshold): eigenvalues <= threshold dropped)."""
    lam, U = np.linalg.eigh(Vmetric)
    keep = lam > threshold
    W = U[:, keep] * lam[keep] ** -0.25
    return X @ (W @ W.T), int(keep.sum())

"""Closed-shell RI-HF of many molecules at once, with every J/K of an iteration
taken from one batched, grouped launch (include/rest_batch.h).

The loop per molecule is REST's scf() (src/scf_io/mod.rs:3190-3268) in the
form the device driver rscf.hip runs it, on the host in FP64 numpy:
  - guess: REST's guess step from the SAD density on the single-problem path
    (rscf_guess_step_factored), whose orbitals start the loop;
  - F = H + J − ½K, E = E_nuc + ½ Σ D (H + F) (:1339-1381, :1715-1802);
  - Pulay DIIS from start_diis_cycle over at most num_max_diis vectors,
    e = F D S − S D F (:3094-3188), with the energy-rise mixing of prev_hamiltonian
    after start_check_oscillation energy records (:3002-3010, :3039-3057);
  - convergence (:1875-1951): from num_iter 2 on, |ΔE| <= scf_acc_etot,
    ‖ΔD‖_F / nbf <= scf_acc_rho, ‖Δε‖₂ <= scf_acc_eev; stop at
    num_iter >= max(max_scf_cycle, 2).
Each molecule's loop is a generator that asks for the J/K of its occupied
orbitals; the driver sets every pending request into the batch, replays the
batch's graph once, and resumes every loop (host work on a thread pool).
Molecules that have stopped are switched off in the batch.  A density that is
not 2 C_occ C_occᵀ (linear mixing before start_diis_cycle, the DIIS-failure
fallback of :3061-3078) needs more K columns than the batch holds: such a
molecule leaves the batch and is finished by the single-problem device SCF.
Molecules go through in waves whose device memory fits a budget
(rjk_batch_plan_waves).  The learned guess's training densities come from here.
"""
from __future__ import annotations

import copy
import threading
import time
from concurrent.futures import ThreadPoolExecutor
from dataclasses import dataclass, field
from typing import List, Optional

import numpy as np

from . import guess as gs
from . import jk
from . import scf as rs


@dataclass
class BatchOutcome:
    energy: float
    converged: bool
    iterations: int
    dm: np.ndarray
    mo_coeff: np.ndarray
    mo_energy: np.ndarray
    batched: bool = True            # False: finished by the single-problem SCF (mixing needed)
    energies: list = field(default_factory=list)


class _NeedsMixing(Exception):
    pass


def _unpack(v: np.ndarray, n: int) -> np.ndarray:
    """Full symmetric matrix of a packed-upper vector (p = j(j+1)/2 + i)."""
    m = np.zeros((n, n))
    iu = np.triu_indices(n)
    order = np.lexsort((iu[0], iu[1]))
    m[iu[0][order], iu[1][order]] = v
    m[iu[1][order], iu[0][order]] = v
    return m


def _diis(errs: list) -> Optional[np.ndarray]:
    """c = B⁻¹1 / 1ᵀB⁻¹1 (diis_solver, :3138-3188); None when B is singular."""
    bm = np.array([[float(np.dot(a, b)) for b in errs] for a in errs])
    try:
        inv = np.linalg.inv(bm)
    except np.linalg.LinAlgError:
        return None
    if not np.all(np.isfinite(inv)):
        return None
    return inv.sum(axis=1) / inv.sum()


def _rhf_loop(S, H, enuc: float, nocc: int, c_guess: np.ndarray, ctrl):
    """REST's closed-shell loop as a generator: yields C_occ, receives (J, K) of D = 2 C_occ C_occᵀ
    (unscaled); returns (E, converged, iterations, D, C, ε, energies)."""
    n = S.shape[0]
    s_val, s_vec = np.linalg.eigh(S)
    X = s_vec / np.sqrt(s_val)[None, :]          # S^{-1/2} (canonical basis, no functions dropped)

    def diag(F):
        e, v = np.linalg.eigh(X.T @ F @ X)
        return e, X @ v

    def density(C):
        co = C[:, :nocc]
        return 2.0 * co @ co.T, co

    D_in, co = density(c_guess)
    J, K = yield co
    F = H + J - 0.5 * K
    E = enuc + 0.5 * float(np.sum(D_in * (H + F)))
    eps, C = diag(F)
    D, co = density(C)
    rec_E, rec_eps, rec_D = E, eps, [D_in, D]
    a = ctrl.mix_param
    num_iter, start_diis = 1, ctrl.start_diis_cycle
    errs, targets, erec, prev_h = [], [], [], [None]
    energies = [E]
    max_cycle = max(ctrl.max_scf_cycle, 2)
    converged = False
    while True:
        if num_iter < start_diis:
            raise _NeedsMixing()
        J, K = yield co
        F = H + J - 0.5 * K
        E = enuc + 0.5 * float(np.sum(D * (H + F)))
        erec.append(E)
        osc = len(erec) >= 2 and len(erec) > ctrl.start_check_oscillation and erec[-1] - erec[-2] > 0.0
        if len(targets) == ctrl.num_max_diis:
            targets.pop(0)
            errs.pop(0)
        fds = F @ D @ S
        errs.append((fds - fds.T).ravel())
        targets.append(F)
        c = _diis(errs)
        if c is None:
            raise _NeedsMixing()
        Fd = sum(ci * t for ci, t in zip(c, targets))
        if osc:
            F = (1.0 - a) * prev_h[0] + a * Fd if prev_h[0] is not None else F
        else:
            F = Fd
        if len(prev_h) == 2:
            prev_h.pop(0)
        prev_h.append(F)
        energies.append(E)
        eps, C = diag(F)
        D_new, co = density(C)
        dE = E - rec_E
        dm_err = float(np.sqrt(np.sum((D_new - rec_D[1]) ** 2))) / n
        eev_err = float(np.sqrt(np.sum((eps - rec_eps) ** 2)))
        converged = (num_iter >= 2 and abs(dE) <= ctrl.scf_acc_etot and dm_err <= ctrl.scf_acc_rho
                     and eev_err <= ctrl.scf_acc_eev)
        at_limit = num_iter >= max_cycle
        rec_E, rec_eps, rec_D = E, eps, [rec_D[1], D_new]
        D = D_new
        num_iter += 1
        if converged or at_limit:
            return E, converged, num_iter - 1, D, C, eps, energies


def _single_bytes(nbf: int, naux: int) -> int:
    """Device bytes held by one molecule's set-up (its FP64 B and the guess step's work space, bounded
    by twice B)."""
    ldb = -(-(nbf * (nbf + 1) // 2) // 64) * 64
    return 2 * 8 * ldb * naux


def rhf_batch(mols, ctrl=None, prec: int = jk.PREC_FP64, budget_bytes: int = 0, threads: int = 16,
              setup_threads: int = 8, log=None, stats: dict = None) -> List[BatchOutcome]:
    """Converge every molecule of `mols` (ri.Molecule, closed shell); returns their outcomes in order.
    prec: the batch's J/K precision; budget_bytes: device memory per wave (0: 90% of the free memory
    less what `setup_threads` concurrent set-ups hold); threads: host threads for the per-molecule
    loops; setup_threads: molecules whose integrals, B and SAD guess step are made at once."""
    from threadpoolctl import threadpool_limits

    ctrl = ctrl or rs.default_ctrl()
    nbf = [m.nbf for m in mols]
    naux = [m.naux for m in mols]
    nocc = [m.nelectron // 2 for m in mols]
    if budget_bytes == 0:
        free, _ = jk.device_memory()
        reserve = setup_threads * max(_single_bytes(a, b) for a, b in zip(nbf, naux)) + (2 << 30)
        budget_bytes = max(1 << 30, int(0.9 * free) - reserve)
    waves = jk.batch_plan_waves(nbf, naux, nocc, prec, budget_bytes)
    for el in sorted({e for m in mols for e in m.symbols}):    # atomic densities cached before threads start
        gs.atomic_density(el, mols[0].basis, mols[0].auxbasis)
    out: List[Optional[BatchOutcome]] = [None] * len(mols)
    st = {"waves": len(waves), "batch_runs": 0, "jk_seconds": 0.0, "host_seconds": 0.0, "setup_seconds": 0.0,
          "fallbacks": 0, "alloc_seconds": 0.0, "setup_int1e_thread_s": 0.0, "setup_b_thread_s": 0.0,
          "setup_guess_thread_s": 0.0}
    tlock = threading.Lock()

    def tadd(key, v):
        with tlock:
            st[key] += v

    def run_wave(batch, i0, i1, pool):
        t0 = time.perf_counter()
        loops, pend = {}, {}

        def setup(k):
            i = i0 + k
            m = mols[i]
            ta = time.perf_counter()
            S, T, V = m.int1e()
            H = T + V
            tb = time.perf_counter()
            with jk.DeviceRI(m.nbf, m.naux) as r:
                m.build_rimatr(r, 1e-8, jk.STORE_FP64)
                tc = time.perf_counter()
                c1 = gs.sad_guess(m, r, S, H).c1
                batch.set_from_ri(k, r)
            td = time.perf_counter()
            tadd("setup_int1e_thread_s", tb - ta)
            tadd("setup_b_thread_s", tc - tb)
            tadd("setup_guess_thread_s", td - tc)
            c1 = c1[0] if c1.ndim == 3 else c1
            loop = _rhf_loop(S, H, m.nuclear_repulsion(), nocc[i], c1, ctrl)
            return k, loop, next(loop)

        with ThreadPoolExecutor(max_workers=max(1, setup_threads)) as ex, \
                threadpool_limits(limits=max(1, threads // max(1, setup_threads))):
            for k, loop, req in ex.map(setup, range(i1 - i0)):
                loops[k] = loop
                pend[k] = req
        st["setup_seconds"] += time.perf_counter() - t0
        occs = {k: np.full(nocc[i0 + k], 2.0) for k in loops}

        def resume(k, jkm):
            try:
                return k, loops[k].send(jkm), None
            except StopIteration as e:
                return k, None, e.value
            except _NeedsMixing:
                return k, None, "mixing"

        while pend:
            t1 = time.perf_counter()
            for k, co in pend.items():
                batch.set_inputs(k, 2.0 * co @ co.T, co, occs[k])
            batch.run(1.0, 1.0)
            batch.synchronize()
            res = {k: batch.get_outputs(k) for k in pend}
            t2 = time.perf_counter()
            st["batch_runs"] += 1
            st["jk_seconds"] += t2 - t1
            with threadpool_limits(limits=1):
                done = list(pool.map(lambda k: resume(k, (_unpack(res[k][0], nbf[i0 + k]),
                                                          _unpack(res[k][1], nbf[i0 + k]))), list(pend)))
            st["host_seconds"] += time.perf_counter() - t2
            new = {}
            for k, req, fin in done:
                if req is not None:
                    new[k] = req
                    continue
                batch.set_active(k, False)
                i = i0 + k
                if fin == "mixing":
                    st["fallbacks"] += 1
                    out[i] = _single(mols[i], ctrl)
                else:
                    E, conv, it, D, C, eps, energies = fin
                    out[i] = BatchOutcome(E, bool(conv), it, D, C, eps, True, energies)
            pend = new

    with ThreadPoolExecutor(max_workers=max(1, threads)) as pool:
        for w, (i0, i1) in enumerate(waves):
            t0 = time.perf_counter()
            batch = jk.Batch(nbf[i0:i1], naux[i0:i1], nocc[i0:i1], prec)
            st["alloc_seconds"] += time.perf_counter() - t0
            try:
                run_wave(batch, i0, i1, pool)
            finally:
                t3 = time.perf_counter()
                batch.close()
                st["alloc_seconds"] += time.perf_counter() - t3
            if log:
                log(f"wave {w + 1}/{len(waves)}: molecules {i0}..{i1 - 1} done, {st['batch_runs']} batch runs so far")
    if stats is not None:
        stats.update(st)
    return out


def _single(mol, ctrl) -> BatchOutcome:
    """The single-problem device SCF of one molecule (the fallback for a mixing step)."""
    o, _ = rs.run_molecule(mol, guess="sad", ctrl=copy.copy(ctrl))
    return BatchOutcome(o.energy, bool(o.converged), o.iterations, o.dm, o.mo_coeff, o.mo_energy, False, [])

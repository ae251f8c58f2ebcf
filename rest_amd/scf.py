"""Python binding of the device-resident RI-HF SCF driver (include/rest_scf.h).

`scf(mol, ...)` mirrors REST's `scf()` (src/scf_io/mod.rs:3190-3268) for
closed-shell RI-HF: B is built into a DeviceRI, the SCF loop runs in HBM.
"""
from __future__ import annotations

import ctypes as C
import time
from dataclasses import dataclass, field
from typing import Optional

import numpy as np

from . import jk

LOG_FIELDS = ("energy", "dE", "diis_max", "dm_err", "eev_err", "prec", "jk_ms", "diag_ms", "iter_ms", "dref",
              "inc_err", "osc", "exact")


class Ctrl(C.Structure):
    _fields_ = [("num_max_diis", C.c_int), ("start_diis_cycle", C.c_int), ("mix_param", C.c_double),
                ("max_scf_cycle", C.c_int), ("scf_acc_rho", C.c_double), ("scf_acc_eev", C.c_double),
                ("scf_acc_etot", C.c_double), ("jk_prec", C.c_int), ("switch_diis_err", C.c_double),
                ("print_level", C.c_int), ("switch_mode", C.c_int), ("jk_prec_start", C.c_int),
                ("start_switch_err", C.c_double), ("guess_energy_max", C.c_double), ("pseudo_diag_err", C.c_double),
                ("jk_prec_increment", C.c_int), ("guess_err_max", C.c_double),
                ("start_check_oscillation", C.c_int), ("inc_refresh", C.c_double), ("inc_check", C.c_int)]


SWITCH_RESTART, SWITCH_SHIFT = 0, 1


class Result(C.Structure):
    _fields_ = [("energy", C.c_double), ("converged", C.c_int), ("iterations", C.c_int),
                ("fp64_iterations", C.c_int), ("wall_seconds", C.c_double), ("jk_seconds", C.c_double),
                ("diag_seconds", C.c_double), ("diis_seconds", C.c_double), ("guess_energy", C.c_double),
                ("guess_rejected", C.c_int), ("incremental_iterations", C.c_int),
                ("guess_error", C.c_double), ("reference_builds", C.c_int)]


def _bind(lib):
    if getattr(lib, "_rscf_bound", False):
        return lib
    vp = C.c_void_p
    lib.rscf_default_ctrl.restype = None
    lib.rscf_default_ctrl.argtypes = [C.POINTER(Ctrl)]
    lib.rscf_rhf.restype = C.c_int
    lib.rscf_rhf.argtypes = [vp, vp, vp, C.c_double, C.c_size_t, vp, C.POINTER(Ctrl), C.POINTER(Result), vp, vp,
                             vp, vp]
    lib.rscf_rhf_orbitals.restype = C.c_int
    lib.rscf_rhf_orbitals.argtypes = [vp, vp, vp, C.c_double, C.c_size_t, vp, C.POINTER(Ctrl), C.POINTER(Result),
                                      vp, vp, vp, vp]
    lib.rscf_uhf.restype = C.c_int
    lib.rscf_uhf.argtypes = [vp, vp, vp, C.c_double, C.c_size_t, C.c_size_t, vp, C.POINTER(Ctrl),
                             C.POINTER(Result), vp, vp, vp, vp]
    lib.rscf_guess_step.restype = C.c_int
    lib.rscf_guess_step.argtypes = [vp, vp, vp, C.c_double, C.c_size_t, vp, vp, C.c_int, C.POINTER(C.c_double), vp,
                                    vp, vp, C.POINTER(C.c_double)]
    lib.rscf_guess_step_factored.restype = C.c_int
    lib.rscf_guess_step_factored.argtypes = [vp, vp, vp, C.c_double, C.c_size_t, vp, vp, vp, vp, vp, C.c_int,
                                             C.POINTER(C.c_double), vp, vp, vp, C.POINTER(C.c_double)]
    lib.rscf_energy.restype = C.c_int
    lib.rscf_energy.argtypes = [vp, vp, vp, C.c_double, C.c_size_t, vp, vp, C.c_int, C.POINTER(C.c_double)]
    lib.rscf_mcweeny.restype = C.c_int
    lib.rscf_mcweeny.argtypes = [vp, vp, vp, C.c_int, vp, C.POINTER(C.c_double)]
    lib._rscf_bound = True
    return lib


def default_ctrl(**kw) -> Ctrl:
    lib = _bind(jk.load_library())
    c = Ctrl()
    lib.rscf_default_ctrl(C.byref(c))
    for k, v in kw.items():
        setattr(c, k, v)
    return c


def mixed_ctrl(**kw) -> Ctrl:
    """The mixed-precision ladder: BF16X2 J/K until max|e| < 1e-4, BF16X3 until < 1e-5, then one
    FP64 reference build and FP64-referenced increments at BF16X2; low-precision iterations below
    max|e| = 1e-2 and every iteration from the reference on pseudo-diagonalise.  (Two planes give
    K within 3e-6 of FP64 and three within 2e-6 at (H2O)64, the FP32 accumulation dominating
    both, so the third plane is kept for the last steps before the FP64 switch only.)"""
    c = default_ctrl(jk_prec=jk.PREC_BF16X3, jk_prec_start=jk.PREC_BF16X2, start_switch_err=1e-4,
                     pseudo_diag_err=1e-2, jk_prec_increment=jk.PREC_BF16X2)
    for k, v in kw.items():
        setattr(c, k, v)
    return c


@dataclass
class ScfOutcome:
    energy: float
    converged: bool
    iterations: int
    fp64_iterations: int
    wall_seconds: float
    jk_seconds: float
    diag_seconds: float
    diis_seconds: float
    mo_energy: np.ndarray = None
    mo_coeff: np.ndarray = None
    dm: np.ndarray = None
    log: list = field(default_factory=list)
    guess_energy: float = None      # energy of the guess Fock build
    guess_rejected: bool = False    # the guess failed ctrl.guess_energy_max or ctrl.guess_err_max: no iteration ran
    incremental_iterations: int = 0  # FP64 iterations built as increments on the reference build
    guess_error: float = None       # max|F D S - S D F| of the guess build
    reference_builds: int = 0       # full FP64 reference builds of the incremental tail


def _colmajor(M):
    return np.ascontiguousarray(np.asarray(M, dtype=np.float64).T)


def _sym(M):
    """A symmetric matrix (overlap, core Hamiltonian, density): its row-major memory is already
    its column-major layout, so no transposing copy is made."""
    return np.ascontiguousarray(M, dtype=np.float64)


def _from_colmajor(a, symmetric: bool):
    """A column-major n x n output as a row-major-indexed array: the buffer itself for a symmetric
    matrix, else its transposed view (no copy either way)."""
    return a if symmetric else a.T


def _run(ri, S, H, enuc, nel, d_guess, ctrl, want_orbitals, c_guess=None) -> ScfOutcome:
    lib = _bind(jk.load_library())
    n = ri.nbf
    ns = len(nel)
    ctrl = ctrl or default_ctrl()
    Sf, Hf = _sym(S), _sym(H)
    dg = None
    if d_guess is not None:
        dgs = [d_guess] if ns == 1 else list(d_guess)
        dg = np.ascontiguousarray(np.stack([_sym(d) for d in dgs]))
    res = Result()
    C_ = np.empty((ns, n, n)) if want_orbitals else None
    eps = np.empty((ns, n))
    dm = np.empty((ns, n, n)) if want_orbitals else None
    log = np.zeros((max(ctrl.max_scf_cycle, 2), len(LOG_FIELDS)))   # REST runs at least two iterations
    ptr = lambda a: None if a is None else a.ctypes.data
    cg = None if c_guess is None else _colmajor(c_guess)
    if ns == 1 and cg is not None:
        rc = lib.rscf_rhf_orbitals(ri.h, ptr(Sf), ptr(Hf), float(enuc), int(nel[0]), ptr(cg), C.byref(ctrl),
                                   C.byref(res), ptr(C_), ptr(eps), ptr(dm), ptr(log))
    elif ns == 1:
        rc = lib.rscf_rhf(ri.h, ptr(Sf), ptr(Hf), float(enuc), int(nel[0]), ptr(dg), C.byref(ctrl), C.byref(res),
                          ptr(C_), ptr(eps), ptr(dm), ptr(log))
    else:
        rc = lib.rscf_uhf(ri.h, ptr(Sf), ptr(Hf), float(enuc), int(nel[0]), int(nel[1]), ptr(dg), C.byref(ctrl),
                          C.byref(res), ptr(C_), ptr(eps), ptr(dm), ptr(log))
    if rc != 0:
        raise jk.RJKError(rc, lib.rjk_last_error(ri.h).decode())
    rows = [dict(zip(LOG_FIELDS, r)) for r in log[:res.iterations]]
    tr = (lambda a, sym: None if a is None else
          (_from_colmajor(a[0], sym) if ns == 1 else (a if sym else a.transpose(0, 2, 1))))
    return ScfOutcome(res.energy, bool(res.converged), res.iterations, res.fp64_iterations, res.wall_seconds,
                      res.jk_seconds, res.diag_seconds, res.diis_seconds, eps[0] if ns == 1 else eps, tr(C_, False), tr(dm, True),
                      rows, res.guess_energy, bool(res.guess_rejected), res.incremental_iterations,
                      res.guess_error, res.reference_builds)


def rhf(ri: "jk.DeviceRI", S: np.ndarray, H: np.ndarray, enuc: float, nelec: int,
        d_guess: Optional[np.ndarray] = None, ctrl: Optional[Ctrl] = None, want_orbitals: bool = True,
        c_guess: Optional[np.ndarray] = None) -> ScfOutcome:
    """Closed-shell RI-HF on ri's resident B: from a guess density (rscf_rhf), guess orbitals
    (c_guess, rscf_rhf_orbitals) or the core Hamiltonian."""
    return _run(ri, S, H, enuc, (nelec,), d_guess, ctrl, want_orbitals, c_guess)


def uhf(ri: "jk.DeviceRI", S: np.ndarray, H: np.ndarray, enuc: float, nalpha: int, nbeta: int,
        d_guess=None, ctrl: Optional[Ctrl] = None, want_orbitals: bool = True) -> ScfOutcome:
    """Unrestricted RI-HF (rscf_uhf); d_guess is (D_α, D_β) or None; orbitals/densities are [2, n, n]."""
    return _run(ri, S, H, enuc, (nalpha, nbeta), d_guess, ctrl, want_orbitals)


def guess_step(ri: "jk.DeviceRI", S, H, enuc: float, nel, d0, prec: int = jk.PREC_FP64, want_e1: bool = False,
               want_outputs: bool = True, factors=None):
    """REST's density-guess step (rscf_guess_step): returns (E(D0), D1, C, ε), plus E(D1) when
    want_e1; nel = (nelec,) or (nα, nβ).  With want_outputs=False D1, C and ε stay on the device
    and come back as None (the energies only).  factors: per spin (C0, w0) with D0 = C0 diag(w0) C0ᵀ
    (rscf_guess_step_factored: the first K by the occupied route over C0's columns)."""
    lib = _bind(jk.load_library())
    n = ri.nbf
    ns = len(nel)
    d0s = [d0] if ns == 1 else list(d0)
    d0f = np.ascontiguousarray(np.stack([_sym(d) for d in d0s]))
    d1 = np.empty((ns, n, n)) if want_outputs else None
    C_ = np.empty((ns, n, n)) if want_outputs else None
    eps = np.empty((ns, n)) if want_outputs else None
    ptr = lambda a: None if a is None else a.ctypes.data
    nl = (C.c_size_t * 2)(*(list(nel) + [0] * (2 - ns)))
    e = C.c_double(0.0)
    e1 = C.c_double(0.0)
    Sf, Hf = _sym(S), _sym(H)      # kept alive across the call
    if factors is None:
        rc = lib.rscf_guess_step(ri.h, Sf.ctypes.data, Hf.ctypes.data, float(enuc), ns, nl, d0f.ctypes.data,
                                 int(prec), C.byref(e), ptr(d1), ptr(C_), ptr(eps), C.byref(e1) if want_e1 else None)
    else:
        fs = [factors] if ns == 1 and len(factors) == 2 and np.ndim(factors[0]) == 2 else list(factors)
        c0 = np.zeros((ns, n, n))
        w0 = np.zeros((ns, n))
        nc = (C.c_size_t * 2)(0, 0)
        for s_, (cs, ws) in enumerate(fs):
            k = cs.shape[1]
            c0[s_, :k, :] = np.asarray(cs, dtype=np.float64).T       # column-major: column j at row j
            w0[s_, :k] = ws
            nc[s_] = k
        rc = lib.rscf_guess_step_factored(ri.h, Sf.ctypes.data, Hf.ctypes.data, float(enuc), ns, nl, d0f.ctypes.data,
                                          c0.ctypes.data, w0.ctypes.data, C.cast(nc, C.c_void_p), int(prec),
                                          C.byref(e), ptr(d1), ptr(C_), ptr(eps), C.byref(e1) if want_e1 else None)
    if rc != 0:
        raise jk.RJKError(rc, lib.rjk_last_error(ri.h).decode())
    if not want_outputs:
        out = (e.value, None, None, None)
        return out + (e1.value,) if want_e1 else out
    C_ = C_.transpose(0, 2, 1)     # column-major outputs; d1 is symmetric
    out = (e.value, (d1[0] if ns == 1 else d1), (C_[0] if ns == 1 else C_), (eps[0] if ns == 1 else eps))
    return out + (e1.value,) if want_e1 else out


def energy(ri: "jk.DeviceRI", S, H, enuc: float, nel, dm, prec: int = jk.PREC_FP64) -> float:
    """HF energy of a density without diagonalising (rscf_energy); nel = (nelec,) or (nα, nβ)."""
    lib = _bind(jk.load_library())
    ns = len(nel)
    dms = [dm] if ns == 1 else list(dm)
    df = np.ascontiguousarray(np.stack([_sym(d) for d in dms]))
    nl = (C.c_size_t * 2)(*(list(nel) + [0] * (2 - ns)))
    Sf, Hf = _sym(S), _sym(H)
    e = C.c_double(0.0)
    rc = lib.rscf_energy(ri.h, Sf.ctypes.data, Hf.ctypes.data, float(enuc), ns, nl, df.ctypes.data, int(prec),
                         C.byref(e))
    if rc != 0:
        raise jk.RJKError(rc, lib.rjk_last_error(ri.h).decode())
    return e.value


def mcweeny(ri: "jk.DeviceRI", S, D, iterations: int = 3):
    """McWeeny-purified closed-shell density (rscf_mcweeny); returns (D, electron count 2 tr(PS))."""
    lib = _bind(jk.load_library())
    out = np.empty_like(np.asarray(D, dtype=np.float64))
    ne = C.c_double(0.0)
    Sf, Df = _sym(S), _sym(D)      # kept alive across the call
    rc = lib.rscf_mcweeny(ri.h, Sf.ctypes.data, Df.ctypes.data, int(iterations), out.ctypes.data, C.byref(ne))
    if rc != 0:
        raise jk.RJKError(rc, lib.rjk_last_error(ri.h).decode())
    return _from_colmajor(out, True), ne.value


def run_molecule(mol, prec: int = jk.PREC_FP64, d_guess=None, ctrl: Optional[Ctrl] = None,
                 threshold: float = 1e-8, keep_b: bool = False, guess: str = "sad", guess_path=None,
                 chkfile=None) -> tuple:
    """Integrals, B build, initial guess and SCF for a rest_amd.ri.Molecule; returns
    (ScfOutcome, timings dict).  guess (REST's initial_guess, src/initial_guess/mod.rs:15-72):
      "sad"   — REST's default (src/ctrl_io/mod.rs:229);
      "hcore" — core Hamiltonian;
      "file"  — the density of a REST guess file (HDF5 dataset "init_guess") at guess_path,
                then the guess step (initial_guess_from_hdf5guess, :75-87);
      "chk"   — the orbitals of a REST chkfile (group "scf") at guess_path (restart, :89-128).
    An explicit d_guess density overrides guess.  chkfile: write REST's chkfile (scf/e_tot,
    scf/mo_coeff, scf/mo_energy; save_chkfile, src/post_scf_analysis/mod.rs:47-103) at the end.
    With keep_b the host copy of the device B is returned as timings['rimatr']."""
    from . import guess as gs
    from . import h5
    t0 = time.perf_counter()
    S, T, V = mol.int1e()
    H = T + V
    t1 = time.perf_counter()
    ri_ = jk.DeviceRI(mol.nbf, mol.naux)
    store = jk.STORE_FP64 | (jk.STORE_FP32 if prec != jk.PREC_FP64 else 0) \
        | (jk.STORE_BF16 if prec in (jk.PREC_BF16X1, jk.PREC_BF16X2, jk.PREC_BF16X3) else 0)
    binfo = mol.build_rimatr(ri_, threshold, store)
    t2 = time.perf_counter()
    c_guess = None
    if d_guess is not None:
        ginfo = {"guess": "density"}
    elif guess == "sad":
        g = gs.sad_guess(mol, ri_, S, H)
        c_guess = g.c1
        ginfo = {"guess": "sad", "e_guess_d0": g.energy_d0}
    elif guess == "file":
        d0 = h5.read_init_guess(guess_path, mol.nbf, 1)[0]
        e0, _, c_guess, _ = guess_step(ri_, S, H, mol.nuclear_repulsion(), (mol.nelectron,), d0)
        ginfo = {"guess": "file", "e_guess_d0": e0}
    elif guess == "chk":
        _, cs, _ = h5.read_chkfile(guess_path, mol.nbf, mol.nbf, 1)
        c_guess = cs[0]
        ginfo = {"guess": "chk"}
    elif guess == "hcore":
        ginfo = {"guess": "hcore"}
    else:
        raise ValueError(f"unknown guess {guess!r}")
    t3 = time.perf_counter()
    ctrl = ctrl or default_ctrl()
    ctrl.jk_prec = prec
    out = rhf(ri_, S, H, mol.nuclear_repulsion(), mol.nelectron, d_guess, ctrl, c_guess=c_guess)
    t4 = time.perf_counter()
    if chkfile is not None:
        h5.write_chkfile(chkfile, out.energy, [out.mo_coeff], [out.mo_energy])
    extra = {"rimatr": ri_.get_rimatr(), "S": S, "H": H} if keep_b else {}
    ri_.close()
    return out, {"t_1e": t1 - t0, "t_B": t2 - t1, "t_guess": t3 - t2, "t_scf": t4 - t3, **ginfo, **binfo, **extra}

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

LOG_FIELDS = ("energy", "dE", "diis_max", "dm_err", "eev_err", "prec", "jk_ms", "diag_ms", "iter_ms")


class Ctrl(C.Structure):
    _fields_ = [("num_max_diis", C.c_int), ("start_diis_cycle", C.c_int), ("mix_param", C.c_double),
                ("max_scf_cycle", C.c_int), ("scf_acc_rho", C.c_double), ("scf_acc_eev", C.c_double),
                ("scf_acc_etot", C.c_double), ("jk_prec", C.c_int), ("switch_diis_err", C.c_double),
                ("print_level", C.c_int)]


class Result(C.Structure):
    _fields_ = [("energy", C.c_double), ("converged", C.c_int), ("iterations", C.c_int),
                ("fp64_iterations", C.c_int), ("wall_seconds", C.c_double), ("jk_seconds", C.c_double),
                ("diag_seconds", C.c_double), ("diis_seconds", C.c_double)]


def _bind(lib):
    if getattr(lib, "_rscf_bound", False):
        return lib
    vp = C.c_void_p
    lib.rscf_default_ctrl.restype = None
    lib.rscf_default_ctrl.argtypes = [C.POINTER(Ctrl)]
    lib.rscf_rhf.restype = C.c_int
    lib.rscf_rhf.argtypes = [vp, vp, vp, C.c_double, C.c_size_t, vp, C.POINTER(Ctrl), C.POINTER(Result), vp, vp,
                             vp, vp]
    lib._rscf_bound = True
    return lib


def default_ctrl(**kw) -> Ctrl:
    lib = _bind(jk.load_library())
    c = Ctrl()
    lib.rscf_default_ctrl(C.byref(c))
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


def rhf(ri: "jk.DeviceRI", S: np.ndarray, H: np.ndarray, enuc: float, nelec: int,
        d_guess: Optional[np.ndarray] = None, ctrl: Optional[Ctrl] = None, want_orbitals: bool = True) -> ScfOutcome:
    lib = _bind(jk.load_library())
    n = ri.nbf
    ctrl = ctrl or default_ctrl()
    Sf = np.ascontiguousarray(np.asarray(S, dtype=np.float64).T)   # column-major
    Hf = np.ascontiguousarray(np.asarray(H, dtype=np.float64).T)
    dg = None if d_guess is None else np.ascontiguousarray(np.asarray(d_guess, dtype=np.float64).T)
    res = Result()
    C_ = np.empty((n, n)) if want_orbitals else None
    eps = np.empty(n)
    dm = np.empty((n, n)) if want_orbitals else None
    log = np.zeros((ctrl.max_scf_cycle, len(LOG_FIELDS)))
    rc = lib.rscf_rhf(ri.h, Sf.ctypes.data, Hf.ctypes.data, float(enuc), int(nelec),
                      None if dg is None else dg.ctypes.data, C.byref(ctrl), C.byref(res),
                      None if C_ is None else C_.ctypes.data, eps.ctypes.data, None if dm is None else dm.ctypes.data,
                      log.ctypes.data)
    if rc != 0:
        raise jk.RJKError(rc, lib.rjk_last_error(ri.h).decode())
    rows = [dict(zip(LOG_FIELDS, r)) for r in log[:res.iterations]]
    return ScfOutcome(res.energy, bool(res.converged), res.iterations, res.fp64_iterations, res.wall_seconds,
                      res.jk_seconds, res.diag_seconds, res.diis_seconds, eps,
                      None if C_ is None else C_.T.copy(), None if dm is None else dm.T.copy(), rows)


def run_molecule(mol, prec: int = jk.PREC_FP64, d_guess=None, ctrl: Optional[Ctrl] = None,
                 threshold: float = 1e-8, keep_b: bool = False) -> tuple:
    """Integrals, B build and SCF for a rest_amd.ri.Molecule; returns (ScfOutcome, timings dict).
    With keep_b the host copy of the device B is returned as timings['rimatr']."""
    t0 = time.perf_counter()
    S, T, V = mol.int1e()
    H = T + V
    t1 = time.perf_counter()
    ri_ = jk.DeviceRI(mol.nbf, mol.naux)
    store = jk.STORE_FP64 | (jk.STORE_FP32 if prec == jk.PREC_FP32 else 0)
    binfo = mol.build_rimatr(ri_, threshold, store)
    t2 = time.perf_counter()
    ctrl = ctrl or default_ctrl()
    ctrl.jk_prec = prec
    out = rhf(ri_, S, H, mol.nuclear_repulsion(), mol.nelectron, d_guess, ctrl)
    t3 = time.perf_counter()
    extra = {"rimatr": ri_.get_rimatr(), "S": S, "H": H} if keep_b else {}
    ri_.close()
    return out, {"t_1e": t1 - t0, "t_B": t2 - t1, "t_scf": t3 - t2, **binfo, **extra}

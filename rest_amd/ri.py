"""Molecules, basis sets and the RI-V B producer (include/rest_ri.h).

Mirrors what REST's Molecule builds before the SCF (src/molecule_io/mod.rs):
basis shells per atom sorted by angular momentum (src/basis_io/mod.rs:133),
coordinates in Å converted with ANG = 0.5291772083 (src/constants/mod.rs:85),
nuclear repulsion (src/geom_io/mod.rs:163-179), the one-electron matrices,
the auxiliary metric and packed B = (μν|Q)·V^{-1/2} written into a DeviceRI.
"""
from __future__ import annotations

import ctypes as C
import json
from dataclasses import dataclass
from functools import lru_cache
from pathlib import Path
from typing import Optional, Sequence

import numpy as np

from . import jk

ANG = 0.5291772083          # bohr radius in Å, as REST's constants
BASIS_FILE = Path(__file__).resolve().parent / "basis" / "basis_sets.json"
ZNUC = {"H": 1, "He": 2, "Li": 3, "Be": 4, "B": 5, "C": 6, "N": 7, "O": 8, "F": 9, "Ne": 10}


@lru_cache(maxsize=1)
def _basis_db() -> dict:
    return json.loads(BASIS_FILE.read_text())["basis"]


def _check(rc: int, what: str):
    if rc != 0:
        lib = jk.load_library()
        raise jk.RJKError(rc, f"{what}: {lib.rsi_last_error().decode()}")


class Basis:
    """Shells of one basis set on a molecule (handle of the C-ABI rsi_basis)."""

    def __init__(self, symbols: Sequence[str], xyz_bohr: np.ndarray, name: str, shells: Optional[list] = None):
        """`shells` (optional) overrides the named set: a list of (atom, l, exponents, coefficients)."""
        atom, lval, nprim, exps, coefs = [], [], [], [], []
        if shells is None:
            db = _basis_db()
            if name not in db:
                raise KeyError(f"basis {name!r} not available (have {sorted(db)})")
            shells = []
            for a, el in enumerate(symbols):
                for s in sorted(db[name][el], key=lambda s: s["l"]):   # stable: REST's per-element order
                    shells.append((a, s["l"], s["exponents"], s["coefficients"]))
        for a, l, ex, co in shells:
            atom.append(a)
            lval.append(l)
            nprim.append(len(ex))
            exps.extend(ex)
            coefs.extend(co)
        self.name = name
        self.shell_atom = np.array(atom, dtype=np.int32)
        self.shell_l = np.array(lval, dtype=np.int32)
        self.shell_nprim = np.array(nprim, dtype=np.int32)
        self._exps = np.array(exps, dtype=np.float64)
        self._coefs = np.array(coefs, dtype=np.float64)
        self._xyz = np.ascontiguousarray(xyz_bohr, dtype=np.float64)
        lib = jk.load_library()
        h = C.c_void_p()
        _check(lib.rsi_basis_create(len(atom), self.shell_atom.ctypes.data, self.shell_l.ctypes.data,
                                    self.shell_nprim.ctypes.data, self._exps.ctypes.data, self._coefs.ctypes.data,
                                    len(symbols), self._xyz.ctypes.data, C.byref(h)), "rsi_basis_create")
        self.h = h
        self.nbf = int(lib.rsi_basis_nbf(h))
        offs, o = [], 0
        for lv in lval:
            offs.append(o)
            o += 2 * lv + 1
        self.shell_off = np.array(offs, dtype=np.int64)

    def __del__(self):
        try:
            if getattr(self, "h", None):
                jk.load_library().rsi_basis_destroy(self.h)
                self.h = None
        except Exception:
            pass

    def shell_boundaries(self) -> np.ndarray:
        return np.append(self.shell_off, self.nbf)


def int1e(orb: Basis, z: np.ndarray = None, xyz_bohr: np.ndarray = None):
    """(S, T, V) of a basis; V over the point charges z at xyz_bohr (None: no V)."""
    n = orb.nbf
    S = np.empty((n, n)); T = np.empty((n, n))
    V = np.empty((n, n)) if z is not None else None
    zz = None if z is None else np.ascontiguousarray(z, dtype=np.float64)
    xx = None if z is None else np.ascontiguousarray(xyz_bohr, dtype=np.float64)
    _check(jk.load_library().rsi_int1e(orb.h, 0 if zz is None else len(zz), None if zz is None else zz.ctypes.data,
                                       None if xx is None else xx.ctypes.data, S.ctypes.data, T.ctypes.data,
                                       None if V is None else V.ctypes.data), "rsi_int1e")
    return S, T, V


def int2c2e(aux: Basis) -> np.ndarray:
    V = np.empty((aux.nbf, aux.nbf))
    _check(jk.load_library().rsi_int2c2e(aux.h, V.ctypes.data), "rsi_int2c2e")
    return V


def int3c2e_packed(orb: Basis, aux: Basis, j0: int = 0, j1: Optional[int] = None) -> np.ndarray:
    """(μν|P) for pair columns j in [j0, j1): array [p, P] with p relative to j0(j0+1)/2."""
    j1 = orb.nbf if j1 is None else j1
    nq = j1 * (j1 + 1) // 2 - j0 * (j0 + 1) // 2
    out = np.empty((aux.nbf, nq))
    _check(jk.load_library().rsi_int3c2e_packed(orb.h, aux.h, j0, j1, out.ctypes.data), "rsi_int3c2e_packed")
    return out.T


def int3c2e_packed_dev(ri: "jk.DeviceRI", orb: Basis, aux: Basis, j0: int = 0, j1: Optional[int] = None):
    """Device kernels' (μν|P) for pair columns [j0, j1) as a torch tensor on ri's device, shape [P, q]."""
    import torch
    j1 = orb.nbf if j1 is None else j1
    nq = j1 * (j1 + 1) // 2 - j0 * (j0 + 1) // 2
    out = torch.empty((aux.nbf, nq), dtype=torch.float64, device="cuda")
    rc = jk.load_library().rsi_int3c2e_packed_dev(ri.h, orb.h, aux.h, j0, j1, out.data_ptr())
    _check(rc, "rsi_int3c2e_packed_dev")
    return out


@dataclass
class Molecule:
    symbols: list
    xyz_bohr: np.ndarray
    basis: str = "def2-SVP"
    auxbasis: str = "def2-SVP-JKFIT"
    charge: int = 0

    @classmethod
    def from_angstrom(cls, symbols, xyz_ang, basis="def2-SVP", auxbasis="def2-SVP-JKFIT", charge=0):
        return cls(list(symbols), np.asarray(xyz_ang, dtype=np.float64) / ANG, basis, auxbasis, charge)

    @classmethod
    def water_cluster(cls, xyz_ang: np.ndarray, basis="def2-SVP", auxbasis="def2-SVP-JKFIT"):
        n = len(xyz_ang) // 3
        return cls.from_angstrom(["O", "H", "H"] * n, xyz_ang, basis, auxbasis)

    def __post_init__(self):
        self.xyz_bohr = np.ascontiguousarray(self.xyz_bohr, dtype=np.float64)
        self.orb = Basis(self.symbols, self.xyz_bohr, self.basis)
        self.aux = Basis(self.symbols, self.xyz_bohr, self.auxbasis)
        self.z = np.array([ZNUC[s] for s in self.symbols], dtype=np.float64)

    @property
    def nbf(self) -> int:
        return self.orb.nbf

    @property
    def naux(self) -> int:
        return self.aux.nbf

    @property
    def nelectron(self) -> int:
        return int(self.z.sum()) - self.charge

    def nuclear_repulsion(self) -> float:
        """E_nuc = Σ_{A<B} Z_A Z_B / R_AB (src/geom_io/mod.rs:163-179)."""
        r = self.xyz_bohr
        e = 0.0
        for a in range(len(self.z)):
            d = np.linalg.norm(r[a + 1:] - r[a], axis=1)
            e += float(np.sum(self.z[a] * self.z[a + 1:] / d))
        return e

    # -- integrals ------------------------------------------------------------
    def int1e(self):
        """(S, T, V) full nbf×nbf (symmetric, so row- and column-major agree)."""
        return int1e(self.orb, self.z, self.xyz_bohr)

    def int1e_aux_overlap(self):
        return int1e(self.aux)[0]

    def int2c2e(self) -> np.ndarray:
        return int2c2e(self.aux)

    def int3c2e_packed(self, j0: int = 0, j1: Optional[int] = None) -> np.ndarray:
        return int3c2e_packed(self.orb, self.aux, j0, j1)

    def build_rimatr(self, ri: "jk.DeviceRI", threshold: float = 1e-8, store: int = jk.STORE_FP64) -> dict:
        """Packed B = (μν|Q)·V^{-1/2} into ri's HBM storage; returns timing and kept eigenvalue count."""
        lib = jk.load_library()
        kept = C.c_size_t(0)
        secs = (C.c_double * 3)()
        rc = lib.rsi_build_rimatr(ri.h, self.orb.h, self.aux.h, float(threshold), store, C.byref(kept), secs)
        if rc != 0:
            raise jk.RJKError(rc, lib.rjk_last_error(ri.h).decode())
        return {"naux_kept": int(kept.value), "t_3c": secs[0], "t_metric": secs[1], "t_gemm_tail": secs[2]}

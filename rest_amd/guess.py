"""Initial guesses for the device SCF (REST src/initial_guess).

SAD (superposition of atomic densities), restating initial_guess_from_sad
(src/initial_guess/sad.rs:8-86):
  - one SCF per unique element, the atom alone at the origin in the
    molecule's basis and auxiliary basis, Hartree-Fock with RI-V, core
    Hamiltonian guess, DIIS, REST's default thresholds (InputKeywords::
    init_ctrl, src/ctrl_io/mod.rs:216-224);
  - spin by element (ctrl_setting_atom_sad, sad.rs:140-150): UHF with the
    listed multiplicity for groups 1, 13-17; otherwise RHF whose density is
    split evenly between the spins;
  - block-diagonal assembly in atom order (block_diag_specific, sad.rs:88-117)
    and D = D_α + D_β for a closed-shell molecule (sad.rs:71-72).
The guess step that follows (F(D0) by the density-only J/K, diagonalise,
new density; src/initial_guess/mod.rs:54-64) is rscf_guess_step in the
device library.  The atomic SCFs run on the device through the same J/K and
SCF code as the molecule.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from . import jk, ri
from . import scf as rs

# (multiplicity, spin channels) per element, sad.rs:140-150
_SAD_SPIN = {}
for _els, _m in ((("H", "Li", "Na", "K", "Rb", "Cs", "Fr"), 2), (("B", "Al", "Ga", "In", "Tl", "Nh"), 2),
                 (("C", "Si", "Ge", "Sn", "Pb", "Fl"), 3), (("N", "P", "As", "Sb", "Bi", "Mc"), 4),
                 (("O", "S", "Se", "Te", "Po", "Lv"), 3), (("F", "Cl", "Br", "I", "At", "Ts"), 2)):
    for _e in _els:
        _SAD_SPIN[_e] = (_m, 2)


def sad_spin(element: str) -> tuple:
    """(multiplicity, spin channels) of the atomic SCF for `element`."""
    return _SAD_SPIN.get(element, (1, 1))


def atomic_ctrl() -> "rs.Ctrl":
    """REST's default SCF controls, as the atomic SCFs use them (src/ctrl_io/mod.rs:216-224)."""
    return rs.default_ctrl(num_max_diis=8, start_diis_cycle=1, mix_param=0.6, max_scf_cycle=100, scf_acc_rho=1e-6,
                           scf_acc_eev=1e-5, scf_acc_etot=1e-8)


_ATOM_CACHE: dict = {}


@dataclass
class AtomicDensity:
    element: str
    d_alpha: np.ndarray
    d_beta: np.ndarray
    energy: float
    iterations: int
    converged: bool


def atomic_density(element: str, basis: str, auxbasis: str) -> AtomicDensity:
    """Converged (D_α, D_β) of the isolated atom (cached per element and basis pair)."""
    key = (element, basis, auxbasis)
    if key in _ATOM_CACHE:
        return _ATOM_CACHE[key]
    atom = ri.Molecule([element], np.zeros((1, 3)), basis, auxbasis)
    S, T, V = atom.int1e()
    H = T + V
    z = ri.ZNUC[element]
    mult, nspin = sad_spin(element)
    with jk.DeviceRI(atom.nbf, atom.naux) as r:
        atom.build_rimatr(r)
        if nspin == 2:
            unpaired = mult - 1
            nb = (z - unpaired) // 2
            out = rs.uhf(r, S, H, 0.0, nb + unpaired, nb, None, atomic_ctrl())
            da, db = out.dm[0], out.dm[1]
        else:
            out = rs.rhf(r, S, H, 0.0, z, None, atomic_ctrl())
            da = db = 0.5 * out.dm
    res = AtomicDensity(element, da, db, out.energy, out.iterations, out.converged)
    _ATOM_CACHE[key] = res
    return res


def sad_density(mol: "ri.Molecule") -> tuple:
    """Block-diagonal (D_α, D_β) of the molecule from the atomic densities."""
    n = mol.nbf
    da = np.zeros((n, n))
    db = np.zeros((n, n))
    off = 0
    for el in mol.symbols:
        a = atomic_density(el, mol.basis, mol.auxbasis)
        k = a.d_alpha.shape[0]
        da[off:off + k, off:off + k] = a.d_alpha
        db[off:off + k, off:off + k] = a.d_beta
        off += k
    if off != n:
        raise RuntimeError(f"SAD blocks cover {off} functions, the molecule has {n}")
    return da, db


def sad_factors(mol: "ri.Molecule", cut: float = 1e-13) -> tuple:
    """(C0, w0) with D_α + D_β = C0 diag(w0) C0ᵀ, from the eigenpairs of every atomic block
    (|w| > cut·max|w| of that block); the columns are zero outside their atom's functions."""
    n = mol.nbf
    cols, ws = [], []
    off = 0
    cache = {}
    for el in mol.symbols:
        if el not in cache:
            a = atomic_density(el, mol.basis, mol.auxbasis)
            lam, U = np.linalg.eigh(a.d_alpha + a.d_beta)
            keep = np.abs(lam) > cut * np.abs(lam).max()
            cache[el] = (lam[keep], U[:, keep])
        lam, U = cache[el]
        k = U.shape[0]
        block = np.zeros((n, U.shape[1]))
        block[off:off + k] = U
        cols.append(block)
        ws.append(lam)
        off += k
    return np.concatenate(cols, axis=1), np.concatenate(ws)


@dataclass
class Guess:
    d0: np.ndarray          # assembled guess density (RHF: D_α + D_β)
    energy_d0: float        # E(D0), REST's "Initial guess HF energy"
    d1: np.ndarray          # density of the orbitals of F(D0): where the SCF starts
    c1: np.ndarray = None   # those orbitals (the SCF's first Fock uses them, as REST's scf() does)
    kind: str = "sad"


def sad_guess(mol: "ri.Molecule", r: "jk.DeviceRI", S: np.ndarray, H: np.ndarray) -> Guess:
    """REST's SAD initial guess for a closed-shell molecule whose B is resident in r."""
    da, db = sad_density(mol)
    d0 = da + db
    e0, d1, c1, _ = rs.guess_step(r, S, H, mol.nuclear_repulsion(), (mol.nelectron,), d0, factors=sad_factors(mol))
    return Guess(d0, e0, d1, c1, "sad")

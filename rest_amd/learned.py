"""Learned initial guess: a per-atom-pair MLP that predicts density-matrix
blocks (a new `initial_guess` branch beside REST's "sad",
src/initial_guess/mod.rs:15-72).

Representation
  - Pairs: every atom pair (A, B), A < B, within R_CUT bohr, and every atom
    with itself.  For each, a local frame (rows e_x, e_y, e_z):
      A ≠ B: e_z along B − A; e_x from the environment vector of A
             v_A = Σ_C w(|C − A|)(C − A) over atoms C ∉ {A, B} (B's if A's is
             parallel to e_z), w(r) = exp(−r²/σ_ref²);
      A = B: e_z along v_A (all neighbours), e_x from the nearest neighbour.
  - Inputs (N_IN numbers): r_AB/10 and 1/r_AB (0 for A = B), the ordered
    element pair one-hot over ELEMENTS² and a same-atom flag, 16-bin Gaussian
    radial neighbour features of A and of B per neighbour element, v_A and v_B
    in the pair frame with their lengths, and the NNEIGH nearest neighbours of
    A and of B (other than A and B) in the pair frame: coordinates / 5 bohr,
    1/r and element one-hot.
  - Outputs: the pair's block of D − D_sph in the pair frame, on a MAXB × MAXB
    grid (rows: A's functions, columns: B's) with a mask for the element
    pair's real block size, scaled by TARGET_SCALE.  D is the converged
    closed-shell density; D_sph is the superposition of spherically averaged
    atomic densities (SAD's atomic blocks averaged over rotations: within an
    atom, blocks between two shells of equal l become tr/(2l+1) times the
    identity, all others zero).  The average makes the target independent of
    the lab-frame axis along which an open-shell atom's SCF put its partly
    filled p shell.
  - Blocks are carried between the molecule frame and the pair frame by the
    real-solid-harmonic rotation matrices of the engine's own convention
    (rsi_cart2sph): D_loc = U_A D U_Bᵀ.

Guess: D = D_sph + predicted blocks (pairs beyond R_CUT keep zero),
symmetrised, scaled to the molecule's electron count (tr D S = N),
McWeeny-purified (rscf_mcweeny), and accepted by one of two guards, else the
SAD guess is used: the guess build's own self-consistency error
max|F D S - S D F| below GUESS_ERR_MAX (learned_scf's default), or its energy
E(D) below the energy of SAD's first orbital density (REST's guess step on
the SAD density; learned_guess, and learned_scf with guard="energy").

The network runs on the device through include/rest_lg.h (BF16 matrix-core
GEMMs, FP32 master weights, Adam); torch only provides device buffers.
"""
from __future__ import annotations

import ctypes as C
import json
import math
import time
from concurrent.futures import ThreadPoolExecutor
from dataclasses import dataclass, field
from pathlib import Path
from typing import Optional

import numpy as np

from . import guess as gs
from . import jk, ri
from . import scf as rs
from . import workload as wl

ELEMENTS = ("H", "C", "N", "O")
NBINS = 16
MU = 1.0 + 0.6 * np.arange(NBINS)       # radial bin centres (bohr)
SIGMA_BIN = 0.6
R_CUT = 10.0                            # pair cut-off (bohr)
R_ENV = 12.0                            # neighbour cut-off of the radial features (bohr)
SIGMA_REF = 2.0                         # width of the environment-vector weights (bohr)
MAXB = 14                               # largest per-atom block (O, C, N in def2-SVP)
TARGET_SCALE = 4.0
NNEIGH = 4                              # nearest neighbours of A and of B described in the pair frame
NB_FEAT = 4 + len(ELEMENTS)             # local x, y, z (/5 bohr), 1/r, element one-hot
N_IN = 2 + len(ELEMENTS) ** 2 + 1 + 2 * len(ELEMENTS) * NBINS + 8 + 2 * NNEIGH * NB_FEAT
N_OUT = MAXB * MAXB
MODEL_DIR = Path(__file__).resolve().parent / "models"


# ---------------------------------------------------------------------------
# rotations of real solid harmonics in the engine's convention
# ---------------------------------------------------------------------------
def _cart_exps(l):
    return [(x, y, l - x - y) for x in range(l, -1, -1) for y in range(l - x, -1, -1)]


def _sphere_moment(a, b, c):
    if a % 2 or b % 2 or c % 2:
        return 0.0
    g = math.gamma
    return 2.0 * g((a + 1) / 2) * g((b + 1) / 2) * g((c + 1) / 2) / g((a + b + c + 3) / 2)


_ROWS: dict = {}


def harmonic_rows(l: int) -> np.ndarray:
    """Engine rows (rsi_cart2sph) of shell l, each scaled to unit norm on the unit sphere."""
    if l in _ROWS:
        return _ROWS[l]
    nc = (l + 1) * (l + 2) // 2
    T = np.empty((2 * l + 1) * nc)
    rc = jk.load_library().rsi_cart2sph(int(l), T.ctypes.data)
    if rc != 0:
        raise RuntimeError(f"rsi_cart2sph({l}) failed")
    T = T.reshape(2 * l + 1, nc)
    ex = _cart_exps(l)
    G = np.array([[_sphere_moment(*(np.add(e1, e2))) for e2 in ex] for e1 in ex])
    T = T / np.sqrt(np.einsum("mi,ij,mj->m", T, G, T))[:, None]
    _ROWS[l] = T
    return T


_PTS = None


def _sample_points():
    global _PTS
    if _PTS is None:
        rng = np.random.default_rng(12345)
        p = rng.standard_normal((48, 3))
        _PTS = p / np.linalg.norm(p, axis=1, keepdims=True)
    return _PTS


def _eval_rows(l, P):
    ex = _cart_exps(l)
    mono = np.stack([P[:, 0] ** a * P[:, 1] ** b * P[:, 2] ** c for a, b, c in ex], axis=1)
    return mono @ harmonic_rows(l).T


_PINV: dict = {}


def sph_rotations(l: int, Rs: np.ndarray) -> np.ndarray:
    """D^l(R) for a stack of rotations Rs [k, 3, 3]: for the molecule rotated by R (r → R r),
    shell-l coefficients transform as c' = D^l c.  Least squares of Y_m(Rᵀ p) = Σ_m' Y_m'(p) D_m'm
    over fixed sample points p on the unit sphere (exact: the rows span an invariant space)."""
    Rs = np.asarray(Rs, dtype=np.float64).reshape(-1, 3, 3)
    if l == 0:
        return np.ones((len(Rs), 1, 1))
    P = _sample_points()
    if l not in _PINV:
        _PINV[l] = np.linalg.pinv(_eval_rows(l, P))
    PR = np.matmul(P[None], Rs)                        # rows (Rᵀ p)ᵀ = pᵀ R
    ex = _cart_exps(l)
    mono = np.stack([PR[..., 0] ** a * PR[..., 1] ** b * PR[..., 2] ** c for a, b, c in ex], axis=-1)
    return _PINV[l] @ (mono @ harmonic_rows(l).T)


def sph_rotation(l: int, R: np.ndarray) -> np.ndarray:
    return sph_rotations(l, np.asarray(R)[None])[0]


def atom_rotations(shell_ls, Rs: np.ndarray) -> np.ndarray:
    """Block-diagonal rotations [k, n, n] of one atom's functions (its shells in basis order)."""
    Rs = np.asarray(Rs, dtype=np.float64).reshape(-1, 3, 3)
    Dl = {l: sph_rotations(l, Rs) for l in set(shell_ls)}
    n = sum(2 * l + 1 for l in shell_ls)
    U = np.zeros((len(Rs), n, n))
    o = 0
    for l in shell_ls:
        k = 2 * l + 1
        U[:, o:o + k, o:o + k] = Dl[l]
        o += k
    return U


def atom_rotation(shell_ls, R: np.ndarray) -> np.ndarray:
    return atom_rotations(shell_ls, np.asarray(R)[None])[0]


def molecule_rotation(mol: "ri.Molecule", R: np.ndarray) -> np.ndarray:
    """U with D(rotated molecule) = U D Uᵀ for the whole orbital basis."""
    n = mol.nbf
    U = np.zeros((n, n))
    for a, (o, ls) in enumerate(atom_layout(mol)):
        Ua = atom_rotation(ls, R)
        U[o:o + Ua.shape[0], o:o + Ua.shape[0]] = Ua
    return U


def spherical_average(mol: "ri.Molecule", D: np.ndarray) -> np.ndarray:
    """Rotation average of the atom-diagonal blocks of D; off-atom blocks are dropped."""
    b = mol.orb
    out = np.zeros_like(D)
    for a in range(len(mol.symbols)):
        idx = np.nonzero(b.shell_atom == a)[0]
        for i in idx:
            for j in idx:
                li, lj = int(b.shell_l[i]), int(b.shell_l[j])
                if li != lj:
                    continue
                oi, oj, k = int(b.shell_off[i]), int(b.shell_off[j]), 2 * li + 1
                blk = D[oi:oi + k, oj:oj + k]
                out[oi:oi + k, oj:oj + k] = np.trace(blk) / k * np.eye(k)
    return out


def atom_layout(mol: "ri.Molecule") -> list:
    """[(first function, [shell l ...]) per atom] of the orbital basis."""
    b = mol.orb
    out = []
    for a in range(len(mol.symbols)):
        idx = np.nonzero(b.shell_atom == a)[0]
        out.append((int(b.shell_off[idx[0]]), [int(l) for l in b.shell_l[idx]]))
    return out


# ---------------------------------------------------------------------------
# pairs, frames and descriptors
# ---------------------------------------------------------------------------
def _unit(v):
    n = np.linalg.norm(v)
    return v / n if n > 0 else v


def _perp(v, z):
    return v - np.dot(v, z) * z


def pair_list(xyz: np.ndarray, cutoff: float = R_CUT) -> np.ndarray:
    """(A, B) with A <= B and |B − A| <= cutoff, diagonal included, in row-major order."""
    d = np.linalg.norm(xyz[:, None, :] - xyz[None, :, :], axis=-1)
    A, B = np.nonzero(np.triu(d <= cutoff))
    return np.stack([A, B], axis=1)


def _env_vector(xyz, a, exclude):
    d = xyz - xyz[a]
    r2 = np.sum(d * d, axis=1)
    w = np.exp(-r2 / SIGMA_REF ** 2)
    w[list(exclude)] = 0.0
    return (w[:, None] * d).sum(axis=0)


def pair_frame(xyz: np.ndarray, a: int, b: int) -> np.ndarray:
    """Rows e_x, e_y, e_z of the local frame of pair (a, b)."""
    if a != b:
        z = _unit(xyz[b] - xyz[a])
        cands = [_env_vector(xyz, a, (a, b)), _env_vector(xyz, b, (a, b)), np.array([1.0, 0.0, 0.0]),
                 np.array([0.0, 1.0, 0.0])]
    else:
        v = _env_vector(xyz, a, (a,))
        d = np.linalg.norm(xyz - xyz[a], axis=1)
        d[a] = np.inf
        order = np.argsort(d, kind="stable")
        z = _unit(v) if np.linalg.norm(v) > 1e-8 else (_unit(xyz[order[0]] - xyz[a]) if len(xyz) > 1
                                                        else np.array([0.0, 0.0, 1.0]))
        cands = [xyz[c] - xyz[a] for c in order[:3] if np.isfinite(d[c])] + [np.array([1.0, 0.0, 0.0]),
                                                                              np.array([0.0, 1.0, 0.0])]
    for v in cands:
        p = _perp(v, z)
        if np.linalg.norm(p) > 1e-3 * max(1.0, np.linalg.norm(v)):
            x = _unit(p)
            return np.stack([x, np.cross(z, x), z])
    raise RuntimeError("no local frame")


def radial_features(symbols, xyz: np.ndarray) -> np.ndarray:
    """[natom, len(ELEMENTS) * NBINS] Gaussian radial neighbour histograms."""
    n = len(symbols)
    el = np.array([ELEMENTS.index(s) for s in symbols])
    d = np.linalg.norm(xyz[:, None, :] - xyz[None, :, :], axis=-1)
    g = np.exp(-((d[:, :, None] - MU[None, None, :]) / SIGMA_BIN) ** 2)
    g[(d > R_ENV) | (d == 0.0)] = 0.0
    out = np.zeros((n, len(ELEMENTS), NBINS))
    for e in range(len(ELEMENTS)):
        out[:, e, :] = g[:, el == e, :].sum(axis=1)
    return out.reshape(n, -1)


def _first_valid(cands: np.ndarray, z: np.ndarray) -> np.ndarray:
    """Per row, unit(perp(first candidate whose component ⊥ z is not negligible)); cands [k, m, 3]."""
    perp = cands - np.einsum("kmi,ki->km", cands, z)[:, :, None] * z[:, None, :]
    pn = np.linalg.norm(perp, axis=-1)
    ok = pn > 1e-3 * np.maximum(1.0, np.linalg.norm(cands, axis=-1))
    if not ok.any(axis=1).all():
        raise RuntimeError("no local frame")
    first = np.argmax(ok, axis=1)
    rows = np.arange(len(cands))
    return perp[rows, first] / pn[rows, first][:, None]


def frames(xyz: np.ndarray, pairs: np.ndarray) -> np.ndarray:
    """Local frames [k, 3, 3] (rows e_x, e_y, e_z) of all pairs; the same frames as pair_frame."""
    n = len(xyz)
    d = xyz[None, :, :] - xyz[:, None, :]                  # d[a, c] = x_c − x_a
    r = np.linalg.norm(d, axis=-1)
    w = np.exp(-r ** 2 / SIGMA_REF ** 2)
    np.fill_diagonal(w, 0.0)
    v_all = np.einsum("ac,aci->ai", w, d)
    A, B = pairs[:, 0], pairs[:, 1]
    F = np.empty((len(pairs), 3, 3))
    ex = np.broadcast_to(np.array([1.0, 0.0, 0.0]), (len(pairs), 3))
    ey = np.broadcast_to(np.array([0.0, 1.0, 0.0]), (len(pairs), 3))
    off = A != B
    if off.any():
        a, b = A[off], B[off]
        z = d[a, b] / r[a, b][:, None]
        va = v_all[a] - w[a, b][:, None] * d[a, b]
        vb = v_all[b] - w[b, a][:, None] * d[b, a]
        x = _first_valid(np.stack([va, vb, ex[off], ey[off]], axis=1), z)
        F[off] = np.stack([x, np.cross(z, x), z], axis=1)
    dg = ~off
    if dg.any():
        a = A[dg]
        rr = r.copy()
        np.fill_diagonal(rr, np.inf)
        order = np.argsort(rr, axis=1, kind="stable")[:, :3]
        vn = np.linalg.norm(v_all[a], axis=1)
        near = d[a, order[a, 0]] if n > 1 else np.broadcast_to(np.array([0.0, 0.0, 1.0]), (len(a), 3))
        zc = np.where((vn > 1e-8)[:, None], v_all[a], near)
        z = zc / np.linalg.norm(zc, axis=1, keepdims=True)
        nb = [d[a, order[a, j]] if j < n - 1 else np.zeros((len(a), 3)) for j in range(3)]
        x = _first_valid(np.stack(nb + [ex[dg], ey[dg]], axis=1), z)
        F[dg] = np.stack([x, np.cross(z, x), z], axis=1)
    return F


def descriptors(symbols, xyz: np.ndarray, pairs: np.ndarray):
    """(features [npair, N_IN] float32, frames [npair, 3, 3])."""
    n = len(symbols)
    ne = len(ELEMENTS)
    el = np.array([ELEMENTS.index(s) for s in symbols])
    rad = radial_features(symbols, xyz)
    F = frames(xyz, pairs)
    A, B = pairs[:, 0], pairs[:, 1]
    k = len(pairs)
    d = xyz[None, :, :] - xyz[:, None, :]
    r = np.linalg.norm(d, axis=-1)
    w = np.exp(-r ** 2 / SIGMA_REF ** 2)
    np.fill_diagonal(w, 0.0)
    v_all = np.einsum("ac,aci->ai", w, d)
    same = A == B
    rab = r[A, B]
    X = np.zeros((k, N_IN))
    X[:, 0] = rab / 10.0
    X[:, 1] = np.where(same, 0.0, 1.0 / np.where(same, 1.0, rab))
    o = 2
    X[np.arange(k), o + el[A] * ne + el[B]] = 1.0
    o += ne * ne
    X[:, o] = same
    o += 1
    X[:, o:o + rad.shape[1]] = rad[A]
    o += rad.shape[1]
    X[:, o:o + rad.shape[1]] = rad[B]
    o += rad.shape[1]
    va = v_all[A] - np.where(same, 0.0, w[A, B])[:, None] * d[A, B]
    vb = v_all[B] - np.where(same, 0.0, w[B, A])[:, None] * d[B, A]
    la = np.einsum("kij,kj->ki", F, va)
    lb = np.einsum("kij,kj->ki", F, vb)
    X[:, o:o + 3] = la
    X[:, o + 3] = np.linalg.norm(la, axis=1)
    X[:, o + 4:o + 7] = lb
    X[:, o + 7] = np.linalg.norm(lb, axis=1)
    o += 8
    # nearest neighbours of A and of B other than A and B, in the pair frame
    rr = r.copy()
    np.fill_diagonal(rr, np.inf)
    order = np.argsort(rr, axis=1, kind="stable")[:, :NNEIGH + 1]      # [n, <= NNEIGH+1]
    if order.shape[1] < NNEIGH + 1:   # fewer atoms: pad with the atom itself (never valid)
        order = np.concatenate([order, np.repeat(np.arange(n)[:, None], NNEIGH + 1 - order.shape[1], axis=1)], axis=1)
    for atom, other in ((A, B), (B, A)):
        cand = order[atom]                                              # [k, NNEIGH+1]
        valid = (cand != other[:, None]) & np.isfinite(rr[atom[:, None], cand])
        sel = np.argsort(~valid, axis=1, kind="stable")[:, :NNEIGH]
        c = np.take_along_axis(cand, sel, axis=1)
        ok = np.take_along_axis(valid, sel, axis=1)
        loc = np.einsum("kij,kmj->kmi", F, d[atom[:, None], c])       # [k, NNEIGH, 3]
        feat = np.zeros((k, NNEIGH, NB_FEAT))
        feat[..., 0:3] = loc / 5.0
        feat[..., 3] = 1.0 / np.maximum(r[atom[:, None], c], 1e-6)
        feat[np.arange(k)[:, None], np.arange(NNEIGH)[None, :], 4 + el[c]] = 1.0
        feat[~ok] = 0.0
        X[:, o:o + NNEIGH * NB_FEAT] = feat.reshape(k, -1)
        o += NNEIGH * NB_FEAT
    assert o == N_IN
    return X.astype(np.float32), F


# ---------------------------------------------------------------------------
# encoding densities as per-pair targets and back
# ---------------------------------------------------------------------------
def _groups(mol: "ri.Molecule", pairs: np.ndarray):
    """Pairs grouped by ordered element pair: (rows, offsets A, offsets B, shells A, shells B)."""
    lay = atom_layout(mol)
    off = np.array([o for o, _ in lay])
    sym = np.array(mol.symbols)
    out = []
    keys = sorted(set(zip(sym[pairs[:, 0]], sym[pairs[:, 1]])))
    for ea, eb in keys:
        rows = np.nonzero((sym[pairs[:, 0]] == ea) & (sym[pairs[:, 1]] == eb))[0]
        la = lay[pairs[rows[0], 0]][1]
        lb = lay[pairs[rows[0], 1]][1]
        out.append((rows, off[pairs[rows, 0]], off[pairs[rows, 1]], la, lb))
    return out


def encode(mol: "ri.Molecule", D: np.ndarray, D_sad: np.ndarray):
    """Training arrays (X, T, M) of one molecule with converged density D and baseline D_sad
    (the spherically averaged SAD density in training)."""
    pairs = pair_list(mol.xyz_bohr)
    X, F = descriptors(mol.symbols, mol.xyz_bohr, pairs)
    T = np.zeros((len(pairs), MAXB, MAXB), dtype=np.float32)
    M = np.zeros_like(T)
    dD = D - D_sad
    for rows, oa, ob, la, lb in _groups(mol, pairs):
        Ua, Ub = atom_rotations(la, F[rows]), atom_rotations(lb, F[rows])
        na, nb = Ua.shape[1], Ub.shape[1]
        if na > MAXB or nb > MAXB:
            raise ValueError(f"atom blocks above {MAXB} functions are not supported")
        ia = oa[:, None] + np.arange(na)[None, :]
        ib = ob[:, None] + np.arange(nb)[None, :]
        blk = dD[ia[:, :, None], ib[:, None, :]]
        T[rows, :na, :nb] = TARGET_SCALE * (Ua @ blk @ np.transpose(Ub, (0, 2, 1)))
        M[rows, :na, :nb] = 1.0
    return X, T.reshape(len(pairs), -1), M.reshape(len(pairs), -1)


def decode(mol: "ri.Molecule", pairs: np.ndarray, frames_: np.ndarray, Y: np.ndarray, D_sad: np.ndarray):
    """D = D_sad + Σ predicted pair blocks rotated back, symmetric.  The pairs are unique with
    A <= B (pair_list), so every matrix element belongs to one block or its mirror: the blocks are
    written, not accumulated."""
    dD = np.zeros_like(D_sad)
    Y = np.asarray(Y, dtype=np.float64).reshape(len(pairs), MAXB, MAXB) / TARGET_SCALE
    for rows, oa, ob, la, lb in _groups(mol, pairs):
        Ua = atom_rotations(la, frames_[rows])
        Ub = Ua if list(lb) == list(la) else atom_rotations(lb, frames_[rows])
        na, nb = Ua.shape[1], Ub.shape[1]
        blk = np.transpose(Ua, (0, 2, 1)) @ Y[rows, :na, :nb] @ Ub
        same = pairs[rows, 0] == pairs[rows, 1]
        if same.any():   # only element-diagonal groups hold atom-diagonal blocks (square)
            blk[same] = 0.5 * (blk[same] + np.transpose(blk[same], (0, 2, 1)))
        ia = oa[:, None] + np.arange(na)[None, :]
        ib = ob[:, None] + np.arange(nb)[None, :]
        dD[ia[:, :, None], ib[:, None, :]] = blk
        nd = ~same
        dD[ib[nd][:, :, None], ia[nd][:, None, :]] = np.transpose(blk[nd], (0, 2, 1))
    return D_sad + dD


# ---------------------------------------------------------------------------
# the device network
# ---------------------------------------------------------------------------
class Net:
    """ctypes owner of an rlg_net (include/rest_lg.h)."""

    def __init__(self, n_in: int = N_IN, n_out: int = N_OUT, hidden: int = 256, layers: int = 3,
                 max_batch: int = 8192, device: int = 0, seed: int = 0):
        self.lib = jk.load_library()
        self.dims = [n_in] + [hidden] * layers + [n_out]
        self.max_batch = max_batch
        h = C.c_void_p()
        rc = self.lib.rlg_create(int(device), n_in, hidden, layers, n_out, max_batch, C.byref(h))
        if rc != 0:
            raise RuntimeError(f"rlg_create failed ({rc})")
        self.h = h
        self.nparam = int(self.lib.rlg_num_params(h))
        self.set_params(self.init_params(seed))
        self.input_mean = np.zeros(n_in, dtype=np.float32)
        self.input_std = np.ones(n_in, dtype=np.float32)

    def init_params(self, seed: int) -> np.ndarray:
        """He-uniform weights for the SiLU layers, zero biases."""
        rng = np.random.default_rng(seed)
        out = []
        for l, (din, dout) in enumerate(zip(self.dims[:-1], self.dims[1:])):
            lim = math.sqrt(6.0 / din) if l + 1 < len(self.dims) - 1 else math.sqrt(3.0 / din) * 0.1
            out.append(rng.uniform(-lim, lim, size=din * dout))
            out.append(np.zeros(dout))
        return np.concatenate(out).astype(np.float32)

    def _check(self, rc, what):
        if rc != 0:
            raise RuntimeError(f"{what}: {self.lib.rlg_last_error(self.h).decode()}")

    def close(self):
        if getattr(self, "h", None):
            self.lib.rlg_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_params(self, p: np.ndarray):
        p = np.ascontiguousarray(p, dtype=np.float32)
        assert p.size == self.nparam
        self._check(self.lib.rlg_set_params(self.h, p.ctypes.data), "rlg_set_params")

    def get_params(self) -> np.ndarray:
        p = np.empty(self.nparam, dtype=np.float32)
        self._check(self.lib.rlg_get_params(self.h, p.ctypes.data), "rlg_get_params")
        return p

    def set_stream(self, stream_ptr: int):
        self._check(self.lib.rlg_set_stream(self.h, C.c_void_p(stream_ptr)), "rlg_set_stream")

    def normalise(self, X: np.ndarray) -> np.ndarray:
        return ((X - self.input_mean) / self.input_std).astype(np.float32)

    def forward_dev(self, x, y):
        """x, y: contiguous float32 torch CUDA tensors [batch, dims]."""
        self._check(self.lib.rlg_forward(self.h, C.c_void_p(x.data_ptr()), x.shape[0], C.c_void_p(y.data_ptr())),
                    "rlg_forward")

    def predict(self, X: np.ndarray) -> np.ndarray:
        """Host features (not yet normalised) -> host outputs."""
        import torch
        Xn = self.normalise(X)
        out = np.empty((len(Xn), self.dims[-1]), dtype=np.float32)
        for i0 in range(0, len(Xn), self.max_batch):
            x = torch.from_numpy(Xn[i0:i0 + self.max_batch]).cuda()
            y = torch.empty((x.shape[0], self.dims[-1]), dtype=torch.float32, device="cuda")
            self.forward_dev(x, y)
            out[i0:i0 + len(x)] = y.cpu().numpy()
        return out

    def loss_grad_dev(self, x, t, m, want_grad: bool = False):
        loss = C.c_double(0.0)
        g = np.empty(self.nparam, dtype=np.float32) if want_grad else None
        self._check(self.lib.rlg_loss_grad(self.h, C.c_void_p(x.data_ptr()), C.c_void_p(t.data_ptr()),
                                           C.c_void_p(m.data_ptr()), x.shape[0], C.byref(loss),
                                           None if g is None else g.ctypes.data), "rlg_loss_grad")
        return loss.value, g

    def adam(self, lr: float, b1: float = 0.9, b2: float = 0.999, eps: float = 1e-8):
        self._check(self.lib.rlg_adam_step(self.h, lr, b1, b2, eps), "rlg_adam_step")

    def reset_optimizer(self):
        self._check(self.lib.rlg_reset_optimizer(self.h), "rlg_reset_optimizer")

    def save(self, path: Path, meta: Optional[dict] = None):
        np.savez(path, params=self.get_params(), dims=np.array(self.dims), input_mean=self.input_mean,
                 input_std=self.input_std, meta=json.dumps(meta or {}))

    @classmethod
    def load(cls, path: Path, max_batch: int = 8192, device: int = 0) -> "Net":
        z = np.load(path, allow_pickle=False)
        dims = [int(d) for d in z["dims"]]
        net = cls(dims[0], dims[-1], dims[1], len(dims) - 2, max_batch, device)
        net.set_params(z["params"])
        net.input_mean = z["input_mean"].astype(np.float32)
        net.input_std = z["input_std"].astype(np.float32)
        net.meta = json.loads(str(z["meta"]))
        return net


def train(net: Net, X: np.ndarray, T: np.ndarray, M: np.ndarray, epochs: int = 20, batch: int = 4096,
          lr: float = 1e-3, lr_min: float = 3e-5, seed: int = 0, val_fraction: float = 0.05, log=print) -> dict:
    """Masked-MSE training with Adam and a cosine learning-rate schedule, all on the device."""
    import torch
    rng = np.random.default_rng(seed)
    perm = rng.permutation(len(X))
    nval = max(1, int(len(X) * val_fraction))
    net.input_mean = X.mean(axis=0).astype(np.float32)
    net.input_std = (X.std(axis=0) + 1e-3).astype(np.float32)
    Xn = net.normalise(X)[perm]
    x_all = torch.from_numpy(Xn).cuda()
    t_all = torch.from_numpy(np.ascontiguousarray(T[perm])).cuda()
    m_all = torch.from_numpy(np.ascontiguousarray(M[perm])).cuda()
    ntr = len(X) - nval
    nb = max(1, ntr // batch)
    steps = epochs * nb
    hist = {"train": [], "val": []}
    step = 0
    t0 = time.time()
    for ep in range(epochs):
        tot = 0.0
        for bi in rng.permutation(nb):
            i0 = nval + bi * batch
            i1 = min(nval + ntr, i0 + batch)
            cur = lr_min + 0.5 * (lr - lr_min) * (1 + math.cos(math.pi * step / max(1, steps - 1)))
            loss, _ = net.loss_grad_dev(x_all[i0:i1], t_all[i0:i1], m_all[i0:i1])
            net.adam(cur)
            tot += loss
            step += 1
        vl = 0.0
        for i0 in range(0, nval, batch):
            i1 = min(nval, i0 + batch)
            vl += net.loss_grad_dev(x_all[i0:i1], t_all[i0:i1], m_all[i0:i1])[0] * (i1 - i0)
        hist["train"].append(tot / nb)
        hist["val"].append(vl / nval)
        log(f"epoch {ep + 1:3d}/{epochs}  train {tot / nb:.3e}  val {vl / nval:.3e}  ({time.time() - t0:.1f}s)")
    return hist


# ---------------------------------------------------------------------------
# the guess
# ---------------------------------------------------------------------------
def predict_density(net: Net, mol: "ri.Molecule", D_sad: np.ndarray) -> np.ndarray:
    pairs = pair_list(mol.xyz_bohr)
    X, F = descriptors(mol.symbols, mol.xyz_bohr, pairs)
    Y = net.predict(X)
    return decode(mol, pairs, F, Y, D_sad)


@dataclass
class LearnedGuess:
    c1: np.ndarray              # orbitals the SCF starts from (rscf_rhf_orbitals)
    kind: str                   # "learned" or "sad"
    e_learned: float            # E(D) of the purified prediction
    e_sad: float                # E(D1) of SAD's first orbital density
    electrons: float            # 2 tr(PS) of the purified prediction
    seconds: dict = field(default_factory=dict)


def learned_guess(net: Net, mol: "ri.Molecule", r: "jk.DeviceRI", S, H, prec: int = jk.PREC_FP64,
                  mcweeny_iterations: int = 6) -> LearnedGuess:
    """Predict, purify, and keep the prediction only if its energy is below that of SAD's first
    orbital density.

    Both energies are of idempotent densities with the molecule's electron count, so both are
    upper bounds of the converged energy and compare fairly (SAD's own D0 is not idempotent and
    its energy can lie below the converged one).  The purified prediction then takes REST's
    guess step (F(D), diagonalise, new density) like SAD's D0 does."""
    t0 = time.perf_counter()
    da, db = gs.sad_density(mol)
    d_sad = da + db
    t1 = time.perf_counter()
    d_pred = predict_density(net, mol, spherical_average(mol, d_sad))
    d_pred *= mol.nelectron / float(np.sum(d_pred * S))
    t2 = time.perf_counter()
    d_pur, ne = rs.mcweeny(r, S, d_pred, mcweeny_iterations)
    t3 = time.perf_counter()
    enuc = mol.nuclear_repulsion()
    nel = (mol.nelectron,)
    _, _, c1_s, _, e_s = rs.guess_step(r, S, H, enuc, nel, d_sad, prec, want_e1=True)
    t4 = time.perf_counter()
    ok = bool(np.all(np.isfinite(d_pur))) and abs(ne - mol.nelectron) < 0.5
    e_l, c1_l = float("inf"), None
    if ok:
        e_l, _, c1_l, _ = rs.guess_step(r, S, H, enuc, nel, d_pur, prec)
        ok = bool(np.isfinite(e_l)) and e_l < e_s
    t5 = time.perf_counter()
    secs = {"sad": t1 - t0, "predict": t2 - t1, "mcweeny": t3 - t2, "sad_energy": t4 - t3, "learned_step": t5 - t4}
    return LearnedGuess(c1_l if ok else c1_s, "learned" if ok else "sad", e_l, e_s, ne, secs)


# Acceptance bar of the "error" guard: max|F D S - S D F| of the purified prediction's guess
# Fock build (BF16X2 J/K).  Measured with the packaged model: 0.044-0.095 on 64 held-out (H2O)64
# and 0.037-0.052 on (H2O)8, 16 and 32, against 0.23-0.27 for SAD's first orbital density (REST's
# guess step) and about 1.9 for the SAD density itself; a prediction at 0.15 or above is no clear
# gain over SAD's step.
GUESS_ERR_MAX = 0.15


@dataclass
class LearnedScf:
    outcome: "rs.ScfOutcome"    # the SCF that was kept
    kind: str                   # "learned" or "sad"
    e_learned: float            # E(D) of the purified prediction (the SCF's guess Fock build)
    e_sad: float                # E(D1) of SAD's first orbital density ("energy" guard; None otherwise)
    electrons: float            # 2 tr(PS) of the purified prediction
    seconds: dict = field(default_factory=dict)
    err_learned: float = None   # max|F D S - S D F| of the prediction's guess build
    band: float = 0.0           # "energy": kept only when e_learned < e_sad - band (the estimates' error bound)
    refined: bool = False       # "energy": e_sad was computed again at FP64 (the first estimate was inconclusive)


# Relative bound of an energy's error when its J/K ran at this precision: used by the energy guard to
# decide on estimates.  Measured at (H2O)64 (E about 4859 Eh): E(D1) of SAD off by 0.73 Eh at BF16X1
# (1.5e-4 relative), 1.2e-3 Eh at BF16X2 (2.5e-7); the bounds keep a factor of two at BF16X1 and four
# at BF16X2.  An estimate outside its bound can only keep or drop a prediction wrongly near the bound;
# the SCF from either start converges to the same energy.
GUARD_BAND_REL = {jk.PREC_FP64: 0.0, jk.PREC_FP32: 1e-6, jk.PREC_BF16X3: 1e-6, jk.PREC_BF16X2: 1e-6,
                  jk.PREC_BF16X1: 3e-4}


def learned_scf(net: Net, mol: "ri.Molecule", r: "jk.DeviceRI", S, H, ctrl=None,
                mcweeny_iterations: int = 6, guard: str = "error",
                guess_err_max: float = GUESS_ERR_MAX, guard_prec: int = None,
                guard_band_rel: dict = None) -> LearnedScf:
    """RHF from the learned guess, with the acceptance rule folded into the SCF.

    The SCF starts from the purified prediction; its guess Fock build F(D_pred) decides, before
    the first iteration, whether the prediction is kept (REST's guess-step sequence continues
    from it without a separate guess-step build) or the SCF stops at once and runs from SAD
    instead (REST's guess step from the SAD density).
      guard "error" (default): kept when the guess build's self-consistency error
        max|F D S - S D F| is below guess_err_max; the test costs two GEMMs.
      guard "energy": SAD's E(D1) is computed first (REST's guess step from the SAD density plus
        the energy of the new orbitals' density, at guard_prec, by default the SCF's starting
        precision); kept when E(D_pred), from the SCF's guess build at its starting precision, is
        below it.  Both are energies of idempotent densities with the molecule's electron count, so
        both bound the converged energy from above and compare fairly.  Each estimate carries an
        error bound (GUARD_BAND_REL of its precision, or guard_band_rel): the prediction is kept
        when E(D_pred) lies below E(D1) by more than both bounds, dropped when it lies above by
        more, and in between E(D1) is computed again at FP64 and decides.
    A prediction that is not finite or misses the electron count by 0.5 or more is never kept."""
    if guard not in ("error", "energy"):
        raise ValueError(f"learned_scf: guard must be 'error' or 'energy', not {guard!r}")
    ctrl = ctrl or rs.default_ctrl()
    prec0 = ctrl.jk_prec_start if ctrl.jk_prec_start >= 0 else ctrl.jk_prec
    t0 = time.perf_counter()
    da, db = gs.sad_density(mol)
    d_sad = da + db
    t1 = time.perf_counter()
    enuc = mol.nuclear_repulsion()
    e_s, band, band_l, refined = None, 0.0, 0.0, False
    rel = {**GUARD_BAND_REL, **(guard_band_rel or {})}
    gp = prec0 if guard_prec is None else guard_prec
    factors = gs.sad_factors(mol) if guard == "energy" else None

    def sad_e1(prec):
        return rs.guess_step(r, S, H, enuc, (mol.nelectron,), d_sad, prec, want_e1=True, want_outputs=False,
                             factors=factors)[4]

    def predict():
        a = time.perf_counter()
        d = predict_density(net, mol, spherical_average(mol, d_sad))
        d *= mol.nelectron / float(np.sum(d * S))
        return d, time.perf_counter() - a

    if guard == "energy":
        # the prediction (host descriptors and decode, the network on its own stream) runs beside
        # SAD's guess step and E(D1) on the J/K context
        with ThreadPoolExecutor(max_workers=1) as ex:
            fut = ex.submit(predict)
            a = time.perf_counter()
            e_s = sad_e1(gp)
            t_e1 = time.perf_counter() - a
            d_pred, t_pred = fut.result()
        band_l = rel[prec0] * abs(e_s)              # E(D_pred) comes from the SCF's first build, at prec0
        band = rel[gp] * abs(e_s) + band_l
    else:
        d_pred, t_pred = predict()
        t_e1 = 0.0
    t2 = time.perf_counter()
    d_pur, ne = rs.mcweeny(r, S, d_pred, mcweeny_iterations)
    t3 = time.perf_counter()
    t4 = t3
    ok = bool(np.all(np.isfinite(d_pur))) and abs(ne - mol.nelectron) < 0.5
    out, e_l, err_l = None, float("inf"), None
    if ok:
        c2 = rs.Ctrl.from_buffer_copy(ctrl)
        if guard == "energy":
            c2.guess_energy_max = e_s - band
        else:
            c2.guess_err_max = guess_err_max
        out = rs.rhf(r, S, H, enuc, mol.nelectron, d_pur, c2)
        e_l, err_l = out.guess_energy, out.guess_error
        if out.guess_rejected or not np.isfinite(e_l):
            out = None
            if guard == "energy" and np.isfinite(e_l) and e_l < e_s + band and gp != jk.PREC_FP64:
                # inconclusive: decide on E(D1) at FP64
                refined = True
                e_s, band = sad_e1(jk.PREC_FP64), band_l
                if e_l < e_s - band:
                    c2.guess_energy_max = e_s - band
                    out = rs.rhf(r, S, H, enuc, mol.nelectron, d_pur, c2)
                    if out.guess_rejected:
                        out = None
    kind = "learned" if out is not None else "sad"
    if out is None:   # SAD, with REST's guess step inside rscf_rhf
        out = rs.rhf(r, S, H, enuc, mol.nelectron, d_sad, ctrl)
    t5 = time.perf_counter()
    secs = {"sad": t1 - t0, "predict": t_pred, "sad_energy": t_e1, "predict_and_sad_energy": t2 - t1,
            "mcweeny": t3 - t2, "scf": t5 - t4}
    return LearnedScf(out, kind, e_l, e_s, ne, secs, err_l, band, refined)


# ---------------------------------------------------------------------------
# training data from this SCF
# ---------------------------------------------------------------------------
def training_cluster(i: int, base_seed: int = 4096, nmin: int = 2, nmax: int = 24):
    """Cluster i of the training set: n uniform in [nmin, nmax], monomers with O–H ±0.05 Å and H–O–H ±5°."""
    r = wl.Xoshiro256((base_seed ^ (i * 0x9E3779B97F4A7C15)) & ((1 << 64) - 1))
    n = nmin + int(r.uniform() * (nmax - nmin + 1))
    seed = r.next()
    return wl.water_cluster(seed, n, oh_jitter=0.05, hoh_jitter_deg=5.0)


def make_training_data_batched(indices, base_seed: int = 4096, basis="def2-SVP", auxbasis="def2-SVP-JKFIT",
                               log=print, stats: dict = None, prec: int = None, budget_bytes: int = 0,
                               chunk: int = 256, setup_threads: int = 16, deadline: float = None):
    """Converge the training clusters with the batched SCF (batch_scf.rhf_batch: REST's loop per
    cluster, every iteration's J/K of all clusters of a wave in one grouped launch, FP64 by default)
    and encode the converged densities.  Clusters go through in chunks of `chunk` (host memory for
    their one-electron matrices), each chunk in device-memory waves; no chunk starts after
    `deadline` (a time.time() value)."""
    from . import batch_scf as bs
    prec = jk.PREC_FP64 if prec is None else prec
    idx = list(indices)
    Xs, Ts, Ms = [], [], []
    st_all = {"clusters_done": 0, "not_converged": 0, "fallbacks": 0, "batch_runs": 0, "waves": 0,
              "jk_seconds": 0.0, "host_seconds": 0.0, "setup_seconds": 0.0}
    t0 = time.time()
    for c0 in range(0, len(idx), chunk):
        if deadline is not None and time.time() > deadline:
            log(f"deadline reached after {c0} clusters")
            break
        part = idx[c0:c0 + chunk]
        mols = [ri.Molecule.water_cluster(training_cluster(i, base_seed), basis, auxbasis) for i in part]
        st = {}
        outs = bs.rhf_batch(mols, prec=prec, budget_bytes=budget_bytes, setup_threads=setup_threads, stats=st,
                            log=log)
        for k, v in st.items():
            st_all[k] = st_all.get(k, 0) + v
        for i, mol, o in zip(part, mols, outs):
            if not o.converged:
                st_all["not_converged"] += 1
                log(f"cluster {i}: SCF did not converge, skipped")
                continue
            da, db = gs.sad_density(mol)
            X, T, M = encode(mol, o.dm, spherical_average(mol, da + db))
            Xs.append(X)
            Ts.append(T)
            Ms.append(M)
        st_all["clusters_done"] += len(part)
        log(f"  {st_all['clusters_done']}/{len(idx)} clusters, {sum(len(x) for x in Xs)} pairs "
            f"({time.time() - t0:.0f}s; J/K {st_all['jk_seconds']:.0f}s, host {st_all['host_seconds']:.0f}s, "
            f"setup {st_all['setup_seconds']:.0f}s)")
    if stats is not None:
        stats.update(st_all)
    if not Xs:
        return (np.zeros((0, N_IN), np.float32), np.zeros((0, MAXB, MAXB), np.float32),
                np.zeros((0, MAXB, MAXB), np.float32))
    return np.concatenate(Xs), np.concatenate(Ts), np.concatenate(Ms)


def make_training_data(indices, base_seed: int = 4096, basis="def2-SVP", auxbasis="def2-SVP-JKFIT", log=print,
                       deadline: float = None, stats: dict = None):
    """Run the FP64 SCF (SAD start) on each training cluster and encode its converged density.
    deadline (time.time() value): stop before the next cluster once it has passed; stats receives
    the number of clusters done."""
    Xs, Ts, Ms = [], [], []
    t0 = time.time()
    done = 0
    for c, i in enumerate(indices):
        if deadline is not None and time.time() > deadline:
            log(f"deadline reached after {c} clusters")
            break
        done = c + 1
        xyz = training_cluster(i, base_seed)
        mol = ri.Molecule.water_cluster(xyz, basis, auxbasis)
        out, _ = rs.run_molecule(mol, guess="sad")
        if not out.converged:
            log(f"cluster {i}: SCF did not converge, skipped")
            continue
        da, db = gs.sad_density(mol)
        X, T, M = encode(mol, out.dm, spherical_average(mol, da + db))
        Xs.append(X)
        Ts.append(T)
        Ms.append(M)
        if (c + 1) % 16 == 0:
            log(f"  {c + 1}/{len(indices)} clusters, {sum(len(x) for x in Xs)} pairs ({time.time() - t0:.0f}s)")
    if stats is not None:
        stats["clusters_done"] = done
    if not Xs:
        return (np.zeros((0, N_IN), np.float32), np.zeros((0, MAXB, MAXB), np.float32),
                np.zeros((0, MAXB, MAXB), np.float32))
    return np.concatenate(Xs), np.concatenate(Ts), np.concatenate(Ms)

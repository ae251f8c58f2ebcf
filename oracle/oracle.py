"""ctypes wrapper of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
CPU-baseline leg of bench.py, never by the rest_amd package.

Same array conventions as rest_amd.jk: matrices [row, col] (passed
column-major), packed B indexed [p, P] with shape (npair, naux), full B
indexed [i, j, P] with shape (nbf, nbf, naux), packed-upper outputs
p = j(j+1)/2 + i.
"""
from __future__ import annotations

import ctypes as C
import glob
import subprocess
from pathlib import Path
from typing import Sequence

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "liboracle.so"

_lib = None


def build() -> Path:
    subprocess.run(["make", "-s", "-B", "-C", str(HERE)], check=True)
    return LIB_PATH


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            build()
        L = C.CDLL(str(LIB_PATH))
        vp, sz, d, u64, i = C.c_void_p, C.c_size_t, C.c_double, C.c_uint64, C.c_int
        sig = {
            "orc_synth_b": (d, [u64, u64, u64]),
            "orc_fill_synth_b": (None, [u64, sz, sz, sz, sz, vp]),
            "orc_water_cluster": (i, [u64, i, d, d, vp]),
            "orc_random_orthonormal": (None, [u64, sz, sz, vp]),
            "orc_vj_upper_rimatr": (None, [vp, sz, sz, vp, sz, d, vp]),
            "orc_gamma_rimatr": (None, [vp, sz, sz, vp, vp]),
            "orc_vj_upper_ri3fn": (None, [vp, sz, sz, vp, sz, d, vp]),
            "orc_vk_upper_occ": (None, [vp, i, sz, sz, vp, sz, vp, vp, sz, d, vp]),
            "orc_vk_upper_dm": (None, [vp, i, sz, sz, vp, sz, d, vp]),
            "orc_ri3fn_from_rimatr": (None, [vp, sz, sz, vp]),
            "orc_density_matrix": (None, [vp, sz, sz, vp, vp, sz, vp]),
            "orc_fock_upper": (None, [vp, vp, vp, sz, sz, vp]),
            "orc_energy_contraction": (d, [vp, vp, sz]),
            "orc_blas_load": (i, [C.c_char_p, C.c_char_p, C.c_char_p]),
            "orc_cpu_baseline_vj": (d, [vp, sz, sz, vp, vp, i]),
            "orc_cpu_baseline_vk": (d, [vp, sz, sz, vp, sz, vp, vp, i]),
            "orc_cpu_baseline_vk_dm": (d, [vp, sz, sz, vp, vp, i]),
            "orc_max_threads": (i, []),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _p(a: np.ndarray):
    return C.c_void_p(a.ctypes.data)


def _cm(mats: Sequence[np.ndarray], ns: int, shape) -> np.ndarray:
    out = np.empty((ns,) + tuple(shape[::-1]))
    for s in range(ns):
        out[s] = np.asarray(mats[s], dtype=np.float64).T
    return out


def _rimatr_buf(b: np.ndarray) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(b, dtype=np.float64).T)  # [P][p]


def _ri3fn_buf(b: np.ndarray) -> np.ndarray:
    return np.ascontiguousarray(np.transpose(np.asarray(b, dtype=np.float64), (2, 1, 0)))  # [P][j][i]


# ---- generators -------------------------------------------------------------
def water_cluster(seed: int, n: int, oh_jitter: float = 0.0, hoh_jitter_deg: float = 0.0) -> np.ndarray:
    """(3n, 3) coordinates in Å, atoms O,H,H per water."""
    xyz = np.empty(9 * n)
    rc = lib().orc_water_cluster(seed, n, oh_jitter, hoh_jitter_deg, _p(xyz))
    if rc < 0:
        raise RuntimeError("water cluster generator exhausted its attempts")
    return xyz.reshape(3 * n, 3)


def random_orthonormal(seed: int, nbf: int, nocc: int) -> np.ndarray:
    c = np.empty(nbf * nocc)
    lib().orc_random_orthonormal(seed, nbf, nocc, _p(c))
    return c.reshape(nocc, nbf).T.copy()


def synth_rimatr(seed: int, nbf: int, naux: int, p_beg: int = 0, p_cnt: int | None = None) -> np.ndarray:
    """Synthetic packed B (same values as rjk_fill_rimatr_synthetic), shape (npair, p_cnt)."""
    npair = nbf * (nbf + 1) // 2
    p_cnt = naux - p_beg if p_cnt is None else p_cnt
    out = np.empty(p_cnt * npair)
    lib().orc_fill_synth_b(seed, npair, naux, p_beg, p_cnt, _p(out))
    return out.reshape(p_cnt, npair).T


# ---- builders ---------------------------------------------------------------
def vj_upper_rimatr(rimatr, dm, spin_channel: int, scale: float) -> np.ndarray:
    b = _rimatr_buf(rimatr)
    naux = b.shape[0]
    nbf = np.asarray(dm[0]).shape[0]
    d = _cm(dm, spin_channel, (nbf, nbf))
    out = np.empty((spin_channel, nbf * (nbf + 1) // 2))
    lib().orc_vj_upper_rimatr(_p(b), nbf, naux, _p(d), spin_channel, scale, _p(out))
    return out


def gamma_rimatr(rimatr, dm) -> np.ndarray:
    b = _rimatr_buf(rimatr)
    nbf = np.asarray(dm).shape[0]
    d = _cm([dm], 1, (nbf, nbf))
    g = np.empty(b.shape[0])
    lib().orc_gamma_rimatr(_p(b), nbf, b.shape[0], _p(d), _p(g))
    return g


def vj_upper_ri3fn(ri3fn, dm, spin_channel: int, scale: float) -> np.ndarray:
    b = _ri3fn_buf(ri3fn)
    naux, nbf = b.shape[0], b.shape[1]
    d = _cm(dm, spin_channel, (nbf, nbf))
    out = np.empty((spin_channel, nbf * (nbf + 1) // 2))
    lib().orc_vj_upper_ri3fn(_p(b), nbf, naux, _p(d), spin_channel, scale, _p(out))
    return out


def vk_upper_occ(b, packed: bool, eigv, homo, occupation, spin_channel: int, scale: float) -> np.ndarray:
    buf = _rimatr_buf(b) if packed else _ri3fn_buf(b)
    naux = buf.shape[0]
    nbf, nstate = np.asarray(eigv[0]).shape
    e = _cm(eigv, spin_channel, (nbf, nstate))
    occ = np.zeros((spin_channel, nstate))
    for s in range(spin_channel):
        occ[s] = np.asarray(occupation[s], dtype=np.float64)[:nstate]
    hm = np.array(list(homo)[:spin_channel] + [0] * (2 - spin_channel), dtype=np.uint64)
    out = np.empty((spin_channel, nbf * (nbf + 1) // 2))
    lib().orc_vk_upper_occ(_p(buf), int(packed), nbf, naux, _p(e), nstate, _p(hm), _p(occ), spin_channel, scale,
                           _p(out))
    return out


def vk_upper_dm(b, packed: bool, dm, spin_channel: int, scale: float) -> np.ndarray:
    buf = _rimatr_buf(b) if packed else _ri3fn_buf(b)
    naux = buf.shape[0]
    nbf = np.asarray(dm[0]).shape[0]
    d = _cm(dm, spin_channel, (nbf, nbf))
    out = np.empty((spin_channel, nbf * (nbf + 1) // 2))
    lib().orc_vk_upper_dm(_p(buf), int(packed), nbf, naux, _p(d), spin_channel, scale, _p(out))
    return out


def ri3fn_from_rimatr(rimatr, nbf: int) -> np.ndarray:
    b = _rimatr_buf(rimatr)
    naux = b.shape[0]
    out = np.empty(naux * nbf * nbf)
    lib().orc_ri3fn_from_rimatr(_p(b), nbf, naux, _p(out))
    return np.transpose(out.reshape(naux, nbf, nbf), (2, 1, 0))


def density_matrix(eigv, homo, occupation, spin_channel: int) -> list:
    nbf, nstate = np.asarray(eigv[0]).shape
    e = _cm(eigv, spin_channel, (nbf, nstate))
    occ = np.zeros((spin_channel, nstate))
    for s in range(spin_channel):
        occ[s] = np.asarray(occupation[s], dtype=np.float64)[:nstate]
    hm = np.array(list(homo)[:spin_channel] + [0] * (2 - spin_channel), dtype=np.uint64)
    out = np.empty(spin_channel * nbf * nbf)
    lib().orc_density_matrix(_p(e), nbf, nstate, _p(hm), _p(occ), spin_channel, _p(out))
    return [out[s * nbf * nbf:(s + 1) * nbf * nbf].reshape(nbf, nbf).T.copy() for s in range(spin_channel)]


def energy_contraction(dm: np.ndarray, a_upper: np.ndarray) -> float:
    nbf = dm.shape[0]
    d = _cm([dm], 1, (nbf, nbf))
    a = np.ascontiguousarray(a_upper, dtype=np.float64)
    return lib().orc_energy_contraction(_p(d), _p(a), nbf)


# ---- CPU baseline -----------------------------------------------------------
def load_host_blas() -> str:
    """Resolve numpy's bundled OpenBLAS (ILP64 `scipy_` symbols); '' if absent."""
    import numpy
    cands = sorted(glob.glob(str(Path(numpy.__file__).parent.parent / "numpy.libs" / "libscipy_openblas64_*.so")))
    for c in cands:
        if lib().orc_blas_load(c.encode(), b"scipy_dgemm_64_", b"scipy_openblas_set_num_threads64_"):
            return Path(c).name
    return ""


def cpu_baseline_vj(rimatr_cols: np.ndarray, nbf: int, dm: np.ndarray, threads: int = 0):
    """Reference-structured J over the given packed columns; returns (seconds, vj)."""
    b = _rimatr_buf(rimatr_cols)
    d = _cm([dm], 1, (nbf, nbf))
    out = np.empty(nbf * (nbf + 1) // 2)
    t = lib().orc_cpu_baseline_vj(_p(b), nbf, b.shape[0], _p(d), _p(out), threads)
    return t, out


def cpu_baseline_vk(rimatr_cols: np.ndarray, nbf: int, c: np.ndarray, occ: np.ndarray, threads: int = 0):
    b = _rimatr_buf(rimatr_cols)
    cc = np.ascontiguousarray(np.asarray(c, dtype=np.float64).T)
    nw = cc.shape[0]
    o = np.ascontiguousarray(occ, dtype=np.float64)
    out = np.empty(nbf * (nbf + 1) // 2)
    t = lib().orc_cpu_baseline_vk(_p(b), nbf, b.shape[0], _p(cc), nw, _p(o), _p(out), threads)
    return t, out


def cpu_baseline_vk_dm(rimatr_cols: np.ndarray, nbf: int, dm: np.ndarray, threads: int = 0):
    """Reference-structured density-only K (REST's use_dm_only route) over the given packed columns;
    returns (seconds, vk)."""
    b = _rimatr_buf(rimatr_cols)
    d = _cm([dm], 1, (nbf, nbf))
    out = np.empty(nbf * (nbf + 1) // 2)
    t = lib().orc_cpu_baseline_vk_dm(_p(b), nbf, b.shape[0], _p(d), _p(out), threads)
    return t, out


def fock_upper(hcore_upper: np.ndarray, vj: np.ndarray, vk: np.ndarray, spin_channel: int) -> np.ndarray:
    """F_s = H + J_0 (+ J_1) + K_s, all packed upper; vj/vk shaped (spin_channel, npair)."""
    h = np.ascontiguousarray(hcore_upper, dtype=np.float64)
    j = np.ascontiguousarray(vj, dtype=np.float64)
    k = np.ascontiguousarray(vk, dtype=np.float64)
    npair = h.shape[0]
    out = np.empty((spin_channel, npair))
    lib().orc_fock_upper(_p(h), _p(j), _p(k), npair, spin_channel, _p(out))
    return out

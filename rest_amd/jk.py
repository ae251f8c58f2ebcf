"""Python host binding of the MI355X RI-V J/K library (rest_amd/lib/librest_jk.so).

The module-level functions mirror REST's vj/vk builders (src/scf_io/mod.rs)
with the same names, argument meaning and return shape:

    vj_upper_with_rimatr_v_sync(ri3fn, dm, spin_channel, scaling_factor)        # :2490
    vj_upper_with_ri_v_sync(ri3fn, dm, spin_channel, scaling_factor)            # :2428
    vk_upper_with_ri_v_sync(ri3fn, eigv, homo, occupation, spin_channel,
                            scaling_factor)                                     # :2758
    vk_upper_with_rimatr_use_dm_only_sync(ri3fn, dm, spin_channel, scaling_factor)  # :2703
    vk_upper_with_ri_v_use_dm_only_sync(ri3fn, dm, spin_channel, scaling_factor)    # :2640

`ri3fn` is a `DeviceRI` (B resident in HBM) or None; like REST, None yields
two 1-element zero vectors.  Every builder returns a list of two packed-upper
vectors (p = j(j+1)/2 + i, i <= j); when spin_channel == 1 the second is
REST's 1-element dummy.

Array conventions: matrices are numpy 2-D arrays indexed [row, col] (they are
handed to the library column-major, like REST's MatrixFull); packed B
("rimatr") is indexed [p, P] with shape (npair, naux); full B ("ri3fn") is
indexed [i, j, P] with shape (nbf, nbf, naux).

There is no CPU fallback: if the native library is absent every call raises.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path
from typing import Optional, Sequence

import numpy as np

LIB_PATH = Path(__file__).resolve().parent / "lib" / "librest_jk.so"

RJK_OK = 0
PREC_FP64, PREC_FP32, PREC_BF16X2, PREC_BF16X3, PREC_BF16X1 = 0, 1, 2, 3, 4
NUM_PREC = 5
STORE_FP64, STORE_FP32, STORE_BF16 = 1, 2, 4
PREC_NAMES = {"fp64": PREC_FP64, "fp32": PREC_FP32, "bf16x2": PREC_BF16X2, "bf16x3": PREC_BF16X3,
              "bf16x1": PREC_BF16X1}
ERRORS = {1: "ARG", 2: "HIP", 3: "NOMEM", 4: "STATE", 5: "NONFINITE", 6: "NOSOLVER"}

# every symbol include/rest_jk.h declares (checked by tests/test_abi.py)
EXPORTS = [
    "rjk_create", "rjk_destroy", "rjk_last_error", "rjk_num_basis", "rjk_num_auxbas", "rjk_num_baspar",
    "rjk_ld_baspar", "rjk_set_stream", "rjk_synchronize", "rjk_set_workspace_limit", "rjk_set_rimatr",
    "rjk_set_rimatr_dev", "rjk_set_ri3fn", "rjk_fill_rimatr_synthetic", "rjk_rimatr_dev", "rjk_get_rimatr",
    "rjk_vj_upper_with_rimatr_v_sync", "rjk_vj_upper_with_ri_v_sync", "rjk_vj_upper_dev", "rjk_vj_upper_dev_noalloc",
    "rjk_vk_upper_with_ri_v_sync", "rjk_vk_upper_occ_dev", "rjk_vk_upper_with_rimatr_use_dm_only_sync",
    "rjk_vk_upper_with_ri_v_use_dm_only_sync", "rjk_vk_upper_dm_dev", "rjk_profile_enable", "rjk_profile_read", "rjk_profile_read_detail", "rjk_copy_dev",
    # include/rest_ri.h
    "rsi_basis_create", "rsi_basis_destroy", "rsi_basis_nbf", "rsi_last_error", "rsi_int1e", "rsi_int2c2e",
    "rsi_int3c2e_packed", "rsi_int3c2e_packed_dev", "rsi_build_rimatr", "rsi_cart2sph",
    # include/rest_scf.h
    "rscf_default_ctrl", "rscf_rhf", "rscf_rhf_orbitals", "rscf_uhf", "rscf_guess_step", "rscf_guess_step_factored",
    "rscf_energy", "rscf_mcweeny",
    # include/rest_batch.h
    "rjk_batch_create", "rjk_batch_destroy", "rjk_batch_last_error", "rjk_batch_size", "rjk_batch_bytes_b",
    "rjk_batch_set_stream", "rjk_batch_set_rimatr", "rjk_batch_fill_synthetic", "rjk_batch_set_inputs",
    "rjk_batch_get_outputs", "rjk_batch_run", "rjk_batch_synchronize", "rjk_batch_autotune", "rjk_batch_config",
    "rjk_batch_set_config", "rjk_batch_device_bytes", "rjk_batch_plan_waves", "rjk_batch_set_rimatr_dev",
    "rjk_batch_profile", "rjk_batch_set_active", "rjk_batch_device_memory",
    # include/rest_lg.h
    "rlg_create", "rlg_destroy", "rlg_last_error", "rlg_num_params", "rlg_set_stream", "rlg_set_params",
    "rlg_get_params", "rlg_forward", "rlg_loss_grad", "rlg_adam_step", "rlg_reset_optimizer", "rlg_gemm_bf16",
]


class RJKError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"rjk error {code} ({ERRORS.get(code, '?')}): {msg}")
        self.code = code


_lib: Optional[C.CDLL] = None


def load_library(path: Optional[Path] = None) -> C.CDLL:
    """Load librest_jk.so (raises if it was not built)."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = Path(path) if path else LIB_PATH
    if not p.exists():
        raise RuntimeError(f"native J/K library not built: {p} (run `python -m rest_amd.build`)")
    # PyTorch wheels bundle their own libamdhip64.so.7 / libhsa-runtime64.so.1.
    # Loading torch first makes this library bind to those same copies (the
    # loader matches by SONAME), so one process never holds two HIP runtimes.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    lib = C.CDLL(str(p))
    vp, sz, d, pd, i = C.c_void_p, C.c_size_t, C.c_double, C.POINTER(C.c_double), C.c_int
    sig = {
        "rjk_create": (i, [i, sz, sz, C.POINTER(vp)]),
        "rjk_destroy": (i, [vp]),
        "rjk_last_error": (C.c_char_p, [vp]),
        "rjk_num_basis": (sz, [vp]),
        "rjk_num_auxbas": (sz, [vp]),
        "rjk_num_baspar": (sz, [vp]),
        "rjk_ld_baspar": (sz, [vp]),
        "rjk_set_stream": (i, [vp, vp]),
        "rjk_synchronize": (i, [vp]),
        "rjk_set_workspace_limit": (i, [vp, sz]),
        "rjk_set_rimatr": (i, [vp, vp, C.c_uint]),
        "rjk_set_rimatr_dev": (i, [vp, vp, C.c_uint]),
        "rjk_set_ri3fn": (i, [vp, vp, C.c_uint]),
        "rjk_fill_rimatr_synthetic": (i, [vp, C.c_uint64, C.c_uint]),
        "rjk_rimatr_dev": (vp, [vp]),
        "rjk_get_rimatr": (i, [vp, vp]),
        "rjk_vj_upper_with_rimatr_v_sync": (i, [vp, vp, sz, d, i, vp]),
        "rjk_vj_upper_with_ri_v_sync": (i, [vp, vp, sz, d, i, vp]),
        "rjk_vj_upper_dev": (i, [vp, vp, sz, d, i, i, vp]),
        "rjk_vj_upper_dev_noalloc": (i, [vp, vp, sz, d, i, i, vp]),
        "rjk_vk_upper_with_ri_v_sync": (i, [vp, vp, sz, vp, vp, sz, d, i, vp]),
        "rjk_vk_upper_occ_dev": (i, [vp, vp, sz, vp, vp, sz, d, i, vp]),
        "rjk_vk_upper_with_rimatr_use_dm_only_sync": (i, [vp, vp, sz, d, i, vp]),
        "rjk_vk_upper_with_ri_v_use_dm_only_sync": (i, [vp, vp, sz, d, i, vp]),
        "rjk_vk_upper_dm_dev": (i, [vp, vp, sz, d, i, vp]),
        "rjk_profile_enable": (i, [vp, i]),
        "rjk_profile_read": (i, [vp, pd, C.POINTER(C.c_uint64)]),
        "rjk_profile_read_detail": (i, [vp, pd, C.POINTER(C.c_uint64), pd]),
        "rjk_copy_dev": (i, [vp, vp, vp, sz]),
        "rsi_basis_create": (i, [sz, vp, vp, vp, vp, vp, sz, vp, C.POINTER(vp)]),
        "rsi_basis_destroy": (i, [vp]),
        "rsi_basis_nbf": (sz, [vp]),
        "rsi_last_error": (C.c_char_p, []),
        "rsi_int1e": (i, [vp, sz, vp, vp, vp, vp, vp]),
        "rsi_int2c2e": (i, [vp, vp]),
        "rsi_int3c2e_packed": (i, [vp, vp, sz, sz, vp]),
        "rsi_int3c2e_packed_dev": (i, [vp, vp, vp, sz, sz, vp]),
        "rsi_cart2sph": (i, [i, vp]),
        "rsi_build_rimatr": (i, [vp, vp, vp, d, C.c_uint, C.POINTER(sz), pd]),
        "rjk_batch_create": (i, [i, sz, vp, vp, vp, i, C.POINTER(vp)]),
        "rjk_batch_destroy": (i, [vp]),
        "rjk_batch_last_error": (C.c_char_p, [vp]),
        "rjk_batch_size": (sz, [vp]),
        "rjk_batch_bytes_b": (sz, [vp]),
        "rjk_batch_set_stream": (i, [vp, vp]),
        "rjk_batch_set_rimatr": (i, [vp, sz, vp]),
        "rjk_batch_fill_synthetic": (i, [vp, C.c_uint64]),
        "rjk_batch_set_inputs": (i, [vp, sz, vp, vp, vp]),
        "rjk_batch_get_outputs": (i, [vp, sz, vp, vp]),
        "rjk_batch_run": (i, [vp, d, d]),
        "rjk_batch_synchronize": (i, [vp]),
        "rjk_batch_autotune": (i, [vp, i, pd]),
        "rjk_batch_config": (i, [vp, C.POINTER(i), C.POINTER(i)]),
        "rjk_batch_set_config": (i, [vp, i, i]),
        "rjk_batch_device_bytes": (sz, [sz, vp, vp, vp, i]),
        "rjk_batch_plan_waves": (i, [sz, vp, vp, vp, i, sz, vp, C.POINTER(sz)]),
        "rjk_batch_set_rimatr_dev": (i, [vp, sz, vp, sz]),
        "rjk_batch_profile": (i, [vp, d, d, pd]),
        "rjk_batch_set_active": (i, [vp, sz, i]),
        "rjk_batch_device_memory": (i, [i, C.POINTER(sz), C.POINTER(sz)]),
        "rlg_create": (i, [i, sz, sz, sz, sz, sz, C.POINTER(vp)]),
        "rlg_destroy": (i, [vp]),
        "rlg_last_error": (C.c_char_p, [vp]),
        "rlg_num_params": (sz, [vp]),
        "rlg_set_stream": (i, [vp, vp]),
        "rlg_set_params": (i, [vp, vp]),
        "rlg_get_params": (i, [vp, vp]),
        "rlg_forward": (i, [vp, vp, sz, vp]),
        "rlg_loss_grad": (i, [vp, vp, vp, vp, sz, pd, vp]),
        "rlg_adam_step": (i, [vp, d, d, d, d]),
        "rlg_reset_optimizer": (i, [vp]),
        "rlg_gemm_bf16": (i, [vp, vp, C.c_long, C.c_long, vp, C.c_long, C.c_long, vp, C.c_long, i, i, i]),
    }
    for name, (res, args) in sig.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    if path is None:
        _lib = lib
    return lib


def _ptr(a: np.ndarray) -> C.c_void_p:
    return C.c_void_p(a.ctypes.data)


def _colmajor_stack(mats: Sequence[np.ndarray], spin_channel: int, shape) -> np.ndarray:
    """Stack spin matrices as contiguous column-major blocks."""
    out = np.empty((spin_channel,) + tuple(shape[::-1]), dtype=np.float64)
    for s in range(spin_channel):
        m = np.asarray(mats[s], dtype=np.float64)
        if m.shape != tuple(shape):
            raise ValueError(f"matrix of spin {s} has shape {m.shape}, expected {tuple(shape)}")
        out[s] = m.T  # row-major of the transpose == column-major of m
    return out


class DeviceRI:
    """B[μν,P] resident in HBM (packed pair layout), one context per device."""

    PHASES = ("j_gamma", "j_acc", "k_half", "k_syrk", "other")

    def __init__(self, num_basis: int, num_auxbas: int, device: int = 0):
        self.lib = load_library()
        h = C.c_void_p()
        rc = self.lib.rjk_create(int(device), int(num_basis), int(num_auxbas), C.byref(h))
        if rc != RJK_OK:
            raise RJKError(rc, self.lib.rjk_last_error(None).decode())
        self.h = h
        self.nbf = int(num_basis)
        self.naux = int(num_auxbas)
        self.npair = self.nbf * (self.nbf + 1) // 2
        self.ldb = int(self.lib.rjk_ld_baspar(h))

    # -- lifecycle -----------------------------------------------------------
    def close(self):
        if getattr(self, "h", None):
            self.lib.rjk_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def _check(self, rc: int):
        if rc != RJK_OK:
            raise RJKError(rc, self.lib.rjk_last_error(self.h).decode())

    # -- B storage -------------------------------------------------------------
    def set_rimatr(self, rimatr: np.ndarray, store: int = STORE_FP64):
        """rimatr indexed [p, P], shape (npair, naux)."""
        b = np.asarray(rimatr, dtype=np.float64)
        if b.shape != (self.npair, self.naux):
            raise ValueError(f"rimatr shape {b.shape}, expected {(self.npair, self.naux)}")
        buf = np.ascontiguousarray(b.T)  # [P][p] == column-major [p, P]
        self._check(self.lib.rjk_set_rimatr(self.h, _ptr(buf), store))

    def set_ri3fn(self, ri3fn: np.ndarray, store: int = STORE_FP64):
        """ri3fn indexed [i, j, P], shape (nbf, nbf, naux)."""
        b = np.asarray(ri3fn, dtype=np.float64)
        if b.shape != (self.nbf, self.nbf, self.naux):
            raise ValueError(f"ri3fn shape {b.shape}, expected {(self.nbf, self.nbf, self.naux)}")
        buf = np.ascontiguousarray(np.transpose(b, (2, 1, 0)))  # [P][j][i]
        self._check(self.lib.rjk_set_ri3fn(self.h, _ptr(buf), store))

    def fill_synthetic(self, seed: int, store: int = STORE_FP64):
        self._check(self.lib.rjk_fill_rimatr_synthetic(self.h, C.c_uint64(seed), store))

    def set_rimatr_dev(self, ptr: int, store: int = STORE_FP64):
        self._check(self.lib.rjk_set_rimatr_dev(self.h, C.c_void_p(ptr), store))

    def get_rimatr(self) -> np.ndarray:
        """Stored FP64 B as a host array indexed [p, P] (REST's rimatr)."""
        buf = np.empty((self.naux, self.npair))
        self._check(self.lib.rjk_get_rimatr(self.h, _ptr(buf)))
        return buf.T

    def rimatr_dev_ptr(self) -> int:
        return int(self.lib.rjk_rimatr_dev(self.h) or 0)

    def set_stream(self, stream_ptr: Optional[int]):
        self._check(self.lib.rjk_set_stream(self.h, C.c_void_p(stream_ptr or 0)))

    def synchronize(self):
        self._check(self.lib.rjk_synchronize(self.h))

    def set_workspace_limit(self, nbytes: int):
        self._check(self.lib.rjk_set_workspace_limit(self.h, int(nbytes)))

    # -- builders (host arrays) -----------------------------------------------
    def vj(self, dm: Sequence[np.ndarray], spin_channel: int, scaling_factor: float = 1.0,
           prec: int = PREC_FP64, full_dm: bool = False) -> np.ndarray:
        d = _colmajor_stack(dm, spin_channel, (self.nbf, self.nbf))
        out = np.empty((spin_channel, self.npair))
        f = self.lib.rjk_vj_upper_with_ri_v_sync if full_dm else self.lib.rjk_vj_upper_with_rimatr_v_sync
        self._check(f(self.h, _ptr(d), spin_channel, float(scaling_factor), prec, _ptr(out)))
        return out

    def vk_occ(self, eigv: Sequence[np.ndarray], homo: Sequence[int], occupation: Sequence[np.ndarray],
               spin_channel: int, scaling_factor: float = 1.0, prec: int = PREC_FP64) -> np.ndarray:
        nstate = np.asarray(eigv[0]).shape[1]
        e = _colmajor_stack(eigv, spin_channel, (self.nbf, nstate))
        occ = np.zeros((spin_channel, nstate))
        for s in range(spin_channel):
            occ[s, :] = np.asarray(occupation[s], dtype=np.float64)[:nstate]
        hm = (C.c_size_t * 2)(*(list(homo)[:spin_channel] + [0] * (2 - spin_channel)))
        out = np.empty((spin_channel, self.npair))
        self._check(self.lib.rjk_vk_upper_with_ri_v_sync(self.h, _ptr(e), nstate, C.cast(hm, C.c_void_p),
                                                         _ptr(occ), spin_channel, float(scaling_factor), prec,
                                                         _ptr(out)))
        return out

    def vk_dm(self, dm: Sequence[np.ndarray], spin_channel: int, scaling_factor: float = 1.0,
              prec: int = PREC_FP64) -> np.ndarray:
        d = _colmajor_stack(dm, spin_channel, (self.nbf, self.nbf))
        out = np.empty((spin_channel, self.npair))
        self._check(self.lib.rjk_vk_upper_with_rimatr_use_dm_only_sync(self.h, _ptr(d), spin_channel,
                                                                       float(scaling_factor), prec, _ptr(out)))
        return out

    # -- builders (device pointers, stream-ordered) -----------------------------
    def vj_dev(self, dm_ptr: int, spin_channel: int, scaling_factor: float, out_ptr: int,
               prec: int = PREC_FP64, full_dm: bool = False):
        self._check(self.lib.rjk_vj_upper_dev(self.h, C.c_void_p(dm_ptr), spin_channel, float(scaling_factor),
                                              prec, int(full_dm), C.c_void_p(out_ptr)))

    def vj_dev_noalloc(self, dm_ptr: int, spin_channel: int, scaling_factor: float, out_ptr: int,
                       prec: int = PREC_FP64, full_dm: bool = False):
        """rjk_vj_upper_dev_noalloc: vj_dev on the library's fixed device scratch (no allocation)."""
        self._check(self.lib.rjk_vj_upper_dev_noalloc(self.h, C.c_void_p(dm_ptr), spin_channel,
                                                      float(scaling_factor), prec, int(full_dm),
                                                      C.c_void_p(out_ptr)))

    def vk_occ_dev(self, eigv_ptr: int, num_state: int, homo: Sequence[int], occ_ptr: int, spin_channel: int,
                   scaling_factor: float, out_ptr: int, prec: int = PREC_FP64):
        hm = (C.c_size_t * 2)(*(list(homo)[:spin_channel] + [0] * (2 - spin_channel)))
        self._check(self.lib.rjk_vk_upper_occ_dev(self.h, C.c_void_p(eigv_ptr), num_state, C.cast(hm, C.c_void_p),
                                                  C.c_void_p(occ_ptr), spin_channel, float(scaling_factor), prec,
                                                  C.c_void_p(out_ptr)))

    def vk_dm_dev(self, dm_ptr: int, spin_channel: int, scaling_factor: float, out_ptr: int,
                  prec: int = PREC_FP64):
        self._check(self.lib.rjk_vk_upper_dm_dev(self.h, C.c_void_p(dm_ptr), spin_channel, float(scaling_factor),
                                                 prec, C.c_void_p(out_ptr)))

    # -- measurement -------------------------------------------------------------
    def profile_enable(self, on: bool = True):
        self._check(self.lib.rjk_profile_enable(self.h, int(on)))

    def profile_read(self) -> dict:
        ms = (C.c_double * 5)()
        n = (C.c_uint64 * 5)()
        self._check(self.lib.rjk_profile_read(self.h, ms, n))
        return {name: (ms[k], int(n[k])) for k, name in enumerate(self.PHASES)}

    def profile_read_detail(self) -> dict:
        """{(prec, phase): (ms, launches, algorithmic work)} with prec "fp64" / "fp32" / "bf16x2" / "bf16x3" /
        "bf16x1"."""
        ms = (C.c_double * (5 * NUM_PREC))()
        n = (C.c_uint64 * (5 * NUM_PREC))()
        w = (C.c_double * (5 * NUM_PREC))()
        self._check(self.lib.rjk_profile_read_detail(self.h, ms, n, w))
        return {(p, name): (ms[5 * pi + k], int(n[5 * pi + k]), w[5 * pi + k])
                for pi, p in enumerate(("fp64", "fp32", "bf16x2", "bf16x3", "bf16x1"))
                for k, name in enumerate(self.PHASES)}

    def copy_dev(self, dst_ptr: int, src_ptr: int, nbytes: int):
        self._check(self.lib.rjk_copy_dev(self.h, C.c_void_p(dst_ptr), C.c_void_p(src_ptr), int(nbytes)))


BATCH_NUM_STAGES = 9
BATCH_STAGES = ("weights", "gamma", "j_acc", "prep", "k_half_narrow", "k_half_wide", "k_syrk_narrow", "k_syrk_wide",
                "k_reduce")


def batch_device_bytes(nbf, naux, nocc, prec: int) -> int:
    """Device bytes a Batch of these problems allocates (rjk_batch_device_bytes)."""
    lib = load_library()
    n = len(nbf)
    arr = lambda v: (C.c_size_t * n)(*[int(x) for x in v])
    return int(lib.rjk_batch_device_bytes(n, arr(nbf), arr(naux), arr(nocc), int(prec)))


def device_memory(device: int = 0) -> tuple:
    """(free, total) bytes of device memory now."""
    lib = load_library()
    fr, tot = C.c_size_t(0), C.c_size_t(0)
    rc = lib.rjk_batch_device_memory(int(device), C.byref(fr), C.byref(tot))
    if rc != RJK_OK:
        raise RJKError(rc, "rjk_batch_device_memory failed")
    return fr.value, tot.value


def batch_plan_waves(nbf, naux, nocc, prec: int, budget_bytes: int = 0) -> list:
    """Consecutive index ranges [(start, stop), ...] whose batches fit budget_bytes of device memory
    (rjk_batch_plan_waves; 0 = 90% of the free device memory)."""
    lib = load_library()
    n = len(nbf)
    arr = lambda v: (C.c_size_t * n)(*[int(x) for x in v])
    starts = (C.c_size_t * (n + 1))()
    nw = C.c_size_t(0)
    rc = lib.rjk_batch_plan_waves(n, arr(nbf), arr(naux), arr(nocc), int(prec), int(budget_bytes), starts,
                                  C.byref(nw))
    if rc != RJK_OK:
        raise RJKError(rc, "rjk_batch_plan_waves failed")
    return [(int(starts[w]), int(starts[w + 1])) for w in range(nw.value)]


class Batch:
    """J/K of many problems in grouped persistent launches replayed from a hipGraph (include/rest_batch.h)."""

    def __init__(self, nbf, naux, nocc, prec: int = PREC_FP64, device: int = 0):
        self.lib = load_library()
        self.nbf = [int(x) for x in nbf]
        self.naux = [int(x) for x in naux]
        self.nocc = [int(x) for x in nocc]
        n = len(self.nbf)
        arr = lambda v: (C.c_size_t * n)(*v)
        h = C.c_void_p()
        rc = self.lib.rjk_batch_create(int(device), n, arr(self.nbf), arr(self.naux), arr(self.nocc), int(prec),
                                       C.byref(h))
        if rc != RJK_OK:
            raise RJKError(rc, "rjk_batch_create failed")
        self.h = h
        self.prec = prec

    def _check(self, rc):
        if rc != RJK_OK:
            raise RJKError(rc, self.lib.rjk_batch_last_error(self.h).decode())

    def close(self):
        if getattr(self, "h", None):
            self.lib.rjk_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    @property
    def bytes_b(self) -> int:
        return int(self.lib.rjk_batch_bytes_b(self.h))

    def set_stream(self, stream_ptr: int):
        self._check(self.lib.rjk_batch_set_stream(self.h, C.c_void_p(stream_ptr)))

    def set_rimatr(self, i: int, rimatr: np.ndarray):
        """rimatr indexed [p, P] (npair, naux), as DeviceRI.set_rimatr."""
        buf = np.ascontiguousarray(np.asarray(rimatr, dtype=np.float64).T)
        self._check(self.lib.rjk_batch_set_rimatr(self.h, i, _ptr(buf)))

    def fill_synthetic(self, seed: int):
        self._check(self.lib.rjk_batch_fill_synthetic(self.h, C.c_uint64(seed)))

    def set_rimatr_dev(self, i: int, ptr: int, ld: int):
        """B of problem i from a device array [naux][ld] (padded rimatr layout)."""
        self._check(self.lib.rjk_batch_set_rimatr_dev(self.h, i, C.c_void_p(ptr), int(ld)))

    def set_from_ri(self, i: int, ri: "DeviceRI"):
        """B of problem i copied on the device from a DeviceRI's FP64 storage (same nbf and naux)."""
        if ri.nbf != self.nbf[i] or ri.naux != self.naux[i]:
            raise ValueError(f"set_from_ri: DeviceRI is {ri.nbf} x {ri.naux}, problem {i} is "
                             f"{self.nbf[i]} x {self.naux[i]}")
        self.set_rimatr_dev(i, ri.rimatr_dev_ptr(), ri.ldb)

    def set_active(self, i: int, active: bool):
        """Include problem i in later runs or skip it (its outputs keep their last values)."""
        self._check(self.lib.rjk_batch_set_active(self.h, i, 1 if active else 0))

    def profile(self, scale_j: float = 1.0, scale_k: float = 1.0) -> list:
        """Milliseconds of each of the BATCH_NUM_STAGES stages of one eager J/K of every problem."""
        ms = (C.c_double * BATCH_NUM_STAGES)()
        self._check(self.lib.rjk_batch_profile(self.h, float(scale_j), float(scale_k), ms))
        return list(ms)

    def set_inputs(self, i: int, dm: np.ndarray, eigv: np.ndarray, occ: np.ndarray):
        d = np.ascontiguousarray(np.asarray(dm, dtype=np.float64).T)
        c = np.ascontiguousarray(np.asarray(eigv, dtype=np.float64)[:, :self.nocc[i]].T)
        o = np.ascontiguousarray(np.asarray(occ, dtype=np.float64)[:self.nocc[i]])
        self._check(self.lib.rjk_batch_set_inputs(self.h, i, _ptr(d), _ptr(c), _ptr(o)))

    def get_outputs(self, i: int):
        n = self.nbf[i] * (self.nbf[i] + 1) // 2
        vj, vk = np.empty(n), np.empty(n)
        self._check(self.lib.rjk_batch_get_outputs(self.h, i, _ptr(vj), _ptr(vk)))
        return vj, vk

    def run(self, scale_j: float = 1.0, scale_k: float = 1.0):
        self._check(self.lib.rjk_batch_run(self.h, float(scale_j), float(scale_k)))

    def synchronize(self):
        self._check(self.lib.rjk_batch_synchronize(self.h))

    def autotune(self, reps: int = 2) -> float:
        ms = C.c_double(0.0)
        self._check(self.lib.rjk_batch_autotune(self.h, int(reps), C.byref(ms)))
        return ms.value

    def config(self):
        cut, w = C.c_int(0), C.c_int(0)
        self._check(self.lib.rjk_batch_config(self.h, C.byref(cut), C.byref(w)))
        return cut.value, w.value

    def set_config(self, khalf_cut: int, waves_per_cu: int):
        self._check(self.lib.rjk_batch_set_config(self.h, int(khalf_cut), int(waves_per_cu)))


# ---------------------------------------------------------------------------
# REST-shaped builders (src/scf_io/mod.rs)
# ---------------------------------------------------------------------------
def _dummy_pair(spin_channel: int, vals: np.ndarray) -> list:
    out = [np.zeros(1), np.zeros(1)]
    for s in range(spin_channel):
        out[s] = vals[s].copy()
    return out


def vj_upper_with_rimatr_v_sync(ri3fn: Optional[DeviceRI], dm, spin_channel: int, scaling_factor: float,
                                prec: int = PREC_FP64) -> list:
    """src/scf_io/mod.rs:2490 — packed B, upper triangle of D."""
    if ri3fn is None:
        return [np.zeros(1), np.zeros(1)]
    return _dummy_pair(spin_channel, ri3fn.vj(dm, spin_channel, scaling_factor, prec, full_dm=False))


def vj_upper_with_ri_v_sync(ri3fn: Optional[DeviceRI], dm, spin_channel: int, scaling_factor: float,
                            prec: int = PREC_FP64) -> list:
    """src/scf_io/mod.rs:2428 — full B, γ over the full D."""
    if ri3fn is None:
        return [np.zeros(1), np.zeros(1)]
    return _dummy_pair(spin_channel, ri3fn.vj(dm, spin_channel, scaling_factor, prec, full_dm=True))


def vk_upper_with_ri_v_sync(ri3fn: Optional[DeviceRI], eigv, homo, occupation, spin_channel: int,
                            scaling_factor: float, prec: int = PREC_FP64) -> list:
    """src/scf_io/mod.rs:2758 — occupied-orbital route."""
    if ri3fn is None:
        return [np.zeros(1), np.zeros(1)]
    return _dummy_pair(spin_channel, ri3fn.vk_occ(eigv, homo, occupation, spin_channel, scaling_factor, prec))


def vk_upper_with_rimatr_use_dm_only_sync(ri3fn: Optional[DeviceRI], dm, spin_channel: int, scaling_factor: float,
                                          prec: int = PREC_FP64) -> list:
    """src/scf_io/mod.rs:2703 — density-only route on packed B."""
    if ri3fn is None:
        return [np.zeros(1), np.zeros(1)]
    return _dummy_pair(spin_channel, ri3fn.vk_dm(dm, spin_channel, scaling_factor, prec))


def vk_upper_with_ri_v_use_dm_only_sync(ri3fn: Optional[DeviceRI], dm, spin_channel: int, scaling_factor: float,
                                        prec: int = PREC_FP64) -> list:
    """src/scf_io/mod.rs:2640 — density-only route on full B (stored through its upper triangle)."""
    return vk_upper_with_rimatr_use_dm_only_sync(ri3fn, dm, spin_channel, scaling_factor, prec)


# ---------------------------------------------------------------------------
# packed-index helpers
# ---------------------------------------------------------------------------
def upper_indices(nbf: int):
    """(i, j) arrays of the packed order p = j(j+1)/2 + i, i <= j."""
    j = np.concatenate([np.full(k + 1, k) for k in range(nbf)])
    i = np.concatenate([np.arange(k + 1) for k in range(nbf)])
    return i, j


def unpack_upper(v: np.ndarray, nbf: int) -> np.ndarray:
    """packed upper vector -> symmetric full matrix"""
    i, j = upper_indices(nbf)
    m = np.zeros((nbf, nbf))
    m[i, j] = v
    m[j, i] = v
    return m

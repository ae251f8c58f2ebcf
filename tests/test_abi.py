"""CPU checks of the C-ABI library: it loads, exports every entry point that
include/*.h declares, and its host-side argument checks work without a
GPU (no compute call is made here)."""
import ctypes as C
import re
from pathlib import Path

import pytest

from rest_amd import jk

ROOT = Path(__file__).resolve().parents[1]


def header_symbols():
    syms = set()
    for h in sorted((ROOT / "include").glob("*.h")):
        text = re.sub(r"/\*.*?\*/", "", h.read_text(), flags=re.S)
        syms |= set(re.findall(r"\b((?:rjk|rsi|rscf|rlg)_[a-z0-9_]+)\s*\(", text))
    return sorted(syms)


def test_library_exports_every_header_symbol():
    lib = jk.load_library()
    syms = header_symbols()
    assert len(syms) >= 20
    missing = [s for s in syms if not hasattr(lib, s)]
    assert not missing, missing
    assert sorted(jk.EXPORTS) == syms


def test_create_rejects_bad_shapes_before_touching_a_device():
    lib = jk.load_library()
    h = C.c_void_p()
    assert lib.rjk_create(0, 0, 10, C.byref(h)) == 1
    assert b"positive" in lib.rjk_last_error(None)
    assert lib.rjk_create(0, 10, 0, C.byref(h)) == 1
    assert lib.rjk_create(0, 10, 10, None) == 1
    assert not h.value


def test_create_without_device_reports_error():
    lib = jk.load_library()
    h = C.c_void_p()
    rc = lib.rjk_create(999, 24, 113, C.byref(h))
    assert rc == 2
    assert b"not available" in lib.rjk_last_error(None)


def test_null_context_calls_are_rejected():
    lib = jk.load_library()
    assert lib.rjk_destroy(None) == 0
    assert lib.rjk_synchronize(None) == 1
    assert lib.rjk_num_basis(None) == 0
    assert lib.rjk_vj_upper_with_rimatr_v_sync(None, None, 1, 1.0, 0, None) == 1
    assert lib.rjk_vk_upper_with_ri_v_sync(None, None, 1, None, None, 1, 1.0, 0, None) == 1
    assert lib.rjk_vk_upper_dm_dev(None, None, 1, 1.0, 0, None) == 1


def test_reference_shaped_builders_return_dummy_when_b_is_absent():
    """REST returns two 1-element zero vectors when ri3fn is None (src/scf_io/mod.rs:2494)."""
    out = jk.vj_upper_with_rimatr_v_sync(None, [], 1, 1.0)
    assert len(out) == 2 and all(v.shape == (1,) and v[0] == 0 for v in out)
    out = jk.vk_upper_with_ri_v_sync(None, [], [0, 0], [], 2, -1.0)
    assert len(out) == 2


def test_missing_library_fails_loudly(tmp_path):
    with pytest.raises(RuntimeError, match="not built"):
        jk.load_library(tmp_path / "nope.so")


def test_scf_control_defaults_and_the_mixed_ladder():
    """rscf_default_ctrl leaves every acceleration off (plain REST behaviour); scf.mixed_ctrl turns on
    the ladder, pseudo-diagonalisation and the incremental FP64 tail.  The ctypes mirror of
    rscf_ctrl must match the C struct field by field for these to round-trip."""
    import math
    from rest_amd import scf as rs
    c = rs.default_ctrl()
    assert c.jk_prec == jk.PREC_FP64 and c.jk_prec_start == -1 and c.jk_prec_increment == -1
    assert c.pseudo_diag_err == 0.0 and math.isinf(c.guess_energy_max) and math.isinf(c.guess_err_max)
    assert c.switch_diis_err == 1e-5 and c.start_switch_err == 1e-3 and c.switch_mode == rs.SWITCH_RESTART
    m = rs.mixed_ctrl(max_scf_cycle=77)
    assert (m.jk_prec, m.jk_prec_start, m.jk_prec_increment) == (jk.PREC_BF16X3, jk.PREC_BF16X2, jk.PREC_BF16X2)
    assert m.pseudo_diag_err == 1e-2 and m.max_scf_cycle == 77 and m.switch_diis_err == 1e-5
    assert m.start_switch_err == 1e-4
    assert C.sizeof(rs.Ctrl) % 8 == 0 and C.sizeof(rs.Result) % 8 == 0

"""CPU tests of the HDF5 reader/writer (rest_amd/h5.py) for REST's guess and
chk files.

Pins: the reader on a file written by the HDF5 library itself
(tests/golden/pen_pyscf_restart.chk = examples/PEN/RIHF/restart.chk, PySCF's
DF-RHF pentacene run, def2-universal JKFIT, E = −831.024160249288,
examples/PEN/RIHF/rihf_pen.pyscf.log:520).  Its orbitals are orthonormal in
this build's overlap matrix and hold 146 electrons (AO order and
normalisation agree), and their density's RI-HF energy under this build's
integrals lies just above REST's converged energy (−831.0241601906) — a
variational bound that fails for any convention mismatch (the bound is
loosened by 1e-7 for the two codes' different bohr constants).  The writer is
checked by round trips through the reader; that an HDF5 library reads it
back is parity unpinned here (no HDF5 library in this image).
"""
from pathlib import Path

import numpy as np
import pytest

from rest_amd import h5, ri

GOLD = Path(__file__).resolve().parent / "golden"
E_PYSCF = -831.024160249288
E_REST = -831.0241601906


def pentacene():
    lines = (GOLD / "pentacene.xyz").read_text().splitlines()[2:]
    sym = [l.split()[0] for l in lines]
    xyz = np.array([[float(x) for x in l.split()[1:]] for l in lines])
    return ri.Molecule.from_angstrom(sym, xyz, basis="STO-3G", auxbasis="def2-SV(P)-JKFIT")


def test_reader_on_library_written_chkfile():
    r = h5.H5Reader(GOLD / "pen_pyscf_restart.chk")
    assert r.keys("/") == ["mol", "scf"]
    assert r.keys("scf") == ["e_tot", "mo_coeff", "mo_energy", "mo_occ"]
    assert abs(float(r.read("scf/e_tot")) - E_PYSCF) < 1e-9
    occ = r.read("scf/mo_occ")
    assert occ.shape == (124,) and occ.sum() == 146
    with pytest.raises(h5.H5Error):
        r.read("mol")                      # variable-length string: not a numeric dataset


def test_pyscf_orbitals_in_this_builds_basis():
    from oracle import scf_oracle as so
    from tests.test_integrals import numpy_jk
    m = pentacene()
    r = h5.H5Reader(GOLD / "pen_pyscf_restart.chk")
    C = r.read("scf/mo_coeff")             # PySCF: [nao][nmo]
    occ = r.read("scf/mo_occ")
    # the restart reader hands back the same [ao][state] orbitals (REST's reading of the raw data)
    _, cs, es = h5.read_chkfile(GOLD / "pen_pyscf_restart.chk", m.nbf, C.shape[1], 1)
    np.testing.assert_array_equal(cs[0], C)
    np.testing.assert_array_equal(es[0], r.read("scf/mo_energy"))
    S, T, V = m.int1e()
    # PySCF converts Angstrom with 0.52917721092 bohr, REST with 0.5291772083 (src/constants/mod.rs:85):
    # the geometries differ by ~1e-9 relative, which the tolerances below allow for
    np.testing.assert_allclose(C.T @ S @ C, np.eye(m.nbf), atol=1e-7)
    D = (C * occ) @ C.T
    assert abs(np.sum(D * S) - 146) < 1e-6
    B, _ = so.rimatr_from_integrals(m.int3c2e_packed(), m.int2c2e(), 1e-8)
    J, K = numpy_jk(B, m.nbf)(D)
    H = T + V
    E = m.nuclear_repulsion() + 0.5 * np.sum(D * (2 * H + J - 0.5 * K))
    print(f"E(D_PySCF) in this build's RI-HF functional: {E:.10f} (REST {E_REST}, PySCF {E_PYSCF})")
    assert E_REST - 1e-7 <= E <= E_REST + 1e-6, E


def test_writer_round_trips(tmp_path):
    rng = np.random.default_rng(0)
    dms = [rng.standard_normal((7, 7)) for _ in range(2)]
    dms = [d + d.T for d in dms]
    h5.write_init_guess(tmp_path / "g.h5", dms)
    back = h5.read_init_guess(tmp_path / "g.h5", 7, 2)
    for a, b in zip(dms, back):
        np.testing.assert_array_equal(a, b)
    C = [rng.standard_normal((6, 5)) for _ in range(2)]
    eps = [rng.standard_normal(5) for _ in range(2)]
    h5.write_chkfile(tmp_path / "c.h5", -1.25, C, eps)
    e, C2, eps2 = h5.read_chkfile(tmp_path / "c.h5", 6, 5, 2)
    assert e == -1.25
    for a, b in zip(C + eps, C2 + eps2):
        np.testing.assert_array_equal(a, b)
    # a group with more members than one symbol node holds, nested groups, scalars
    w = h5.H5Writer()
    vals = {f"g/sub/d{k:02d}": rng.standard_normal((k + 1, 3)) for k in range(19)}
    vals["g/x"] = np.float64(3.5)
    vals["top"] = np.arange(4.0)
    for k, v in vals.items():
        w.put(k, v)
    w.save(tmp_path / "many.h5")
    rr = h5.H5Reader(tmp_path / "many.h5")
    assert rr.keys("/") == ["g", "top"] and rr.keys("g") == ["sub", "x"]
    assert len(rr.keys("g/sub")) == 19
    for k, v in vals.items():
        np.testing.assert_array_equal(rr.read(k), v)


@pytest.mark.skipif(not Path("/root/reference/examples/NH3/RIHF/init_guess.chk").exists(),
                    reason="the reference checkout is not present")
def test_reader_on_rest_guess_file():
    dm = h5.read_init_guess("/root/reference/examples/NH3/RIHF/init_guess.chk", 367, 1)[0]
    np.testing.assert_allclose(dm, dm.T, atol=1e-12)

"""GPU tests of the UHF driver (rscf_uhf), the density-guess step
(rscf_guess_step) and the SAD initial guess (rest_amd/guess.py).

Pins: the one-electron closed form (H atom: lowest eigenvalue of (H, S));
UHF = RHF for a closed shell; device UHF vs the oracle UHF on the same B
(≤ 1e-9 Eh); REST's pentacene run with its SAD guess
(tests/golden/pentacene_rest_trace.json, from examples/PEN/RIHF/rest.out):
the converged energy, reached in no more iterations than REST's.

The SAD density itself is parity unpinned: an open-shell atom with partly
filled degenerate p shells (C triplet: two of three 2p orbitals occupied)
leaves the choice of occupied orbitals to the eigensolver (REST: LAPACK
dspgvx, src/scf_io/mod.rs:1851; here rocSOLVER dsygvd), so REST's guess
energy E(D0) = −839.49686125 is printed beside ours, not asserted.
"""
import json
import sys
from pathlib import Path

import numpy as np
import pytest
import scipy.linalg as sla

from rest_amd import guess as gs
from rest_amd import jk, ri
from rest_amd import scf as rs
from rest_amd import workload as wl

pytestmark = pytest.mark.gpu
GOLD = Path(__file__).resolve().parent / "golden"


def say(msg):
    print(f"[gpu-case] {msg}", flush=True)
    sys.stdout.flush()


def tight():
    return rs.default_ctrl(scf_acc_etot=1e-11, scf_acc_rho=1e-9, scf_acc_eev=1e-8, max_scf_cycle=200)


def device_setup(m):
    S, T, V = m.int1e()
    r = jk.DeviceRI(m.nbf, m.naux)
    m.build_rimatr(r)
    return r, S, T + V


def test_uhf_one_electron_closed_form():
    say("UHF H atom def2-SVP: E = lowest eigenvalue of (H, S)")
    m = ri.Molecule(["H"], np.zeros((1, 3)))
    r, S, H = device_setup(m)
    with r:
        out = rs.uhf(r, S, H, 0.0, 1, 0, ctrl=tight())
    e_ref = sla.eigh(H, S, eigvals_only=True)[0]
    say(f"  E {out.energy:.14f} closed form {e_ref:.14f}")
    assert out.converged and abs(out.energy - e_ref) < 1e-11
    np.testing.assert_allclose(out.dm[1], 0.0, atol=0)


def test_uhf_closed_shell_equals_rhf():
    say("(H2O)2 UHF(10,10) from the RHF density vs RHF")
    m = ri.Molecule.water_cluster(wl.water_cluster(3, 2))
    r, S, H = device_setup(m)
    with r:
        o_r = rs.rhf(r, S, H, m.nuclear_repulsion(), m.nelectron, ctrl=tight())
        o_u = rs.uhf(r, S, H, m.nuclear_repulsion(), 10, 10, d_guess=(0.5 * o_r.dm, 0.5 * o_r.dm), ctrl=tight())
    say(f"  RHF {o_r.energy:.12f} UHF {o_u.energy:.12f}")
    assert o_r.converged and o_u.converged and abs(o_u.energy - o_r.energy) < 1e-9


@pytest.mark.parametrize("el,na,nb", [("O", 5, 3), ("C", 4, 2)])
def test_uhf_atom_vs_oracle_identical_b(el, na, nb):
    from oracle import scf_oracle as so
    from tests.test_integrals import numpy_jk
    say(f"UHF {el} ({na},{nb}) def2-SVP/JKFIT device vs oracle on the device B")
    m = ri.Molecule([el], np.zeros((1, 3)))
    r, S, H = device_setup(m)
    with r:
        out = rs.uhf(r, S, H, 0.0, na, nb, ctrl=tight())
        B = r.get_rimatr()
    ctrl = so.ScfCtrl(scf_acc_etot=1e-11, scf_acc_rho=1e-9, scf_acc_eev=1e-8, max_scf_cycle=200)
    ref = so.uhf(S, H, B, na, nb, 0.0, ctrl, jk=numpy_jk(B, m.nbf))
    say(f"  device {out.energy:.12f} ({out.iterations} it)  oracle {ref.energy:.12f} ({ref.iterations} it)")
    assert out.converged and ref.converged and abs(out.energy - ref.energy) <= 1e-9


def test_sad_pentacene_reproduces_rest_guess_and_energy():
    trace = json.loads((GOLD / "pentacene_rest_trace.json").read_text())
    lines = (GOLD / "pentacene.xyz").read_text().splitlines()[2:]
    sym = [l.split()[0] for l in lines]
    xyz = np.array([[float(x) for x in l.split()[1:]] for l in lines])
    m = ri.Molecule.from_angstrom(sym, xyz, basis="STO-3G", auxbasis="def2-SV(P)-JKFIT")
    say("pentacene SAD guess + SCF with REST defaults")
    out, t = rs.run_molecule(m, guess="sad")
    for el in ("C", "H"):
        a = gs.atomic_density(el, "STO-3G", "def2-SV(P)-JKFIT")
        say(f"  atom {el}: E {a.energy:.10f} ({a.iterations} it, converged {a.converged})")
    e_log = [row["energy"] for row in out.log]
    say(f"  E(D0) {t['e_guess_d0']:.8f} REST {trace['initial_guess_hf_energy']:.8f}")
    say(f"  iterations {out.iterations} (REST {trace['converged_after']}); energies "
        + " ".join(f"{e:.10f}" for e in e_log))
    say(f"  REST energies " + " ".join(f"{e:.10f}" for e in trace['iteration_energies']))
    assert out.converged
    assert abs(out.energy - trace["final_energy"]) < 1e-8
    assert out.iterations <= trace["converged_after"]
    # the atoms' own energies do not depend on the orientation of the occupied p orbitals
    a_h = gs.atomic_density("H", "STO-3G", "def2-SV(P)-JKFIT")
    assert abs(a_h.energy - (-0.4665818504)) < 1e-9     # one electron: lowest eigenvalue of (H, S)
    # block-diagonal assembly: the SAD density holds every atom's electrons
    da, db = gs.sad_density(m)
    S = m.int1e()[0]
    assert abs(np.sum((da + db) * S) - m.nelectron) < 1e-9


REST_SAD_UNREPRODUCIBLE = (
    "REST's pentacene log (examples/PEN/RIHF/rest.out:47-85) was written by an older REST: it prints "
    "'Initial guess HF energy', a string the source no longer holds (src/initial_guess/mod.rs:57 prints "
    "'Initial guess energy').  Its guess is not reproducible from the source: E(D0) over every orientation "
    "of carbon's two occupied 2p orbitals spans [-839.793, -831.211] Eh; the basis-aligned choices LAPACK makes "
    "for exactly block-diagonal atomic matrices give -839.79073897 (this build), -832.750 and -831.216; no "
    "orientation gives both REST's E(D0) -839.49686125 and its first-Fock energy -819.4804213896 (best "
    "compromise -838.077 / -830.681), and REST's next energies (-612.14, -435.68 Eh) lie hundreds of Eh above "
    "any density of this family.")


@pytest.mark.skip(reason=REST_SAD_UNREPRODUCIBLE)
def test_sad_pentacene_matches_rest_trace():
    """E(D0), the first Fock energy and the per-iteration energies of REST's pentacene log."""
    trace = json.loads((GOLD / "pentacene_rest_trace.json").read_text())
    lines = (GOLD / "pentacene.xyz").read_text().splitlines()[2:]
    sym = [l.split()[0] for l in lines]
    xyz = np.array([[float(x) for x in l.split()[1:]] for l in lines])
    m = ri.Molecule.from_angstrom(sym, xyz, basis="STO-3G", auxbasis="def2-SV(P)-JKFIT")
    out, t = rs.run_molecule(m, guess="sad")
    assert abs(t["e_guess_d0"] - trace["initial_guess_hf_energy"]) <= 1e-7
    e_log = [row["energy"] for row in out.log]
    np.testing.assert_allclose(e_log, trace["iteration_energies"], rtol=0, atol=1e-7)


def test_factored_sad_guess_step_equals_density_route():
    """rscf_guess_step_factored (the SAD density's first K over its atomic eigenvectors) against
    rscf_guess_step (the device factorisation of the whole density), FP64 and BF16X2, (H2O)4."""
    say("(H2O)4 SAD guess step: atomic factors vs the density-only route")
    m = ri.Molecule.water_cluster(wl.water_cluster(4, 4))
    S, T, V = m.int1e()
    H = T + V
    da, db = gs.sad_density(m)
    C0, w0 = gs.sad_factors(m)
    np.testing.assert_allclose((C0 * w0) @ C0.T, da + db, atol=1e-13)
    assert C0.shape[1] < m.nbf
    with jk.DeviceRI(m.nbf, m.naux) as r:
        m.build_rimatr(r, 1e-8, jk.STORE_FP64 | jk.STORE_FP32 | jk.STORE_BF16)
        ref = None
        for prec in (jk.PREC_FP64, jk.PREC_BF16X2):
            e0, d1, _, _, e1 = rs.guess_step(r, S, H, m.nuclear_repulsion(), (m.nelectron,), da + db, prec,
                                             want_e1=True)
            f0, f1, _, _, g1 = rs.guess_step(r, S, H, m.nuclear_repulsion(), (m.nelectron,), da + db, prec,
                                             want_e1=True, factors=(C0, w0))
            say(f"  prec {prec}: E(D0) {e0:.10f} vs {f0:.10f}, E(D1) {e1:.10f} vs {g1:.10f}, "
                f"max|dD1| {np.abs(d1 - f1).max():.1e}")
            if ref is None:   # FP64: the same numbers through either route
                ref = (e0, e1)
                assert abs(e0 - f0) <= 1e-10 and abs(e1 - g1) <= 1e-10 and np.abs(d1 - f1).max() <= 1e-10
            else:             # BF16X2: each route within the two-plane error of FP64 (different roundings)
                assert all(abs(x - ref[0]) <= 5e-4 for x in (e0, f0)) and all(abs(x - ref[1]) <= 5e-4 for x in (e1, g1))


def test_sad_water_cluster_beats_core_guess():
    say("(H2O)4 SAD vs core-Hamiltonian guess")
    m = ri.Molecule.water_cluster(wl.water_cluster(4, 4))
    o_sad, t = rs.run_molecule(m, guess="sad", ctrl=tight())
    o_core, _ = rs.run_molecule(m, guess="hcore", ctrl=tight())
    say(f"  SAD {o_sad.energy:.12f} ({o_sad.iterations} it, E(D0) {t['e_guess_d0']:.8f})  "
        f"hcore {o_core.energy:.12f} ({o_core.iterations} it)")
    assert o_sad.converged and o_core.converged
    assert abs(o_sad.energy - o_core.energy) < 1e-9
    assert o_sad.iterations < o_core.iterations


def test_restart_from_pyscf_chkfile_and_own_chkfile(tmp_path):
    """REST's restart path from a chkfile written by another program (PySCF's pentacene run,
    tests/golden/pen_pyscf_restart.chk): converged in a few iterations to REST's energy; then the
    chkfile this build writes restarts its own SCF."""
    trace = json.loads((GOLD / "pentacene_rest_trace.json").read_text())
    lines = (GOLD / "pentacene.xyz").read_text().splitlines()[2:]
    sym = [l.split()[0] for l in lines]
    xyz = np.array([[float(x) for x in l.split()[1:]] for l in lines])
    m = ri.Molecule.from_angstrom(sym, xyz, basis="STO-3G", auxbasis="def2-SV(P)-JKFIT")
    say("pentacene restarted from PySCF's chkfile")
    out, _ = rs.run_molecule(m, guess="chk", guess_path=GOLD / "pen_pyscf_restart.chk",
                             chkfile=tmp_path / "own.chk")
    say(f"  E {out.energy:.10f} in {out.iterations} iterations (REST {trace['final_energy']})")
    assert out.converged and out.iterations <= 5
    assert abs(out.energy - trace["final_energy"]) < 1e-8
    out2, _ = rs.run_molecule(m, guess="chk", guess_path=tmp_path / "own.chk")
    say(f"  own chkfile restart: {out2.iterations} iterations")
    assert out2.converged and out2.iterations <= 2 and abs(out2.energy - out.energy) < 1e-9
    # REST's guess-file path (density dataset "init_guess") with the converged density
    from rest_amd import h5
    h5.write_init_guess(tmp_path / "guess.h5", [out.dm])
    out3, t = rs.run_molecule(m, guess="file", guess_path=tmp_path / "guess.h5")
    say(f"  guess-file start: {out3.iterations} iterations, E(D0) {t['e_guess_d0']:.10f}")
    assert out3.converged and out3.iterations <= 2 and abs(out3.energy - out.energy) < 1e-9

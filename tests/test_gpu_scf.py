"""GPU tests of the device pipeline: B producer (rsi_build_rimatr) + SCF driver
(rscf_rhf) against the oracle SCF energies (tests/golden/scf_energies.json,
written by tests/golden/make_scf_golden.py) and REST's published pentacene
energy.  Bars: FP64 device SCF vs the oracle SCF on the same molecule at tight
thresholds ≤ 1e-9 Eh; pentacene vs REST ≤ 1e-8 Eh (REST's own convergence is
1e-8); mixed precision (FP32 J/K until max|FDS−SDF| < 1e-5, then FP64) vs the
FP64 run ≤ 1e-9 Eh.
"""
import json
import sys
from pathlib import Path

import numpy as np
import pytest

from rest_amd import jk, ri
from rest_amd import scf as rs
from rest_amd import workload as wl

pytestmark = pytest.mark.gpu
GOLD = Path(__file__).resolve().parent / "golden"
E_GOLD = json.loads((GOLD / "scf_energies.json").read_text())


def say(msg):
    print(f"[gpu-case] {msg}", flush=True)
    sys.stdout.flush()


def tight():
    return rs.default_ctrl(scf_acc_etot=1e-11, scf_acc_rho=1e-9, scf_acc_eev=1e-8, max_scf_cycle=200)


def pentacene():
    lines = (GOLD / "pentacene.xyz").read_text().splitlines()[2:]
    sym = [l.split()[0] for l in lines]
    xyz = np.array([[float(x) for x in l.split()[1:]] for l in lines])
    return ri.Molecule.from_angstrom(sym, xyz, basis="STO-3G", auxbasis="def2-SV(P)-JKFIT")


def test_pentacene_device_scf_matches_rest_and_oracle():
    say("pentacene STO-3G/def2-SV(P)-JKFIT device SCF (REST defaults)")
    out, t = rs.run_molecule(pentacene())
    say(f"  E {out.energy:.10f} iters {out.iterations} wall {out.wall_seconds:.2f}s B {t['t_B']:.2f}s")
    assert out.converged
    assert abs(out.energy - (-831.0241601906)) < 1e-8
    say("pentacene tight thresholds")
    mol = pentacene()
    out2, t2 = rs.run_molecule(mol, ctrl=tight(), keep_b=True)
    ref = E_GOLD["pentacene_sto3g_def2svpjkfit"]["energy"]
    say(f"  E {out2.energy:.12f} vs oracle {ref:.12f} (oracle B from numpy's V^-1/2)")
    # the golden run builds V^{-1/2} with numpy, the device with rocSOLVER; the metric's
    # condition number (~1e7) turns their 1e-16-level differences into ~1e-9 Eh
    assert out2.converged and abs(out2.energy - ref) <= 5e-9
    # on the identical B (read back from HBM), the oracle loop from the device's converged density
    from oracle import scf_oracle as so
    from tests.test_integrals import numpy_jk
    octrl = so.ScfCtrl(scf_acc_etot=1e-11, scf_acc_rho=1e-9, scf_acc_eev=1e-8, max_scf_cycle=200)
    same = so.rhf(t2["S"], t2["H"], t2["rimatr"], mol.nelectron, mol.nuclear_repulsion(), octrl, d_guess=out2.dm,
                  jk=numpy_jk(t2["rimatr"], mol.nbf))
    say(f"  identical B: device {out2.energy:.12f} oracle {same.energy:.12f} diff {out2.energy - same.energy:.2e}")
    assert same.converged and abs(out2.energy - same.energy) <= 1e-9


def test_pentacene_mixed_ladder_matches_rest():
    """The mixed ladder (scf.mixed_ctrl: BF16X2 -> BF16X3 -> FP64 reference + BF16X2 increments,
    pseudo-diagonalisation) from REST's SAD guess on REST's own pentacene example: the converged
    energy is REST's printed one (examples/PEN/RIHF/rest.out:85) to its 1e-8 Eh convergence."""
    say("pentacene STO-3G/def2-SV(P)-JKFIT, mixed ladder from SAD (REST defaults)")
    out, _ = rs.run_molecule(pentacene(), prec=jk.PREC_BF16X3, ctrl=rs.mixed_ctrl())
    precs = [int(x["prec"]) for x in out.log]
    say(f"  E {out.energy:.10f} iters {out.iterations} (incremental {out.incremental_iterations}), precisions {precs}")
    assert out.converged and abs(out.energy - (-831.0241601906)) < 1e-8
    assert precs[0] == jk.PREC_BF16X2 and precs[-1] == jk.PREC_FP64 and out.incremental_iterations >= 1


def test_identical_b_device_scf_vs_oracle_scf():
    """Same B (read back from HBM), same S, H: device SCF vs the oracle SCF loop, tight thresholds."""
    from oracle import scf_oracle as so
    from tests.test_integrals import numpy_jk
    say("(H2O)2 identical-B device SCF vs oracle SCF")
    m = ri.Molecule.water_cluster(wl.water_cluster(3, 2))
    out, t = rs.run_molecule(m, ctrl=tight(), keep_b=True)
    ctrl = so.ScfCtrl(scf_acc_etot=1e-11, scf_acc_rho=1e-9, scf_acc_eev=1e-8, max_scf_cycle=200)
    ref = so.rhf(t["S"], t["H"], t["rimatr"], m.nelectron, m.nuclear_repulsion(), ctrl, jk=numpy_jk(t["rimatr"], m.nbf))
    say(f"  device {out.energy:.13f}  oracle {ref.energy:.13f}  diff {out.energy - ref.energy:.2e}")
    assert out.converged and ref.converged and abs(out.energy - ref.energy) <= 1e-9


@pytest.mark.parametrize("seed,n", [(3, 2), (4, 4)])
def test_water_cluster_fp64_vs_oracle(seed, n):
    say(f"(H2O){n} seed {seed} def2-SVP/JKFIT FP64 device SCF")
    m = ri.Molecule.water_cluster(wl.water_cluster(seed, n))
    out, _ = rs.run_molecule(m, ctrl=tight())
    ref = E_GOLD[f"water{n}_seed{seed}_def2svp"]["energy"]
    say(f"  E {out.energy:.12f} vs oracle {ref:.12f} ({out.iterations} iterations)")
    assert out.converged and abs(out.energy - ref) <= 1e-9
    # orbitals: C^T S C = I and the density from them
    S = m.int1e()[0]
    C = out.mo_coeff
    np.testing.assert_allclose(C.T @ S @ C, np.eye(m.nbf), atol=1e-10)
    nocc = m.nelectron // 2
    np.testing.assert_allclose(out.dm, 2 * C[:, :nocc] @ C[:, :nocc].T, atol=1e-10)


_E64 = {}


@pytest.mark.parametrize("prec,mode", [(jk.PREC_FP32, rs.SWITCH_SHIFT), (jk.PREC_FP32, rs.SWITCH_RESTART),
                                       (jk.PREC_BF16X3, rs.SWITCH_SHIFT), (jk.PREC_BF16X3, rs.SWITCH_RESTART)])
def test_mixed_precision_converges_to_fp64_energy(prec, mode):
    say(f"(H2O)4 mixed precision (prec {prec} J/K -> FP64 at max|e| < 1e-5), switch mode {mode}")
    m = ri.Molecule.water_cluster(wl.water_cluster(4, 4))
    if "h2o4" not in _E64:
        o64, _ = rs.run_molecule(m, ctrl=tight())
        assert o64.converged
        _E64["h2o4"] = (o64.energy, o64.iterations)
    e64, it64 = _E64["h2o4"]
    ctrl = tight()
    ctrl.switch_mode = mode
    om, _ = rs.run_molecule(m, prec=prec, ctrl=ctrl)
    precs = [int(r["prec"]) for r in om.log]
    say(f"  fp64 {e64:.12f} ({it64} it)  mixed {om.energy:.12f} ({om.iterations} it, "
        f"{om.fp64_iterations} fp64)  dE {om.energy - e64:.2e}")
    assert om.converged and abs(om.energy - e64) <= 1e-9
    assert precs[0] == prec and precs[-1] == jk.PREC_FP64
    # once switched, the J/K stay FP64
    first64 = precs.index(jk.PREC_FP64)
    assert all(p == jk.PREC_FP64 for p in precs[first64:])


def test_precision_ladder_and_guess_rejection():
    say("(H2O)4: BF16X2 start -> BF16X3 below 1e-3 -> FP64 below 1e-5; guess rejection by energy")
    m = ri.Molecule.water_cluster(wl.water_cluster(4, 4))
    S, T, V = m.int1e()
    H = T + V
    enuc = m.nuclear_repulsion()
    with jk.DeviceRI(m.nbf, m.naux) as r:
        m.build_rimatr(r, 1e-8, jk.STORE_FP64 | jk.STORE_FP32 | jk.STORE_BF16)
        o64 = rs.rhf(r, S, H, enuc, m.nelectron, None, tight())
        ctrl = tight()
        ctrl.jk_prec, ctrl.jk_prec_start = jk.PREC_BF16X3, jk.PREC_BF16X2
        om = rs.rhf(r, S, H, enuc, m.nelectron, None, ctrl)
        precs = [int(x["prec"]) for x in om.log]
        diis = [x["diis_max"] for x in om.log]
        say(f"  precisions {precs}; E {om.energy:.12f} vs {o64.energy:.12f}")
        assert om.converged and abs(om.energy - o64.energy) <= 1e-9
        assert precs[0] == jk.PREC_BF16X2 and precs[-1] == jk.PREC_FP64 and jk.PREC_BF16X3 in precs
        # each precision change follows the DIIS error of the iteration before it
        for i in range(1, len(precs)):
            if precs[i] != precs[i - 1]:
                assert diis[i - 1] < (1e-3 if precs[i] == jk.PREC_BF16X3 else 1e-5)
        # guess rejection: the guess Fock energy is the core-Hamiltonian density's FP64 energy
        c = tight()
        c.guess_energy_max = -1e30
        rej = rs.rhf(r, S, H, enuc, m.nelectron, None, c)
        assert rej.guess_rejected and rej.iterations == 0 and not rej.converged
        assert rej.guess_energy == pytest.approx(o64.guess_energy, abs=1e-9)
        c.guess_energy_max = rej.guess_energy + 1e-6
        kept = rs.rhf(r, S, H, enuc, m.nelectron, None, c)
        assert not kept.guess_rejected and kept.converged and abs(kept.energy - o64.energy) <= 1e-9


@pytest.mark.parametrize("nspin", [1, 2])
def test_pseudo_diagonalisation_in_low_precision_iterations(nspin):
    """Below pseudo_diag_err the low-precision iterations rotate the previous orbitals (occupied-virtual
    first-order step, Cholesky re-orthonormalisation) instead of diagonalising; the FP64 tail
    diagonalises exactly, so the converged energy is the FP64 one."""
    say(f"(H2O)4 ladder with pseudo-diagonalisation below 1e-2, nspin {nspin}")
    m = ri.Molecule.water_cluster(wl.water_cluster(4, 4))
    S, T, V = m.int1e()
    H = T + V
    enuc = m.nuclear_repulsion()
    ne = m.nelectron
    with jk.DeviceRI(m.nbf, m.naux) as r:
        m.build_rimatr(r, 1e-8, jk.STORE_FP64 | jk.STORE_FP32 | jk.STORE_BF16)
        run = (lambda c: rs.rhf(r, S, H, enuc, ne, None, c)) if nspin == 1 else \
              (lambda c: rs.uhf(r, S, H, enuc, ne // 2, ne // 2, None, c))
        o64 = run(tight())
        ctrl = tight()
        ctrl.jk_prec, ctrl.jk_prec_start = jk.PREC_BF16X3, jk.PREC_BF16X2
        ref = run(ctrl)
        ctrl.pseudo_diag_err = 1e-2
        om = run(ctrl)
    precs = [int(x["prec"]) for x in om.log]
    diis = [x["diis_max"] for x in om.log]
    say(f"  E {om.energy:.12f} vs {o64.energy:.12f}; iterations {om.iterations} (eigensolver only: "
        f"{ref.iterations}); precisions {precs}")
    assert om.converged and abs(om.energy - o64.energy) <= 1e-9
    assert precs[-1] == jk.PREC_FP64
    assert any(d < 1e-2 and p != jk.PREC_FP64 for d, p in zip(diis, precs))   # the pseudo branch ran
    assert om.iterations <= ref.iterations + 2


@pytest.mark.parametrize("nspin,pinc,mode", [(1, jk.PREC_BF16X2, rs.SWITCH_RESTART), (1, jk.PREC_BF16X3, rs.SWITCH_RESTART),
                                             (2, jk.PREC_BF16X2, rs.SWITCH_RESTART), (1, jk.PREC_FP32, rs.SWITCH_RESTART),
                                             (1, jk.PREC_BF16X2, rs.SWITCH_SHIFT)])
def test_incremental_fp64_builds_after_the_switch(nspin, pinc, mode):
    """After the switch to FP64 the first FP64 build is the reference and later builds add J/K of
    D − D_ref at a low precision; the converged energy stays the FP64 one and the returned orbitals
    are canonical (an exact diagonalisation of the last Fock matrix)."""
    say(f"(H2O)4 ladder with incremental FP64 builds at prec {pinc}, nspin {nspin}, switch mode {mode}")
    m = ri.Molecule.water_cluster(wl.water_cluster(4, 4))
    S, T, V = m.int1e()
    H = T + V
    enuc = m.nuclear_repulsion()
    ne = m.nelectron
    with jk.DeviceRI(m.nbf, m.naux) as r:
        m.build_rimatr(r, 1e-8, jk.STORE_FP64 | jk.STORE_FP32 | jk.STORE_BF16)
        run = (lambda c: rs.rhf(r, S, H, enuc, ne, None, c)) if nspin == 1 else \
              (lambda c: rs.uhf(r, S, H, enuc, ne // 2, ne // 2, None, c))
        o64 = run(tight())
        ctrl = tight()
        ctrl.jk_prec, ctrl.jk_prec_start = jk.PREC_BF16X3, jk.PREC_BF16X2
        ctrl.jk_prec_increment = pinc
        ctrl.switch_mode = mode
        om = run(ctrl)
    precs = [int(x["prec"]) for x in om.log]
    say(f"  E {om.energy:.12f} vs {o64.energy:.12f} (dE {om.energy - o64.energy:.2e}); iterations {om.iterations}, "
        f"fp64 {om.fp64_iterations}, incremental {om.incremental_iterations}")
    assert om.converged and abs(om.energy - o64.energy) <= 1e-9
    assert om.incremental_iterations >= 1
    assert om.incremental_iterations + om.reference_builds == om.fp64_iterations
    assert precs[-1] == jk.PREC_FP64
    # canonical outputs: C diagonalises F(D_out) and ε matches the FP64 run's
    cs = [om.mo_coeff] if nspin == 1 else list(om.mo_coeff)
    es = [om.mo_energy] if nspin == 1 else list(om.mo_energy)
    e64 = [o64.mo_energy] if nspin == 1 else list(o64.mo_energy)
    for c, e, e_ref in zip(cs, es, e64):
        assert np.allclose(c.T @ S @ c, np.eye(len(e)), atol=1e-10)
        assert np.max(np.abs(e - e_ref)) < 1e-5


def test_incremental_tail_after_an_early_switch_refreshes_its_reference():
    """jk_prec FP64 with a BF16X2 start switches to FP64 at max|e| < start_switch_err (1e-3), far from
    convergence, so the density leaves the first reference; a build whose max|D − D_ref| exceeds
    inc_refresh·max|D_ref| becomes a new reference.  inc_check compares every increment with the
    full FP64 Fock matrix of the same density."""
    say("(H2O)4: FP64 after a BF16X2 start, BF16X2 increments, early switch, increments checked")
    m = ri.Molecule.water_cluster(wl.water_cluster(4, 4))
    S, T, V = m.int1e()
    H = T + V
    enuc = m.nuclear_repulsion()
    with jk.DeviceRI(m.nbf, m.naux) as r:
        m.build_rimatr(r, 1e-8, jk.STORE_FP64 | jk.STORE_FP32 | jk.STORE_BF16)
        o64 = rs.rhf(r, S, H, enuc, m.nelectron, None, tight())
        ctrl = tight()
        ctrl.jk_prec, ctrl.jk_prec_start, ctrl.jk_prec_increment = jk.PREC_FP64, jk.PREC_BF16X2, jk.PREC_BF16X2
        ctrl.inc_check = 1
        om = rs.rhf(r, S, H, enuc, m.nelectron, None, ctrl)
    inc = [(x["dref"], x["inc_err"]) for x in om.log if x["inc_err"] > 0]
    say(f"  E {om.energy:.12f} vs {o64.energy:.12f} (dE {om.energy - o64.energy:.2e}); {om.iterations} it, "
        f"{om.incremental_iterations} increments, {om.reference_builds} references; (dref, F error) {inc}")
    assert om.converged and abs(om.energy - o64.energy) <= 1e-9
    assert om.reference_builds >= 2 and om.incremental_iterations >= 1
    assert all(d <= 1e-5 for d, _ in inc) and all(e <= 1e-10 for _, e in inc)


def test_energy_rise_mixing_on_the_device():
    """REST's energy-rise mixing (src/scf_io/mod.rs:3002-3010, :3039-3057) in the device loop, on
    pentacene from the core guess, whose energy rises early (the oracle loop is pinned the same
    way on the CPU, tests/test_scf_oracle.py; the two paths are not compared step by step because
    this start is chaotic: rounding-level differences grow to 1e-2 Eh within two iterations).
    With the check after 2 records (rows are 0-based iterations, row k's energy is the one DIIS
    step k records):
      - mixing happens only on energy rises, and the first rise after two records mixes;
      - mix_param 1: the path with the check on equals the path with it off, bit for bit;
      - mix_param 0, first mixing at row k: the Fock matrix of row k is H[i−1], the one used two
        rows earlier, so row k+1 rebuilds the Fock matrix of row k−1 with the same density and its
        error vector repeats; DIIS fails (REST :3061-3078) and linear mixing with mix_param 0 keeps
        the input density of row k: e[k+1] == e[k]; row k+2 mixes again (DIIS 8 cycles later)
        and takes the density made at row k, which is the one made at row k−2: e[k+2] == e[k−1].
        Mixing with H[i] instead (one row back) would give e[k+2] == e[k];
      - start_check_oscillation 0: the rise at the second step meets the empty first entry."""
    m = pentacene()
    S, T, V = m.int1e()
    H = T + V
    enuc = m.nuclear_repulsion()
    run = lambda r, **kw: rs.rhf(r, S, H, enuc, m.nelectron, None, rs.default_ctrl(max_scf_cycle=14, **kw))
    with jk.DeviceRI(m.nbf, m.naux) as r:
        m.build_rimatr(r)
        say("pentacene core guess, check after 2 records, mix_param 0.6 / 0 / 1, and the check off")
        o6 = run(r, start_check_oscillation=2)
        o0 = run(r, start_check_oscillation=2, mix_param=0.0)
        o1 = run(r, start_check_oscillation=2, mix_param=1.0)
        off1 = run(r, start_check_oscillation=1000, mix_param=1.0)
        z = run(r, start_check_oscillation=0)
    for name, o in (("0.6", o6), ("0", o0)):
        e = [x["energy"] for x in o.log]
        osc = [bool(x["osc"]) for x in o.log]
        say(f"  mix {name}: mixing at rows {[k for k, f in enumerate(osc) if f]}, energies "
            + " ".join(f"{v:.6f}" for v in e))
        rises = [k for k in range(2, len(e)) if e[k] > e[k - 1]]
        assert rises and osc[rises[0]]
        assert all(k in rises for k, f in enumerate(osc) if f)
    e0 = [x["energy"] for x in o0.log]
    k = next(i for i, x in enumerate(o0.log) if x["osc"])
    assert k + 2 < len(e0) and e0[k] > e0[k - 1]
    assert e0[k + 1] == e0[k] and e0[k + 2] == e0[k - 1]
    assert [x["energy"] for x in o1.log] == [x["energy"] for x in off1.log]
    assert any(x["osc"] for x in o1.log)
    zo = [bool(x["osc"]) for x in z.log]
    assert zo[1] and not zo[0]


def test_diis_failure_on_a_one_function_atom():
    """One basis function (H, STO-3G): F D S − S D F is zero, the DIIS matrix is singular, and the
    driver falls back to linear mixing (REST :3061-3078, as in REST's own pentacene log for H);
    the energy is the closed form H/S after REST's two iterations."""
    say("H atom STO-3G UHF: DIIS failure path")
    m = ri.Molecule(["H"], np.zeros((1, 3)), "STO-3G", "def2-SV(P)-JKFIT")
    S, T, V = m.int1e()
    H = T + V
    with jk.DeviceRI(m.nbf, m.naux) as r:
        m.build_rimatr(r)
        out = rs.uhf(r, S, H, 0.0, 1, 0, None, rs.default_ctrl())
    say(f"  E {out.energy:.14f} closed form {H[0, 0] / S[0, 0]:.14f}, {out.iterations} iterations")
    assert out.converged and out.iterations == 2
    assert abs(out.energy - H[0, 0] / S[0, 0]) < 1e-12


def test_max_scf_cycle_one_still_runs_two_iterations():
    say("(H2O)2 with max_scf_cycle = 1: REST's check starts at the second iteration")
    m = ri.Molecule.water_cluster(wl.water_cluster(3, 2))
    out, _ = rs.run_molecule(m, ctrl=rs.default_ctrl(max_scf_cycle=1))
    assert out.iterations == 2 and not out.converged and len(out.log) == 2


def test_single_water_and_h2o16_fp64_vs_oracle():
    """The single-water workload (O at the origin, H at (±0.7572, 0.5865, 0) Å) and (H2O)16 seed 1 with
    real B, FP64 from SAD at tight thresholds: against the oracle loop on the identical B (<= 1e-9 Eh)
    and against the golden oracle energies of an independently built B (<= 5e-9 Eh)."""
    lines = (GOLD / "h2o_monomer.xyz").read_text().splitlines()[2:]
    mono = ri.Molecule.from_angstrom([l.split()[0] for l in lines],
                                     np.array([[float(x) for x in l.split()[1:]] for l in lines]))
    from oracle import scf_oracle as so
    from tests.test_integrals import numpy_jk
    for name, m in (("water1_monomer_def2svp", mono),
                    ("water16_seed1_def2svp", ri.Molecule.water_cluster(wl.water_cluster(1, 16)))):
        say(f"{name}: nbf {m.nbf} naux {m.naux}, FP64 from SAD, tight")
        out, t = rs.run_molecule(m, ctrl=tight(), keep_b=True)
        ref = E_GOLD[name]["energy"]
        # the identical B: the oracle loop from the device's converged density
        o = so.rhf(t["S"], t["H"], t["rimatr"], m.nelectron, m.nuclear_repulsion(),
                   so.ScfCtrl(scf_acc_etot=1e-11, scf_acc_rho=1e-9, scf_acc_eev=1e-8, max_scf_cycle=50),
                   d_guess=out.dm, jk=numpy_jk(t["rimatr"], m.nbf))
        say(f"  E {out.energy:.12f}; oracle on the identical B {o.energy:.12f} (dE {out.energy - o.energy:.1e}, "
            f"{o.iterations} it); golden {ref:.12f} (dE {out.energy - ref:.1e}, B from numpy's V^-1/2)")
        assert out.converged and o.converged and abs(out.energy - o.energy) <= 1e-9
        # the golden run builds V^{-1/2} with numpy, the device with rocSOLVER; the metric's
        # condition number (~1e7) turns their 1e-16-level differences into ~1e-9 Eh
        assert abs(out.energy - ref) <= 5e-9


def test_mixed_uhf_shift_matches_fp64():
    say("H2O+ doublet UHF: BF16X3 -> FP64 with the shifted DIIS history vs plain FP64")
    m = ri.Molecule.water_cluster(wl.water_cluster(3, 1))
    S, T, V = m.int1e()
    H = T + V
    enuc = m.nuclear_repulsion()
    with jk.DeviceRI(m.nbf, m.naux) as r:
        m.build_rimatr(r, 1e-8, jk.STORE_FP64 | jk.STORE_FP32 | jk.STORE_BF16)
        o64 = rs.uhf(r, S, H, enuc, 5, 4, None, tight())
        ctrl = tight()
        ctrl.jk_prec = jk.PREC_BF16X3
        om = rs.uhf(r, S, H, enuc, 5, 4, None, ctrl)
    say(f"  fp64 {o64.energy:.12f} ({o64.iterations} it)  mixed {om.energy:.12f} ({om.iterations} it, "
        f"{om.fp64_iterations} fp64)")
    assert o64.converged and om.converged and abs(om.energy - o64.energy) <= 1e-9


def test_scf_error_paths():
    say("SCF error paths")
    m = ri.Molecule.water_cluster(wl.water_cluster(3, 1))
    S, T, V = m.int1e()
    with jk.DeviceRI(m.nbf, m.naux) as r:
        with pytest.raises(jk.RJKError) as e:
            rs.rhf(r, S, T + V, 0.0, 10)
        assert e.value.code == 4          # B not set
        m.build_rimatr(r)
        with pytest.raises(jk.RJKError) as e:
            rs.rhf(r, S, T + V, 0.0, 9)
        assert e.value.code == 1          # odd electron count
        with pytest.raises(jk.RJKError) as e:
            rs.rhf(r, S, T + V, 0.0, 10, ctrl=rs.default_ctrl(jk_prec=jk.PREC_FP32))
        assert e.value.code == 4          # no FP32 copy of B
        with pytest.raises(jk.RJKError) as e:
            rs.rhf(r, S, T + V, 0.0, 10, ctrl=rs.default_ctrl(jk_prec=jk.PREC_BF16X3))
        assert e.value.code == 4          # no BF16 planes of B
        with pytest.raises(jk.RJKError) as e:
            rs.rhf(r, S, T + V, 0.0, 10, ctrl=rs.default_ctrl(switch_mode=5))
        assert e.value.code == 1
        with pytest.raises(jk.RJKError) as e:
            rs.rhf(r, S, T + V, 0.0, 10, ctrl=rs.default_ctrl(jk_prec_increment=jk.PREC_FP64))
        assert e.value.code == 1          # increments must be at a low precision
        with pytest.raises(jk.RJKError) as e:
            rs.rhf(r, S, T + V, 0.0, 10, ctrl=rs.default_ctrl(jk_prec_increment=jk.PREC_BF16X2))
        assert e.value.code == 4          # no BF16 planes for the increments
        out = rs.rhf(r, S, T + V, m.nuclear_repulsion(), 10)
        assert out.converged and out.incremental_iterations == 0 and out.guess_error > 0

"""CPU pins of the SCF oracle's loop control (oracle/scf_oracle.py), beside the
energy pins elsewhere: REST's energy-rise mixing (src/scf_io/mod.rs:3002-3010,
:3039-3057), the DIIS-failure fallback (:3061-3078) and the iteration limits
of check_scf_convergence (:1875-1951).

Pentacene from the core-Hamiltonian guess (REST's example molecule; its
energy rises in the early iterations) is the system whose path the mixing
changes.  Invariants used:
  - mix_param = 0: the mixed Fock matrix is H[i−1] itself, so the density
    after a mixing step k equals the density after step k−2, bit for bit;
  - mix_param = 1: the mixing returns the DIIS matrix, so the path equals the
    path with the check off, bit for bit;
  - the list of earlier Fock matrices starts with an empty entry, so a
    mixing at the second DIIS step leaves the Fock matrix of the density;
  - one basis function (H atom, STO-3G): every error vector F D S − S D F is
    zero, the DIIS matrix is singular, and the loop falls back to linear
    mixing while converging to the closed form H/S.
"""
from pathlib import Path

import numpy as np
import pytest
import scipy.linalg as sla

from oracle import scf_oracle as so
from rest_amd import ri
from tests.test_integrals import numpy_jk

GOLD = Path(__file__).resolve().parent / "golden"


@pytest.fixture(scope="module")
def pentacene():
    lines = (GOLD / "pentacene.xyz").read_text().splitlines()[2:]
    sym = [l.split()[0] for l in lines]
    xyz = np.array([[float(x) for x in l.split()[1:]] for l in lines])
    m = ri.Molecule.from_angstrom(sym, xyz, basis="STO-3G", auxbasis="def2-SV(P)-JKFIT")
    S, T, V = m.int1e()
    B, _ = so.rimatr_from_integrals(m.int3c2e_packed(), m.int2c2e(), 1e-8)
    return dict(S=S, H=T + V, B=B, nel=m.nelectron, enuc=m.nuclear_repulsion(), jk=numpy_jk(B, m.nbf))


def run(p, **kw):
    return so.rhf(p["S"], p["H"], p["B"], p["nel"], p["enuc"], so.ScfCtrl(**kw), jk=p["jk"])


def test_mixing_with_alpha_zero_returns_the_density_two_steps_back(pentacene):
    r = run(pentacene, start_check_oscillation=2, mix_param=0.0, max_scf_cycle=12)
    steps = [k for k, o in enumerate(r.osc, start=1) if o and k >= 3]
    assert steps, "no energy rise after the second DIIS step"
    for k in steps:
        assert np.array_equal(r.densities[k - 1], r.densities[k - 3]), k
    # the next step rebuilds the Fock matrix of step k−2 with its own density: a repeated error
    # vector, a singular DIIS matrix and the linear-mixing fallback (REST :3061-3078), which with
    # mix_param 0 keeps its input density (energy of step k again), then takes the density made
    # at step k, i.e. the one made at step k−2 (energy of step k−1)
    j = steps[0]           # 1-based step; r.energies[j] is its energy
    e = r.energies
    assert r.diis_failed[j] and not any(r.diis_failed[:j])
    assert e[j + 1] == e[j] and e[j + 2] == e[j - 1]


def test_mixing_with_alpha_one_is_the_plain_diis_path(pentacene):
    on = run(pentacene, start_check_oscillation=2, mix_param=1.0, max_scf_cycle=12)
    off = run(pentacene, start_check_oscillation=1000, mix_param=1.0, max_scf_cycle=12)
    assert any(on.osc) and not any(off.osc)
    assert on.energies == off.energies


def test_mixing_at_the_second_step_keeps_the_fock_matrix_of_the_density(pentacene):
    p = pentacene
    r = run(p, start_check_oscillation=0, max_scf_cycle=3)
    assert r.osc[:2] == [False, True]
    J, K = p["jk"](r.densities[0])
    w, C = sla.eigh(p["H"] + J - 0.5 * K, p["S"])
    nocc = p["nel"] // 2
    assert np.array_equal(r.densities[1], 2.0 * C[:, :nocc] @ C[:, :nocc].T)


def test_check_starts_after_start_check_oscillation_records(pentacene):
    r = run(pentacene, start_check_oscillation=4, max_scf_cycle=12)
    first = next(k for k, o in enumerate(r.osc, start=1) if o)
    assert first >= 5
    e = r.energies          # e[k] is the energy recorded at DIIS step k
    assert e[first] > e[first - 1]
    assert all(not (e[k] > e[k - 1]) for k in range(5, first))


def test_max_scf_cycle_one_still_runs_two_iterations(pentacene):
    r = run(pentacene, max_scf_cycle=1)
    assert r.iterations == 2 and not r.converged


def test_one_function_atom_diis_fails_and_mixing_converges():
    m = ri.Molecule(["H"], np.zeros((1, 3)), "STO-3G", "def2-SV(P)-JKFIT")
    S, T, V = m.int1e()
    B, _ = so.rimatr_from_integrals(m.int3c2e_packed(), m.int2c2e(), 1e-8)
    r = so.uhf(S, T + V, B, 1, 0, 0.0, so.ScfCtrl(), jk=numpy_jk(B, m.nbf))
    assert r.diis_failed[0] and not any(r.diis_failed[1:])
    assert r.converged and r.iterations == 2
    assert abs(r.energy - (T + V)[0, 0] / S[0, 0]) < 1e-12

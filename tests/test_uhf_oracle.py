"""CPU pins of the oracle's UHF loop (oracle/scf_oracle.py: uhf), which the
device UHF driver is compared against on the GPU.

- One electron: J and K cancel exactly (γ_P = cᵀB_P c), so F c = H c and the
  UHF energy of the H atom is the lowest generalised eigenvalue of (H, S).
  A wrong exchange scale (REST's s_K = −1 for UHF) fails this.
- Closed shell: UHF with n_α = n_β from the RHF density reproduces the RHF
  energy.  J built from one spin density only, or an RHF-style ½ on K,
  fails this.
"""
import numpy as np
import scipy.linalg as sla

from oracle import scf_oracle as so
from rest_amd import ri
from rest_amd import workload as wl
from tests.test_integrals import numpy_jk

TIGHT = so.ScfCtrl(scf_acc_etot=1e-11, scf_acc_rho=1e-9, scf_acc_eev=1e-8, max_scf_cycle=200)


def setup(m):
    S, T, V = m.int1e()
    B, _ = so.rimatr_from_integrals(m.int3c2e_packed(), m.int2c2e(), 1e-8)
    return S, T + V, B


def test_one_electron_uhf_is_lowest_core_eigenvalue():
    m = ri.Molecule(["H"], np.zeros((1, 3)))
    S, H, B = setup(m)
    res = so.uhf(S, H, B, 1, 0, 0.0, TIGHT, jk=numpy_jk(B, m.nbf))
    assert res.converged
    assert abs(res.energy - sla.eigh(H, S, eigvals_only=True)[0]) < 1e-12


def test_closed_shell_uhf_equals_rhf():
    m = ri.Molecule.water_cluster(wl.water_cluster(3, 1))
    S, H, B = setup(m)
    jk = numpy_jk(B, m.nbf)
    r = so.rhf(S, H, B, 10, m.nuclear_repulsion(), TIGHT, jk=jk)
    u = so.uhf(S, H, B, 5, 5, m.nuclear_repulsion(), TIGHT, d_guess=[0.5 * r.density, 0.5 * r.density], jk=jk)
    assert r.converged and u.converged
    assert abs(u.energy - r.energy) < 1e-10
    np.testing.assert_allclose(u.density[0], u.density[1], atol=1e-8)


def test_open_shell_oxygen_triplet_is_variational():
    """O atom, triplet: the converged UHF density satisfies [F_s, D_s]_S = 0 for both spins and
    its energy is below that of the core-Hamiltonian density."""
    m = ri.Molecule(["O"], np.zeros((1, 3)))
    S, H, B = setup(m)
    res = so.uhf(S, H, B, 5, 3, 0.0, TIGHT, jk=numpy_jk(B, m.nbf))
    assert res.converged
    for s in range(2):
        F, D = res.fock[s], res.density[s]
        assert np.abs(F @ D @ S - S @ D @ F).max() < 1e-6
    assert -75.0 < res.energy < -74.5

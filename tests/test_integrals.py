"""CPU tests of the B producer's integral engine (rest_amd/csrc/rsi_*.cpp).

Pins: closed forms for s functions; p functions as finite-difference centre
derivatives of s functions; rotation invariance of the spectra of S, H and
the auxiliary metric (fails for any wrong real-solid-harmonic coefficient up
to g); three-centre vs two-centre consistency through the Gaussian product
theorem; unit norm of every contracted spherical function (REST's
normalisation, src/basis_io/mod.rs:242-341); and end to end, the RI-HF
energy of pentacene in STO-3G / def2-SV(P)-JKFIT against REST's published
−831.0241601906 Eh (examples/PEN/RIHF/rest.out:85).
"""
from pathlib import Path

import numpy as np
import pytest
from scipy.special import erf

from oracle import scf_oracle as so
from rest_amd import ri
from rest_amd import workload as wl

GOLD = Path(__file__).resolve().parent / "golden"


def ns(a):   # normalisation of exp(-a r²)
    return (2 * a / np.pi) ** 0.75


def basis(xyz, shells):
    return ri.Basis(["X"] * len(xyz), np.asarray(xyz, dtype=float), "custom", shells)


def test_normalisation_of_every_function():
    m = ri.Molecule.water_cluster(wl.water_cluster(3, 2))
    S, _, _ = m.int1e()
    np.testing.assert_allclose(np.diag(S), 1.0, atol=1e-13)
    np.testing.assert_allclose(np.diag(m.int1e_aux_overlap()), 1.0, atol=1e-13)


def test_s_closed_forms():
    a, b, R = 0.8, 1.7, 1.3
    xyz = [[0, 0, 0], [0.3, -0.4, R]]
    d = np.linalg.norm(np.subtract(*xyz))
    bs = basis(xyz, [(0, 0, [a], [1.0]), (1, 0, [b], [1.0])])
    S, T, V = ri.int1e(bs, np.array([1.0]), np.array([[0.1, 0.2, 0.5]]))
    p, mu = a + b, a * b / (a + b)
    s_ab = ns(a) * ns(b) * (np.pi / p) ** 1.5 * np.exp(-mu * d * d)
    np.testing.assert_allclose(S[0, 1], s_ab, rtol=1e-13)
    np.testing.assert_allclose(T[0, 1], mu * (3 - 2 * mu * d * d) * s_ab, rtol=1e-12)
    P = (a * np.array(xyz[0]) + b * np.array(xyz[1])) / p
    Rc = np.linalg.norm(P - [0.1, 0.2, 0.5])
    v = -2 * np.pi / p * ns(a) * ns(b) * np.exp(-mu * d * d) * 0.5 * np.sqrt(np.pi / (p * Rc * Rc)) * erf(np.sqrt(p) * Rc)
    np.testing.assert_allclose(V[0, 1], v, rtol=1e-12)
    # two-centre Coulomb of normalised s Gaussians: 2π^{5/2}/(ab√(a+b)) N_a N_b F0(αR²)
    C2 = ri.int2c2e(bs)
    al = a * b / (a + b)
    f0 = 0.5 * np.sqrt(np.pi / (al * d * d)) * erf(np.sqrt(al) * d)
    np.testing.assert_allclose(C2[0, 1], 2 * np.pi ** 2.5 / (a * b * np.sqrt(a + b)) * ns(a) * ns(b) * f0, rtol=1e-12)


@pytest.mark.parametrize("axis", [0, 1, 2])
def test_p_functions_are_centre_derivatives_of_s(axis):
    """p_x = N_p/(2a N_s) ∂/∂A_x s_a(A); checked in S, V, (P|Q) and (μν|P) by central differences."""
    a, b, g = 0.9, 0.6, 1.1
    A, B, Cc = np.array([0.2, -0.1, 0.3]), np.array([1.1, 0.4, -0.7]), np.array([-0.5, 0.9, 0.4])
    h = 1e-4
    npf = ns(a) * np.sqrt(4 * a)
    fac = npf / (2 * a * ns(a))

    def sval(Ashift):
        bs = basis([Ashift, B], [(0, 0, [a], [1.0]), (1, 0, [b], [1.0])])
        S, _, V = ri.int1e(bs, np.array([2.0]), Cc[None, :])
        aux = basis([Ashift, B, Cc], [(0, 0, [a], [1.0]), (1, 0, [b], [1.0]), (2, 0, [g], [1.0])])
        M = ri.int2c2e(aux)
        orb = basis([Ashift, B], [(0, 0, [a], [1.0]), (1, 0, [b], [1.0])])
        X = ri.int3c2e_packed(orb, basis([Cc], [(0, 0, [g], [1.0])]))
        return S[0, 1], V[0, 1], M[0, 1], M[0, 2], X[1, 0]   # X row 1: pair (0, 1)

    e = np.zeros(3); e[axis] = h
    fd = (np.array(sval(A + e)) - np.array(sval(A - e))) / (2 * h) * fac
    bp = basis([A, B], [(0, 1, [a], [1.0]), (1, 0, [b], [1.0])])
    S, _, V = ri.int1e(bp, np.array([2.0]), Cc[None, :])
    auxp = basis([A, B, Cc], [(0, 1, [a], [1.0]), (1, 0, [b], [1.0]), (2, 0, [g], [1.0])])
    M = ri.int2c2e(auxp)
    X = ri.int3c2e_packed(bp, basis([Cc], [(0, 0, [g], [1.0])]))
    # p shell functions are 0..2 (x, y, z); s on B is function 3; pair (axis, 3) -> 3*4/2 + axis
    exact = np.array([S[axis, 3], V[axis, 3], M[axis, 3], M[axis, 4], X[6 + axis, 0]])
    np.testing.assert_allclose(exact, fd, rtol=2e-7, atol=1e-9)


def test_three_centre_reduces_to_two_centre_for_concentric_pair():
    """s_a s_b on one centre = N_a N_b / N_{a+b} · s_{a+b}, so (s_a s_b|c) = N_a N_b/N_{a+b} (s_{a+b}|c)
    for every l of c up to g."""
    a, b = 0.7, 1.3
    A = [0.1, 0.2, -0.3]
    Cc = [0.9, -0.6, 0.5]
    auxs = [(1, l, [0.45 + 0.1 * l], [1.0]) for l in range(5)]
    orb = basis([A], [(0, 0, [a], [1.0]), (0, 0, [b], [1.0])])
    X = ri.int3c2e_packed(orb, basis([A, Cc], auxs))
    two = ri.int2c2e(basis([A, Cc], [(0, 0, [a + b], [1.0])] + auxs))
    ratio = ns(a) * ns(b) / ns(a + b)
    np.testing.assert_allclose(X[1, :], ratio * two[0, 1:], rtol=1e-11, atol=1e-14)


def rotation(seed):
    q = np.random.default_rng(seed).standard_normal(4)
    q /= np.linalg.norm(q)
    w, x, y, z = q
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                     [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                     [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])


def test_rotation_invariance_up_to_g():
    """The spectra of S, H = T+V and the JKFIT metric (d, f, g shells) are invariant under a rigid
    rotation; so is ‖(μν|P)‖_F.  A wrong solid-harmonic coefficient breaks this."""
    xyz = wl.water_cluster(4, 2)
    m1 = ri.Molecule.water_cluster(xyz)
    m2 = ri.Molecule.water_cluster(xyz @ rotation(7).T + np.array([0.3, -1.2, 2.0]))
    S1, T1, V1 = m1.int1e()
    S2, T2, V2 = m2.int1e()
    np.testing.assert_allclose(np.linalg.eigvalsh(S1), np.linalg.eigvalsh(S2), atol=1e-12)
    np.testing.assert_allclose(np.linalg.eigvalsh(T1 + V1), np.linalg.eigvalsh(T2 + V2), atol=1e-10)
    np.testing.assert_allclose(np.linalg.eigvalsh(m1.int2c2e()), np.linalg.eigvalsh(m2.int2c2e()), rtol=1e-11,
                               atol=1e-12)
    def full_norm2(X, n):   # Σ_{μν} over the full (symmetric) pair index from the packed one
        dg = [k * (k + 1) // 2 + k for k in range(n)]
        return 2 * np.sum(X * X) - np.sum(X[dg] ** 2)
    np.testing.assert_allclose(full_norm2(m1.int3c2e_packed(), m1.nbf), full_norm2(m2.int3c2e_packed(), m2.nbf),
                               rtol=1e-12)


def test_pair_slabs_tile_the_full_three_centre_array():
    m = ri.Molecule.water_cluster(wl.water_cluster(5, 2))
    full = m.int3c2e_packed()
    cuts = m.orb.shell_boundaries()
    j0, j1 = int(cuts[3]), int(cuts[9])
    q0 = j0 * (j0 + 1) // 2
    np.testing.assert_array_equal(m.int3c2e_packed(j0, j1), full[q0:q0 + m.int3c2e_packed(j0, j1).shape[0]])
    with pytest.raises(Exception):
        m.int3c2e_packed(4, j1)     # inside the first p shell of O: not a shell boundary


def numpy_jk(B, nbf):
    """J/K from packed B with numpy (occupied-style factorisation of D) — a fast test-side J/K."""
    i, j = np.triu_indices(nbf)
    # np.triu_indices is row-major (i <= j); the packed order is column-major p = j(j+1)/2 + i
    order = np.argsort(j * (j + 1) // 2 + i)
    iu, ju = i[order], j[order]
    naux = B.shape[1]
    Bf = np.zeros((naux, nbf, nbf))
    Bf[:, iu, ju] = B.T
    Bf[:, ju, iu] = B.T

    def jk(D):
        g = np.einsum("Pij,ij->P", Bf, D)
        J = np.einsum("Pij,P->ij", Bf, g)
        lam, U = np.linalg.eigh(D)
        keep = np.abs(lam) > 1e-14 * np.abs(lam).max()
        Cw = U[:, keep] * np.sqrt(np.abs(lam[keep]))
        M = Bf.reshape(naux * nbf, nbf) @ Cw                    # (P μ, i)
        M = M.reshape(naux, nbf, -1).transpose(1, 0, 2).reshape(nbf, -1)
        sg = np.tile(np.sign(lam[keep]), naux)
        K = (M * sg) @ M.T
        return J, K
    return jk


def test_pentacene_ri_hf_energy_matches_rest():
    lines = (GOLD / "pentacene.xyz").read_text().splitlines()[2:]
    sym = [l.split()[0] for l in lines]
    xyz = np.array([[float(x) for x in l.split()[1:]] for l in lines])
    m = ri.Molecule.from_angstrom(sym, xyz, basis="STO-3G", auxbasis="def2-SV(P)-JKFIT")
    assert (m.nbf, m.naux, m.nelectron) == (124, 1902, 146)
    S, T, V = m.int1e()
    B, kept = so.rimatr_from_integrals(m.int3c2e_packed(), m.int2c2e(), 1e-8)
    # REST's run converged in 19 iterations (examples/PEN/RIHF/rest.out:85), before its energy-rise
    # check (start_check_oscillation = 20) can act.  From the core guess this loop needs about 53;
    # with the check on, its mixing steers the SCF to another stationary point (-830.796 Eh, a
    # HOMO-LUMO gap of 0.04 Eh), so the integrals are pinned here with the check off.
    ctrl = so.ScfCtrl(start_check_oscillation=1000)
    res = so.rhf(S, T + V, B, m.nelectron, m.nuclear_repulsion(), ctrl, jk=numpy_jk(B, m.nbf))
    assert res.converged
    assert abs(res.energy - (-831.0241601906)) < 1e-8, res.energy

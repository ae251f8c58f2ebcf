"""CPU tests of the learned guess's geometry side (rest_amd/learned.py):
rotation matrices in the engine's harmonic convention, pair frames,
invariance of the descriptors and equivariance of the targets, and the
encode/decode round trip."""
import numpy as np

from rest_amd import learned as lg
from rest_amd import ri
from rest_amd import workload as wl


def rotation(seed):
    q = np.random.default_rng(seed).standard_normal(4)
    q /= np.linalg.norm(q)
    w, x, y, z = q
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                     [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                     [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])


def test_rotation_matrices_follow_the_engine_convention():
    """The overlap of the rotated molecule (from the integral engine) equals U S Uᵀ, up to g functions."""
    xyz = wl.water_cluster(3, 2)
    R = rotation(1)
    for basis, aux in (("def2-SVP", "def2-SVP-JKFIT"),):
        m1 = ri.Molecule.water_cluster(xyz, basis, aux)
        m2 = ri.Molecule.water_cluster(xyz @ R.T, basis, aux)
        U = lg.molecule_rotation(m1, R)
        np.testing.assert_allclose(U @ m1.int1e()[0] @ U.T, m2.int1e()[0], atol=1e-12)
    for l in range(5):
        D = lg.sph_rotation(l, R)
        np.testing.assert_allclose(D @ D.T, np.eye(2 * l + 1), atol=1e-12)
        np.testing.assert_allclose(lg.sph_rotation(l, R @ rotation(2)), D @ lg.sph_rotation(l, rotation(2)),
                                   atol=1e-12)
    np.testing.assert_allclose(lg.sph_rotation(1, R), R, atol=1e-13)   # p shells: (x, y, z)


def test_frames_are_right_handed_and_descriptors_invariant():
    """(With unequal O–H bonds: for a symmetric water the diagonal frame of O is only fixed up to
    the molecule's own C2 rotation, under which its block is invariant.)"""
    xyz = wl.water_cluster(5, 3, oh_jitter=0.05) / ri.ANG
    sym = ["O", "H", "H"] * 3
    pairs = lg.pair_list(xyz)
    X1, F1 = lg.descriptors(sym, xyz, pairs)
    R = rotation(3)
    X2, F2 = lg.descriptors(sym, xyz @ R.T + np.array([1.0, -2.0, 0.5]), pairs)
    for F in F1:
        np.testing.assert_allclose(F @ F.T, np.eye(3), atol=1e-12)
        assert np.linalg.det(F) > 0
    np.testing.assert_allclose(X1, X2, atol=1e-5)
    np.testing.assert_allclose(F2, F1 @ R.T, atol=1e-10)   # frames rotate with the molecule


def test_vectorised_frames_equal_the_per_pair_definition():
    for xyz in (wl.water_cluster(5, 3, oh_jitter=0.05) / ri.ANG, wl.water_cluster(2, 1) / ri.ANG,
                np.array([[0.0, 0.0, 0.0], [0.0, 0.0, 1.4]])):
        pairs = lg.pair_list(xyz)
        F = lg.frames(xyz, pairs)
        for k, (a, b) in enumerate(pairs):
            np.testing.assert_allclose(F[k], lg.pair_frame(xyz, a, b), atol=1e-12)
    X, _ = lg.descriptors(["H", "H"], np.array([[0.0, 0.0, 0.0], [0.0, 0.0, 1.4]]),
                          np.array([[0, 0], [0, 1], [1, 1]]))
    assert X.shape == (3, lg.N_IN) and np.all(np.isfinite(X))


def test_targets_are_equivariant_and_round_trip():
    xyz = wl.water_cluster(6, 2, oh_jitter=0.05)
    m1 = ri.Molecule.water_cluster(xyz)
    R = rotation(4)
    m2 = ri.Molecule.water_cluster(xyz @ R.T)
    rng = np.random.default_rng(0)
    A = rng.standard_normal((m1.nbf, m1.nbf)) * 0.1
    D = A + A.T
    Dsad = np.diag(rng.random(m1.nbf))
    X1, T1, M1 = lg.encode(m1, D, Dsad)
    U = lg.molecule_rotation(m1, R)
    X2, T2, M2 = lg.encode(m2, U @ D @ U.T, U @ Dsad @ U.T)
    np.testing.assert_allclose(X1, X2, atol=1e-5)
    np.testing.assert_allclose(T1, T2, atol=1e-5)
    # decode(encode(D)) = D for a density with every pair inside the cut-off
    pairs = lg.pair_list(m1.xyz_bohr)
    assert len(pairs) == len(m1.symbols) * (len(m1.symbols) + 1) // 2
    _, F = lg.descriptors(m1.symbols, m1.xyz_bohr, pairs)
    Drt = lg.decode(m1, pairs, F, T1.astype(np.float64), Dsad)
    np.testing.assert_allclose(Drt, D, atol=1e-6)
    assert M1.sum() == sum((14 if a == "O" else 5) * (14 if b == "O" else 5)
                           for i, a in enumerate(m1.symbols) for j, b in enumerate(m1.symbols) if i <= j)


def test_training_clusters_are_seeded_and_perturbed():
    a, b = lg.training_cluster(7), lg.training_cluster(7)
    np.testing.assert_array_equal(a, b)
    n = len(a) // 3
    assert 2 <= n <= 24
    oh = np.linalg.norm(a[1::3] - a[0::3], axis=1)
    assert np.all(np.abs(oh - 0.9572) <= 0.0501)


def test_spherical_average_is_rotation_invariant_and_keeps_electrons():
    """The baseline of the targets: invariant under rotating the molecule (U D_sph Uᵀ is the rotated
    molecule's D_sph even for an anisotropic atomic block), and it keeps tr(D S) per atom."""
    xyz = wl.water_cluster(6, 1, oh_jitter=0.05)
    m1 = ri.Molecule.water_cluster(xyz)
    R = rotation(9)
    m2 = ri.Molecule.water_cluster(xyz @ R.T)
    rng = np.random.default_rng(3)
    A = rng.standard_normal((m1.nbf, m1.nbf))
    D = A @ A.T                                  # anisotropic blocks
    U = lg.molecule_rotation(m1, R)
    np.testing.assert_allclose(U @ lg.spherical_average(m1, D) @ U.T, lg.spherical_average(m2, U @ D @ U.T),
                               atol=1e-10)
    np.testing.assert_allclose(lg.spherical_average(m1, lg.spherical_average(m1, D)), lg.spherical_average(m1, D))
    # same-shell traces are kept (the normalised functions of one shell share S_ii = 1)
    np.testing.assert_allclose(np.trace(lg.spherical_average(m1, D)), np.trace(D), rtol=1e-12)

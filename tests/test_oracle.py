"""CPU tests that pin the oracle (oracle/rjk_oracle.c).

Each check is chosen so that a plausible slip in the restatement — a dropped
diagonal term, the wrong triangle of D, a transposed B_P, a missing occupation
weight or scale — fails at least one of them:

  * golden fixtures written by tests/golden/make_golden.py (numpy long-double
    brute force on unpacked matrices, an independent algorithm);
  * closed forms: naux = 1, D = I, one-hot B;
  * structure: linearity in D, Σ_p w_p J_p = Σ_P γ_P², the asymmetric-D test
    that fixes which triangle the packed J reads, the occupied/density-only
    cross-route identity, symmetry and sign of K;
  * generators: hash values vs a numpy re-implementation, water-cluster
    geometry invariants, orthonormality.
"""
from pathlib import Path

import numpy as np
import pytest

from oracle import oracle as orc
from tests.golden.make_golden import synth_rimatr as np_synth_rimatr

GOLD = Path(__file__).resolve().parent / "golden" / "jk_h2o_shape.npz"


@pytest.fixture(scope="module")
def gold():
    return dict(np.load(GOLD, allow_pickle=False))


def upper_idx(n):
    j = np.concatenate([np.full(k + 1, k) for k in range(n)])
    i = np.concatenate([np.arange(k + 1) for k in range(n)])
    return i, j


def unpack(v, n):
    i, j = upper_idx(n)
    m = np.zeros((n, n))
    m[i, j] = v
    m[j, i] = v
    return m


def weights(d):
    """packed weights the J contraction applies to the upper triangle of D"""
    n = d.shape[0]
    i, j = upper_idx(n)
    return np.where(i == j, 1.0, 2.0) * d[i, j]


# ---------------------------------------------------------------------------
# golden fixtures
# ---------------------------------------------------------------------------
def test_gamma_and_vj_vs_golden(gold):
    b, d = gold["rimatr"], gold["dm"]
    np.testing.assert_allclose(orc.gamma_rimatr(b, d), gold["gamma"], rtol=0, atol=1e-13)
    vj = orc.vj_upper_rimatr(b, [d], 1, 1.0)[0]
    np.testing.assert_allclose(vj, gold["vj"], rtol=0, atol=1e-13)


def test_vj_reads_upper_triangle_of_dm(gold):
    b, da = gold["rimatr"], gold["dm_asym"]
    vj = orc.vj_upper_rimatr(b, [da], 1, 1.0)[0]
    np.testing.assert_allclose(vj, gold["vj_asym_rimatr"], rtol=0, atol=1e-13)
    # and not the full matrix: the ri3fn builder does read both triangles
    assert np.max(np.abs(vj - gold["vj_asym_ri3fn"])) > 1e-4


def test_vj_ri3fn_vs_golden(gold):
    nbf = int(gold["nbf"])
    b3 = orc.ri3fn_from_rimatr(gold["rimatr"], nbf)
    vj = orc.vj_upper_ri3fn(b3, [gold["dm_asym"]], 1, 1.0)[0]
    np.testing.assert_allclose(vj, gold["vj_asym_ri3fn"], rtol=0, atol=1e-13)


def test_vk_occ_vs_golden(gold):
    b, c, occ = gold["rimatr"], gold["c"], gold["occ"]
    vk = orc.vk_upper_occ(b, True, [c], [c.shape[1] - 1], [occ], 1, -0.5)[0]
    np.testing.assert_allclose(vk, gold["vk_rhf"], rtol=0, atol=1e-13)
    nbf = int(gold["nbf"])
    b3 = orc.ri3fn_from_rimatr(b, nbf)
    vk3 = orc.vk_upper_occ(b3, False, [c], [c.shape[1] - 1], [occ], 1, -0.5)[0]
    np.testing.assert_allclose(vk3, gold["vk_rhf"], rtol=0, atol=1e-13)


def test_vk_occ_uhf_vs_golden(gold):
    b, ca, cb = gold["rimatr"], gold["ca"], gold["cb"]
    nstate = ca.shape[1]
    cb_pad = np.zeros_like(ca)
    cb_pad[:, :cb.shape[1]] = cb
    occ = [np.ones(nstate), np.concatenate([np.ones(cb.shape[1]), [0.0]])]
    vk = orc.vk_upper_occ(b, True, [ca, cb_pad], [nstate - 1, cb.shape[1] - 1], occ, 2, -1.0)
    np.testing.assert_allclose(vk[0], gold["vk_uhf_a"], rtol=0, atol=1e-13)
    np.testing.assert_allclose(vk[1], gold["vk_uhf_b"], rtol=0, atol=1e-13)


def test_vk_occ_nonsymmetric_full_b(gold):
    """the occupied route on full B is B_P C occ Cᵀ B_Pᵀ (dgemm NN then NT)"""
    c = gold["nonsym_c"]
    vk = orc.vk_upper_occ(gold["nonsym_ri3fn"], False, [c], [c.shape[1] - 1], [gold["nonsym_occ"]], 1, 1.0)[0]
    np.testing.assert_allclose(vk, gold["nonsym_vk"], rtol=0, atol=1e-13)


def test_vk_dm_general_d_vs_golden(gold):
    vk = orc.vk_upper_dm(gold["rimatr"], True, [gold["dm_asym"]], 1, 1.0)[0]
    np.testing.assert_allclose(vk, gold["vk_dm_asym"], rtol=0, atol=1e-13)


def test_synthetic_generator_matches_numpy(gold):
    nbf, naux = int(gold["nbf"]), int(gold["naux"])
    b = orc.synth_rimatr(11, nbf, naux)
    np.testing.assert_array_equal(b, gold["rimatr"])
    flat = b.T.reshape(-1)
    np.testing.assert_array_equal(flat[gold["synth_idx"]], gold["synth_val"])
    # sub-range generation (the CPU baseline regenerates column slices)
    np.testing.assert_array_equal(orc.synth_rimatr(11, nbf, naux, 40, 7), b[:, 40:47])


# ---------------------------------------------------------------------------
# closed forms
# ---------------------------------------------------------------------------
def test_vj_naux1_closed_form():
    rng = np.random.default_rng(1)
    nbf = 9
    npair = nbf * (nbf + 1) // 2
    b = rng.standard_normal((npair, 1))
    d = rng.standard_normal((nbf, nbf))
    d = d + d.T
    g = 2 * np.dot(b[:, 0], unpack(np.ones(npair), nbf)[upper_idx(nbf)] * d[upper_idx(nbf)]) - \
        np.dot(b[[k * (k + 1) // 2 + k for k in range(nbf)], 0], np.diag(d))
    vj = orc.vj_upper_rimatr(b, [d], 1, 0.7)[0]
    np.testing.assert_allclose(vj, 0.7 * g * b[:, 0], rtol=1e-13, atol=1e-14)


def test_gamma_identity_dm_is_trace():
    rng = np.random.default_rng(2)
    nbf, naux = 11, 5
    npair = nbf * (nbf + 1) // 2
    b = rng.standard_normal((npair, naux))
    g = orc.gamma_rimatr(b, np.eye(nbf))
    diag = [k * (k + 1) // 2 + k for k in range(nbf)]
    np.testing.assert_allclose(g, b[diag, :].sum(axis=0), rtol=1e-14, atol=1e-14)


def test_vj_one_hot_b():
    nbf, naux = 6, 4
    npair = nbf * (nbf + 1) // 2
    rng = np.random.default_rng(3)
    d = rng.standard_normal((nbf, nbf))
    for p0, P0 in [(0, 0), (4, 2), (npair - 1, 3), (7, 1)]:
        b = np.zeros((npair, naux))
        b[p0, P0] = 1.0
        vj = orc.vj_upper_rimatr(b, [d], 1, 1.0)[0]
        expect = np.zeros(npair)
        expect[p0] = weights(d)[p0]
        np.testing.assert_allclose(vj, expect, rtol=0, atol=1e-15)


def test_vj_linearity_and_energy_identity():
    rng = np.random.default_rng(4)
    nbf, naux = 14, 23
    npair = nbf * (nbf + 1) // 2
    b = rng.standard_normal((npair, naux))
    d1 = rng.standard_normal((nbf, nbf)); d1 += d1.T
    d2 = rng.standard_normal((nbf, nbf)); d2 += d2.T
    j1 = orc.vj_upper_rimatr(b, [d1], 1, 1.0)[0]
    j2 = orc.vj_upper_rimatr(b, [d2], 1, 1.0)[0]
    j12 = orc.vj_upper_rimatr(b, [0.3 * d1 - 1.7 * d2], 1, 1.0)[0]
    np.testing.assert_allclose(j12, 0.3 * j1 - 1.7 * j2, rtol=1e-12, atol=1e-12)
    g = orc.gamma_rimatr(b, d1)
    np.testing.assert_allclose(np.dot(weights(d1), j1), np.dot(g, g), rtol=1e-12)


def test_vj_spin_channels_independent():
    rng = np.random.default_rng(5)
    nbf, naux = 8, 6
    npair = nbf * (nbf + 1) // 2
    b = rng.standard_normal((npair, naux))
    da = rng.standard_normal((nbf, nbf)); da += da.T
    db = rng.standard_normal((nbf, nbf)); db += db.T
    both = orc.vj_upper_rimatr(b, [da, db], 2, 2.0)
    np.testing.assert_allclose(both[0], orc.vj_upper_rimatr(b, [da], 1, 2.0)[0], rtol=1e-14)
    np.testing.assert_allclose(both[1], orc.vj_upper_rimatr(b, [db], 1, 2.0)[0], rtol=1e-14)


def test_vk_cross_route_identity_and_sign():
    """occupied route (C, occ) == density-only route with D = C occ Cᵀ; K is PSD for PSD D."""
    rng = np.random.default_rng(6)
    nbf, naux, nocc = 13, 19, 4
    npair = nbf * (nbf + 1) // 2
    b = rng.standard_normal((npair, naux))
    c = np.linalg.qr(rng.standard_normal((nbf, nocc)))[0]
    occ = np.array([2.0, 2.0, 1.0, 0.5])
    d = (c * occ) @ c.T
    ko = orc.vk_upper_occ(b, True, [c], [nocc - 1], [occ], 1, 1.0)[0]
    kd = orc.vk_upper_dm(b, True, [d], 1, 1.0)[0]
    np.testing.assert_allclose(ko, kd, rtol=1e-12, atol=1e-12)
    kfull = unpack(ko, nbf)
    assert np.linalg.eigvalsh(kfull).min() > -1e-10
    # exchange energy with the RHF scale is non-positive
    kr = orc.vk_upper_occ(b, True, [c], [nocc - 1], [occ], 1, -0.5)[0]
    assert orc.energy_contraction(d, kr) <= 0.0


def test_vk_homo_limits_columns():
    """only the first homo+1 columns of eigv are used (nw = homo + 1)"""
    rng = np.random.default_rng(7)
    nbf, naux, nstate = 10, 7, 6
    npair = nbf * (nbf + 1) // 2
    b = rng.standard_normal((npair, naux))
    c = rng.standard_normal((nbf, nstate))
    occ = np.full(nstate, 2.0)
    k2 = orc.vk_upper_occ(b, True, [c], [2], [occ], 1, 1.0)[0]
    c_cut = c.copy()
    c_cut[:, 3:] = 1e3  # garbage beyond homo must not matter
    np.testing.assert_array_equal(orc.vk_upper_occ(b, True, [c_cut], [2], [occ], 1, 1.0)[0], k2)
    d = (c[:, :3] * 2.0) @ c[:, :3].T
    np.testing.assert_allclose(k2, orc.vk_upper_dm(b, True, [d], 1, 1.0)[0], rtol=1e-11, atol=1e-11)


def test_vk_dm_one_hot_b():
    nbf, naux = 5, 2
    npair = nbf * (nbf + 1) // 2
    rng = np.random.default_rng(8)
    d = rng.standard_normal((nbf, nbf))
    b = np.zeros((npair, naux))
    b[1 * 2 // 2 + 0, 1] = 1.0   # (i=0, j=1)
    bp = np.zeros((nbf, nbf)); bp[0, 1] = bp[1, 0] = 1.0
    expect = (bp @ d @ bp)[upper_idx(nbf)]
    np.testing.assert_allclose(orc.vk_upper_dm(b, True, [d], 1, 1.0)[0], expect, atol=1e-15)


def test_density_matrix_and_idempotency():
    rng = np.random.default_rng(9)
    nbf, nstate = 12, 8
    c = np.linalg.qr(rng.standard_normal((nbf, nstate)))[0]
    occ = np.array([2.0] * 5 + [0.0] * 3)
    d = orc.density_matrix([c], [4], [occ], 1)[0]
    np.testing.assert_allclose(d, 2.0 * c[:, :5] @ c[:, :5].T, atol=1e-14)
    np.testing.assert_allclose(d @ d, 2.0 * d, atol=1e-13)   # S = I: DSD = 2D


def test_energy_contraction_matches_full_trace():
    rng = np.random.default_rng(10)
    n = 9
    d = rng.standard_normal((n, n)); d += d.T
    a = rng.standard_normal((n, n)); a += a.T
    np.testing.assert_allclose(orc.energy_contraction(d, a[upper_idx(n)]), np.sum(d * a), rtol=1e-13)


# ---------------------------------------------------------------------------
# generators
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("seed,n", [(1, 16), (2, 64), (3, 2), (77, 27)])
def test_water_cluster_invariants(seed, n):
    xyz = orc.water_cluster(seed, n)
    assert xyz.shape == (3 * n, 3)
    o = xyz[0::3]
    dist = np.linalg.norm(o[:, None, :] - o[None, :, :], axis=-1)
    assert dist[np.triu_indices(n, 1)].min() >= 2.5
    for w in range(n):
        h1, h2 = xyz[3 * w + 1] - o[w], xyz[3 * w + 2] - o[w]
        np.testing.assert_allclose([np.linalg.norm(h1), np.linalg.norm(h2)], np.hypot(0.7572, 0.5865), rtol=1e-12)
        cosang = np.dot(h1, h2) / (np.linalg.norm(h1) * np.linalg.norm(h2))
        ref = (-0.7572 ** 2 + 0.5865 ** 2) / (0.7572 ** 2 + 0.5865 ** 2)
        np.testing.assert_allclose(cosang, ref, rtol=1e-12)
    # O on its lattice site within the jitter
    k = 1
    while k ** 3 < n:
        k += 1
    sites = np.array([[w % k, (w // k) % k, w // (k * k)] for w in range(n)]) * 3.1
    assert np.abs(o - sites).max() <= 0.2 + 1e-12
    np.testing.assert_array_equal(orc.water_cluster(seed, n), xyz)  # deterministic


def test_water_cluster_perturbed_monomers():
    xyz = orc.water_cluster(5, 8, 0.05, 5.0)
    o = xyz[0::3]
    r = np.concatenate([np.linalg.norm(xyz[1::3] - o, axis=1), np.linalg.norm(xyz[2::3] - o, axis=1)])
    r0 = np.hypot(0.7572, 0.5865)
    assert np.all(np.abs(r - r0) <= 0.05 + 1e-12) and np.ptp(r) > 1e-3


def test_random_orthonormal():
    c = orc.random_orthonormal(1, 384, 80)
    np.testing.assert_allclose(c.T @ c, np.eye(80), atol=1e-13)
    np.testing.assert_array_equal(orc.random_orthonormal(1, 384, 80), c)
    assert not np.allclose(orc.random_orthonormal(2, 384, 80), c)


def test_numpy_synth_equals_c_synth_other_shape():
    np.testing.assert_array_equal(orc.synth_rimatr(3, 7, 5), np_synth_rimatr(3, 7, 5))


def _energy_functional(bfull, h, dms, uhf):
    """RI-HF electronic energy from unpacked B, computed directly:
    RHF  E = tr(DH) + ½Σ_P γ_P² − ¼Σ_P tr(D B_P D B_P)
    UHF  E = Σ_s tr(D_s H) + ½Σ_P (γ_α+γ_β)² − ½Σ_s Σ_P tr(D_s B_P D_s B_P)"""
    naux = bfull.shape[2]
    e = sum(np.sum(d * h) for d in dms)
    g = sum(np.einsum("ijP,ij->P", bfull, d) for d in dms)
    e += 0.5 * np.dot(g, g)
    kx = 0.0
    for d in dms:
        for P in range(naux):
            bp = bfull[:, :, P]
            kx += np.sum((bp @ d @ bp) * d)
    return e - (0.5 if uhf else 0.25) * kx


@pytest.mark.parametrize("uhf", [False, True])
def test_fock_is_energy_gradient(uhf):
    """F (H + J_0 + J_1 + s_K·K with s_K = −0.5 RHF / −1 UHF) is dE/dD of the RI
    energy functional: checked by central finite differences."""
    rng = np.random.default_rng(11 + uhf)
    nbf, naux = 7, 9
    npair = nbf * (nbf + 1) // 2
    b = rng.standard_normal((npair, naux)) * 0.3
    bfull = np.stack([unpack(b[:, P], nbf) for P in range(naux)], axis=2)
    h = rng.standard_normal((nbf, nbf)); h = h + h.T
    ns = 2 if uhf else 1
    dms = []
    for _ in range(ns):
        d = rng.standard_normal((nbf, nbf)); dms.append(d + d.T)
    sk = -1.0 if uhf else -0.5
    vj = orc.vj_upper_rimatr(b, dms, ns, 1.0)
    vk = orc.vk_upper_dm(b, True, dms, ns, sk)
    f = orc.fock_upper(h[upper_idx(nbf)], vj, vk, ns)
    for s in range(ns):
        dd = rng.standard_normal((nbf, nbf)); dd = dd + dd.T
        eps = 1e-5
        dp = [x.copy() for x in dms]; dp[s] = dp[s] + eps * dd
        dmn = [x.copy() for x in dms]; dmn[s] = dmn[s] - eps * dd
        fd = (_energy_functional(bfull, h, dp, uhf) - _energy_functional(bfull, h, dmn, uhf)) / (2 * eps)
        analytic = np.sum(unpack(f[s], nbf) * dd)
        np.testing.assert_allclose(analytic, fd, rtol=1e-7, atol=1e-7)


def test_cpu_baseline_routes_equal_the_oracle():
    """The reference-structured CPU J/K that bench.py times (per-P BLAS calls, serial accumulation)
    compute the same J, occupied-route K and density-only K as the compensated oracle."""
    nbf, naux, nocc = 37, 23, 6
    b = orc.synth_rimatr(5, nbf, naux)
    c = orc.random_orthonormal(6, nbf, nocc)
    occ = np.full(nocc, 2.0)
    d = (c * occ) @ c.T
    orc.load_host_blas()
    _, vj = orc.cpu_baseline_vj(b, nbf, d, 2)
    _, vk = orc.cpu_baseline_vk(b, nbf, c, occ, 2)
    _, vkd = orc.cpu_baseline_vk_dm(b, nbf, d, 2)
    rj = orc.vj_upper_rimatr(b, [d], 1, 1.0)[0]
    rk = orc.vk_upper_occ(b, True, [c], [nocc - 1], [occ], 1, 1.0)[0]
    rkd = orc.vk_upper_dm(b, True, [d], 1, 1.0)[0]
    for got, ref in ((vj, rj), (vk, rk), (vkd, rkd), (vkd, rk)):
        assert np.abs(got - ref).max() <= 1e-12 * np.abs(ref).max()

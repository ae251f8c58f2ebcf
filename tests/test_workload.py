"""The package's Python generators agree with the oracle's independent C ones."""
import numpy as np
import pytest

from oracle import oracle as orc
from rest_amd import workload as wl


@pytest.mark.parametrize("seed,n", [(1, 16), (2, 5), (9, 1)])
def test_water_cluster_matches_oracle(seed, n):
    np.testing.assert_allclose(wl.water_cluster(seed, n), orc.water_cluster(seed, n), rtol=0, atol=1e-13)


def test_perturbed_cluster_matches_oracle():
    np.testing.assert_allclose(wl.water_cluster(4, 6, 0.05, 5.0), orc.water_cluster(4, 6, 0.05, 5.0), atol=1e-13)


def test_orthonormal_matches_oracle():
    a = wl.random_orthonormal(3, 96, 20)
    b = orc.random_orthonormal(3, 96, 20)
    np.testing.assert_allclose(a, b, rtol=0, atol=1e-12)


def test_shapes():
    assert (wl.H2O.nbf, wl.H2O.naux, wl.H2O.npair, wl.H2O.nocc) == (24, 113, 300, 5)
    assert (wl.H2O_16.nbf, wl.H2O_16.naux, wl.H2O_16.npair) == (384, 1808, 73920)
    assert (wl.H2O_64.nbf, wl.H2O_64.naux, wl.H2O_64.npair, wl.H2O_64.nocc) == (1536, 7232, 1180416, 320)
    # algorithmic work figures quoted for these shapes
    assert abs(wl.jk_bytes(wl.H2O_64, 8) - 136.59e9) / 136.59e9 < 1e-3
    assert abs(wl.k_flops(wl.H2O_64)["total"] - 1.638e13) / 1.638e13 < 1e-3
    assert abs(wl.k_flops(wl.H2O_16)["total"] - 6.404e10) / 6.404e10 < 1e-3


def _read_xyz(name):
    from pathlib import Path
    lines = (Path(__file__).resolve().parent / "golden" / name).read_text().splitlines()[2:]
    return [l.split()[0] for l in lines], np.array([[float(v) for v in l.split()[1:]] for l in lines])


def test_benchmark_geometries_match_fixtures():
    """The (H2O)16 seed-1 and (H2O)64 seed-2 clusters the benchmarks run are the committed .xyz fixtures;
    every monomer keeps the reference geometry (O–H 0.95777559 Å, H–O–H 104.51°) and no two oxygens
    are closer than 2.5 Å."""
    for name, seed, n in (("h2o16_seed1.xyz", 1, 16), ("h2o64_seed2.xyz", 2, 64)):
        sym, ref = _read_xyz(name)
        xyz = wl.water_cluster(seed, n)
        assert sym == ["O", "H", "H"] * n
        np.testing.assert_allclose(xyz, ref, rtol=0, atol=1e-11)
        o, h1, h2 = xyz[0::3], xyz[1::3], xyz[2::3]
        d1, d2 = np.linalg.norm(h1 - o, axis=1), np.linalg.norm(h2 - o, axis=1)
        np.testing.assert_allclose(np.r_[d1, d2], np.hypot(0.7572, 0.5865), atol=1e-12)
        cos = np.sum((h1 - o) * (h2 - o), axis=1) / (d1 * d2)
        np.testing.assert_allclose(np.degrees(np.arccos(cos)), np.degrees(2 * np.arctan2(0.7572, 0.5865)),
                                   atol=1e-9)
        doo = np.linalg.norm(o[:, None] - o[None], axis=2) + np.eye(n) * 1e9
        assert doo.min() >= 2.5
    sym, mono = _read_xyz("h2o_monomer.xyz")
    assert sym == ["O", "H", "H"] and np.allclose(mono[1], [0.7572, 0.5865, 0.0])

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

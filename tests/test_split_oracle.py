"""Pins of the split-BF16 oracle (oracle/split_oracle.py) on the CPU."""
import numpy as np

from oracle import oracle as orc
from oracle import split_oracle as so


def test_bf16_rounding_known_values():
    # 1 + 2^-8 is a tie between 1 and 1 + 2^-7: ties go to the even mantissa (1.0);
    # 1 + 3·2^-8 ties between 1 + 2^-7 and 1 + 2^-6 and goes to 1 + 2^-6
    x = np.array([1.0, 1 + 2.0**-8, 1 + 3 * 2.0**-8, -2.5, 3.0e38, 1e-30])
    r = so.bf16(x)
    assert r[0] == 1.0 and r[1] == 1.0 and r[2] == 1 + 2.0**-6 and r[3] == -2.5
    assert abs(r[5] - 1e-30) <= 2.0**-9 * 1e-30


def test_planes_sum_back_to_x():
    rng = np.random.default_rng(3)
    x = rng.standard_normal(10000) * 10.0 ** rng.uniform(-6, 3, 10000)
    for s in (1, 2, 3, 4):
        pl = so.split_planes(x, s)
        rel = np.max(np.abs(sum(pl) - x) / np.abs(x))
        # 8 significant bits per plane, half an ulp of the leading bit each: residual <= 2^-8S
        assert rel <= 2.0 ** -(8 * s) * 1.01, (s, rel)
    # values with at most 8 significant bits are one exact plane
    y = np.array([0.75, -3.0, 96.0, 2.0**-20 * 255])
    pl = so.split_planes(y, 3)
    assert np.array_equal(pl[0], y) and not pl[1].any() and not pl[2].any()


def test_plane_pairs():
    assert so.pairs(2) == [(0, 0), (0, 1), (1, 0)]
    assert so.pairs(3) == [(0, 0), (0, 1), (1, 0), (0, 2), (1, 1), (2, 0)]


def test_emulated_k_approaches_float64_k():
    nbf, naux, nocc = 24, 20, 5
    b = orc.synth_rimatr(7, nbf, naux)
    c = orc.random_orthonormal(3, nbf, nocc)
    occ = np.array([2.0, 2.0, 1.5, -0.5, 2.0])
    ref = orc.vk_upper_occ(b, True, [c], [nocc - 1], [occ], 1, -0.5)[0]
    scale = np.abs(ref).max()
    errs = {s: np.abs(so.vk_split(b, c, occ, s, -0.5) - ref).max() / scale for s in (1, 2, 3, 4)}
    assert errs[1] > 1e-4                      # one plane: BF16 alone
    assert 1e-7 < errs[2] < 1e-4               # hi/lo
    assert errs[3] < 3e-7                      # three planes: FP32-level (the M' split of the FP32 value)
    assert errs[4] <= errs[3] * 1.5            # no worse with a fourth plane (FP32 M' bounds it)
    assert errs[2] > 10 * errs[3]
    # exact inputs: B and C' with few significant bits make S = 2 exact up to the M' rounding
    bq = np.round(b * 64) / 64
    cq = np.round(c * 16) / 16
    occ1 = np.ones(nocc)
    refq = orc.vk_upper_occ(bq, True, [cq], [nocc - 1], [occ1], 1, 1.0)[0]
    assert np.abs(so.vk_split(bq, cq, occ1, 3, 1.0) - refq).max() <= 1e-6 * np.abs(refq).max()

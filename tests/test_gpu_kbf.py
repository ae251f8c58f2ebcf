"""GPU K on split BF16 planes (RJK_PREC_BF16X2 / RJK_PREC_BF16X3) through the C ABI.

Bars:
  * BF16X2 against the split-plane oracle (oracle/split_oracle.py, same plane
    roundings and plane pairs, float64 sums): the device differs from it by
    FP32 accumulation only, so |K_dev − K_emul| must stay below a tenth of the
    oracle's own distance from the float64 K (a dropped or extra plane pair
    fails this);
  * BF16X3 against the float64 oracle: relative max error <= 2e-6 (the three
    planes carry 24 significant bits; FP32 accumulation over at most 4096
    products per split before the FP64 sum);
  * the same shapes as the FP64 suite's ragged, two-spin, signed-occupation,
    density-only and chunked cases at that relative bar.
"""
import sys

import numpy as np
import pytest

from oracle import oracle as orc
from oracle import split_oracle as so
from rest_amd import jk

pytestmark = pytest.mark.gpu
REL3 = 2e-6
STORE = jk.STORE_FP64 | jk.STORE_FP32 | jk.STORE_BF16


def say(msg):
    print(f"[gpu-case] {msg}", flush=True)
    sys.stdout.flush()


def rel(a, ref):
    return float(np.max(np.abs(np.asarray(a) - ref)) / max(np.abs(ref).max(), 1e-300))


def test_bf16x2_matches_split_oracle_and_x3_is_fp32_level():
    nbf, naux, nocc = 130, 40, 17
    say(f"BF16 split K nbf={nbf} naux={naux} nocc={nocc}")
    b = orc.synth_rimatr(21, nbf, naux)
    c = orc.random_orthonormal(5, nbf, nocc)
    occ = np.full(nocc, 2.0)
    ref = orc.vk_upper_occ(b, True, [c], [nocc - 1], [occ], 1, -0.5)[0]
    emul2 = so.vk_split(b, c, occ, 2, -0.5)
    with jk.DeviceRI(nbf, naux) as ri:
        ri.fill_synthetic(21, STORE)
        k2 = ri.vk_occ([c], [nocc - 1], [occ], 1, -0.5, prec=jk.PREC_BF16X2)[0]
        k3 = ri.vk_occ([c], [nocc - 1], [occ], 1, -0.5, prec=jk.PREC_BF16X3)[0]
    e_oracle = rel(emul2, ref)
    e_dev = rel(k2, emul2)
    e3 = rel(k3, ref)
    say(f"  x2: oracle vs FP64 {e_oracle:.2e}, device vs oracle {e_dev:.2e}; x3 vs FP64 {e3:.2e}")
    assert e_dev < 0.1 * e_oracle
    assert e3 < REL3


def test_bf16x1_matches_split_oracle():
    """One BF16 plane per operand, one product: the device equals the one-plane emulation up to FP32
    accumulation (a tenth of the emulation's distance from FP64), and sits near 2^-8 of FP64."""
    nbf, naux, nocc = 130, 40, 17
    say(f"BF16X1 split K nbf={nbf} naux={naux} nocc={nocc}")
    b = orc.synth_rimatr(21, nbf, naux)
    c = orc.random_orthonormal(5, nbf, nocc)
    occ = np.full(nocc, 2.0)
    ref = orc.vk_upper_occ(b, True, [c], [nocc - 1], [occ], 1, -0.5)[0]
    emul1 = so.vk_split(b, c, occ, 1, -0.5)
    with jk.DeviceRI(nbf, naux) as ri:
        ri.fill_synthetic(21, STORE)
        k1 = ri.vk_occ([c], [nocc - 1], [occ], 1, -0.5, prec=jk.PREC_BF16X1)[0]
    e_oracle, e_dev = rel(emul1, ref), rel(k1, emul1)
    say(f"  x1: oracle vs FP64 {e_oracle:.2e}, device vs oracle {e_dev:.2e}")
    assert e_dev < 0.1 * e_oracle and e_oracle < 2e-2


@pytest.mark.parametrize("nbf,naux,nw,ns", [
    (1, 1, (1, 1), 1), (5, 3, (2, 1), 2), (17, 29, (3, 3), 1), (130, 77, (17, 5), 2),
    (48, 226, (10, 9), 2), (129, 40, (33, 31), 1), (200, 50, (40, 1), 2),
])
def test_bf16x3_ragged_shapes(nbf, naux, nw, ns):
    say(f"BF16X3 ragged nbf={nbf} naux={naux} nw={nw} ns={ns}")
    rng = np.random.default_rng(nbf * 1000 + naux)
    npair = nbf * (nbf + 1) // 2
    b = rng.standard_normal((npair, naux)) / np.sqrt(naux)
    nstate = max(nw) + 2
    cs, occs = [], []
    for s in range(ns):
        cs.append(rng.standard_normal((nbf, nstate)) / np.sqrt(nbf))
        occ = rng.uniform(0.2, 2.0, nstate)
        if nw[s] > 2:
            occ[1] = 0.0
            occ[2] = -0.7
        occs.append(occ)
    homo = [nw[s] - 1 for s in range(ns)]
    with jk.DeviceRI(nbf, naux) as ri:
        ri.set_rimatr(b, STORE)
        vk = ri.vk_occ(cs, homo, occs, ns, -1.0, prec=jk.PREC_BF16X3)
    ref = orc.vk_upper_occ(b, True, cs, homo, occs, ns, -1.0)
    for s in range(ns):
        e = rel(vk[s], ref[s])
        say(f"  spin {s}: rel {e:.2e}")
        assert e < REL3


def test_bf16x3_density_only_and_chunking():
    nbf, naux, nocc = 200, 333, 21
    say(f"BF16X3 density-only + chunked nbf={nbf} naux={naux}")
    rng = np.random.default_rng(5)
    npair = nbf * (nbf + 1) // 2
    b = rng.standard_normal((npair, naux)) / np.sqrt(naux)
    c = np.linalg.qr(rng.standard_normal((nbf, nocc)))[0]
    occ = np.full(nocc, 2.0)
    d = (c * occ) @ c.T
    ref = orc.vk_upper_occ(b, True, [c], [nocc - 1], [occ], 1, 1.0)[0]
    with jk.DeviceRI(nbf, naux) as ri:
        ri.set_rimatr(b, STORE)
        one = ri.vk_occ([c], [nocc - 1], [occ], 1, 1.0, prec=jk.PREC_BF16X3)[0]
        kd = ri.vk_dm([d], 1, 1.0, prec=jk.PREC_BF16X3)[0]
        ri.set_workspace_limit(8 << 20)   # forces many P chunks
        many = ri.vk_occ([c], [nocc - 1], [occ], 1, 1.0, prec=jk.PREC_BF16X3)[0]
    for name, k in (("single pass", one), ("density-only", kd), ("chunked", many)):
        e = rel(k, ref)
        say(f"  {name}: rel {e:.2e}")
        assert e < REL3


def test_bf16x3_h2o16_shape():
    nbf, naux, nocc = 384, 1808, 80
    say(f"BF16X3 (H2O)16 shape nbf={nbf} naux={naux} nocc={nocc}")
    b = orc.synth_rimatr(101, nbf, naux)
    c = orc.random_orthonormal(1, nbf, nocc)
    occ = np.full(nocc, 2.0)
    ref = orc.vk_upper_occ(b, True, [c], [nocc - 1], [occ], 1, -0.5)[0]
    with jk.DeviceRI(nbf, naux) as ri:
        ri.fill_synthetic(101, STORE)
        k3 = ri.vk_occ([c], [nocc - 1], [occ], 1, -0.5, prec=jk.PREC_BF16X3)[0]
        k64 = ri.vk_occ([c], [nocc - 1], [occ], 1, -0.5)[0]
    assert np.max(np.abs(k64 - ref)) <= 1e-11
    e = rel(k3, ref)
    say(f"  rel {e:.2e}")
    assert e < REL3


def test_bf16_j_runs_on_the_stored_copy_and_error_paths():
    nbf, naux = 40, 60
    say(f"BF16 J copy selection nbf={nbf} naux={naux}")
    b = orc.synth_rimatr(9, nbf, naux)
    c = orc.random_orthonormal(4, nbf, 6)
    d = 2.0 * c @ c.T
    with jk.DeviceRI(nbf, naux) as ri:
        ri.fill_synthetic(9, jk.STORE_FP64 | jk.STORE_BF16)
        # no FP32 copy: J at a BF16 precision runs on the FP64 copy
        j = ri.vj([d], 1, 1.0, prec=jk.PREC_BF16X3)[0]
        assert np.max(np.abs(j - orc.vj_upper_rimatr(b, [d], 1, 1.0)[0])) <= 1e-11
        ri.fill_synthetic(9, STORE)
        j3 = ri.vj([d], 1, 1.0, prec=jk.PREC_BF16X3)[0]
        j32 = ri.vj([d], 1, 1.0, prec=jk.PREC_FP32)[0]
        assert np.array_equal(j3, j32)
        ri.fill_synthetic(9, jk.STORE_FP64)
        with pytest.raises(jk.RJKError) as e:
            ri.vk_occ([c], [5], [np.full(6, 2.0)], 1, 1.0, prec=jk.PREC_BF16X3)
        assert e.value.code == 4   # planes not stored
        with pytest.raises(jk.RJKError) as e:
            ri.vk_occ([c], [5], [np.full(6, 2.0)], 1, 1.0, prec=7)
        assert e.value.code == 1


def test_profile_rows_per_precision():
    """rjk_profile_read_detail keeps one row per RJK_PREC_* value; K work is the algorithmic FLOP of the
    padded columns (2·nbf²·nk·naux half-transform); J at a BF16 precision runs and is counted as FP32."""
    nbf, naux, nocc = 130, 40, 17
    say(f"profile rows nbf={nbf} naux={naux}")
    c = orc.random_orthonormal(5, nbf, nocc)
    occ = np.full(nocc, 2.0)
    with jk.DeviceRI(nbf, naux) as ri:
        ri.fill_synthetic(21, STORE)
        ri.profile_enable(True)
        ri.profile_read_detail()
        for p in (jk.PREC_FP64, jk.PREC_BF16X3, jk.PREC_BF16X2, jk.PREC_BF16X1):
            ri.vk_occ([c], [nocc - 1], [occ], 1, 1.0, prec=p)
        ri.vj([c @ c.T], 1, 1.0, prec=jk.PREC_BF16X3)
        d = ri.profile_read_detail()
    nk = 64   # 17 columns, padded to the column granule and then to 64
    for name in ("fp64", "bf16x3", "bf16x2", "bf16x1"):
        ms, n, work = d[(name, "k_half")]
        assert n == 1 and ms > 0 and work == pytest.approx(2.0 * nbf * nbf * nk * naux), name
        assert d[(name, "k_syrk")][1] == 1
    assert d[("fp32", "j_gamma")][1] == 1 and d[("fp32", "k_half")][1] == 0
    assert d[("fp64", "j_gamma")][1] == 0

"""GPU parity of the HIP J/K library against the CPU oracle (through the C ABI).

Bar (FP64): max-abs difference <= 1e-11 against the oracle on the same B, D
and C.  Shapes: the one-water fixture (nbf 24, naux 113), (H2O)16 dimensions
(nbf 384, naux 1808, nocc 80, D = C Cᵀ from the seeded orthonormal C), ragged
and degenerate shapes, both spin channels, signed occupations, the
density-only route, workspace chunking, and at the (H2O)64 dimensions
(nbf 1536, naux 7232) the closed form of a rank-3 B.
"""
import sys
from pathlib import Path

import numpy as np
import pytest

from oracle import oracle as orc
from rest_amd import jk

pytestmark = pytest.mark.gpu
GOLD = Path(__file__).resolve().parent / "golden" / "jk_h2o_shape.npz"
TOL = 1e-11


def say(msg):
    print(f"[gpu-case] {msg}", flush=True)
    sys.stdout.flush()


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


def maxabs(a, b):
    return float(np.max(np.abs(np.asarray(a) - np.asarray(b))))


# ---------------------------------------------------------------------------
def test_h2o_shape_golden(gold):
    nbf, naux = int(gold["nbf"]), int(gold["naux"])
    say(f"h2o golden nbf={nbf} naux={naux}")
    with jk.DeviceRI(nbf, naux) as ri:
        ri.set_rimatr(gold["rimatr"])
        vj = jk.vj_upper_with_rimatr_v_sync(ri, [gold["dm"]], 1, 1.0)
        assert vj[1].shape == (1,)
        assert maxabs(vj[0], gold["vj"]) <= TOL
        assert maxabs(ri.vj([gold["dm_asym"]], 1, 1.0)[0], gold["vj_asym_rimatr"]) <= TOL
        c, occ = gold["c"], gold["occ"]
        vk = jk.vk_upper_with_ri_v_sync(ri, [c], [c.shape[1] - 1, 0], [occ], 1, -0.5)
        assert maxabs(vk[0], gold["vk_rhf"]) <= TOL
        vkd = jk.vk_upper_with_rimatr_use_dm_only_sync(ri, [gold["dm"]], 1, -0.5)
        assert maxabs(vkd[0], gold["vk_rhf"]) <= TOL
        # REST's density-only K reads a general D in full (src/scf_io/mod.rs:2730-2732)
        vka = jk.vk_upper_with_rimatr_use_dm_only_sync(ri, [gold["dm_asym"]], 1, 1.0)
        assert maxabs(vka[0], gold["vk_dm_asym"]) <= TOL
        ca, cb = gold["ca"], gold["cb"]
        cbp = np.zeros_like(ca)
        cbp[:, :cb.shape[1]] = cb
        occs = [np.ones(ca.shape[1]), np.concatenate([np.ones(cb.shape[1]), [0.0]])]
        vku = ri.vk_occ([ca, cbp], [ca.shape[1] - 1, cb.shape[1] - 1], occs, 2, -1.0)
        assert maxabs(vku[0], gold["vk_uhf_a"]) <= TOL
        assert maxabs(vku[1], gold["vk_uhf_b"]) <= TOL


def test_h2o_shape_ri3fn_entry_points(gold):
    nbf, naux = int(gold["nbf"]), int(gold["naux"])
    say("h2o ri3fn upload + full-D J")
    b3 = orc.ri3fn_from_rimatr(gold["rimatr"], nbf)
    with jk.DeviceRI(nbf, naux) as ri:
        ri.set_ri3fn(b3)
        vj = jk.vj_upper_with_ri_v_sync(ri, [gold["dm_asym"]], 1, 1.0)
        assert maxabs(vj[0], gold["vj_asym_ri3fn"]) <= TOL
        vk = jk.vk_upper_with_ri_v_use_dm_only_sync(ri, [gold["dm"]], 1, -0.5)
        assert maxabs(vk[0], gold["vk_rhf"]) <= TOL
        vka = jk.vk_upper_with_ri_v_use_dm_only_sync(ri, [gold["dm_asym"]], 1, 1.0)
        assert maxabs(vka[0], gold["vk_dm_asym"]) <= TOL


def test_density_only_general_d_ragged():
    """General (asymmetric) D through the density-only route: both spins, a ragged shape whose
    nbf is no multiple of any tile, several auxiliary chunks (small workspace), against the oracle."""
    nbf, naux = 77, 131
    say(f"density-only general D nbf={nbf} naux={naux}")
    rng = np.random.default_rng(81)
    npair = nbf * (nbf + 1) // 2
    b = rng.standard_normal((npair, naux)) / np.sqrt(naux)
    da = rng.standard_normal((nbf, nbf))
    db = rng.standard_normal((nbf, nbf))
    db = db @ db.T / nbf
    db[3, 5] += 0.25                                   # one asymmetric spin makes the call general
    ref = orc.vk_upper_dm(b, True, [da, db], 2, -1.0)
    with jk.DeviceRI(nbf, naux) as ri:
        ri.set_rimatr(b)
        ri.set_workspace_limit(6 << 20)                # forces several auxiliary chunks
        got = ri.vk_dm([da, db], 2, -1.0)
        assert maxabs(got, ref) <= TOL
        # a requested low precision does not change the general route (FP64)
        ri.set_rimatr(b, jk.STORE_FP64 | jk.STORE_FP32 | jk.STORE_BF16)
        assert maxabs(ri.vk_dm([da, db], 2, -1.0, prec=jk.PREC_BF16X2), ref) <= TOL


def test_device_synthetic_fill_equals_host_values(gold):
    nbf, naux = int(gold["nbf"]), int(gold["naux"])
    say("synthetic fill")
    with jk.DeviceRI(nbf, naux) as a, jk.DeviceRI(nbf, naux) as b:
        a.fill_synthetic(11)
        b.set_rimatr(gold["rimatr"])
        d = gold["dm"]
        np.testing.assert_array_equal(a.vj([d], 1, 1.0), b.vj([d], 1, 1.0))


def c2_inputs():
    nbf, naux, nocc = 384, 1808, 80
    c = orc.random_orthonormal(1, nbf, nocc)
    return nbf, naux, nocc, c, c @ c.T


def test_h2o16_shape_parity():
    nbf, naux, nocc, c, d = c2_inputs()
    say(f"(H2O)16 shape nbf={nbf} naux={naux} nocc={nocc}")
    b = orc.synth_rimatr(101, nbf, naux)
    with jk.DeviceRI(nbf, naux) as ri:
        ri.fill_synthetic(101)
        vj = ri.vj([d], 1, 1.0)[0]
        ref_j = orc.vj_upper_rimatr(b, [d], 1, 1.0)[0]
        err_j = maxabs(vj, ref_j)
        say(f"  J max-abs err {err_j:.3e} (max|J| {np.abs(ref_j).max():.3e})")
        assert err_j <= TOL
        occ = np.ones(nocc)
        vk = ri.vk_occ([c], [nocc - 1], [occ], 1, 1.0)[0]
        ref_k = orc.vk_upper_occ(b, True, [c], [nocc - 1], [occ], 1, 1.0)[0]
        err_k = maxabs(vk, ref_k)
        say(f"  K max-abs err {err_k:.3e} (max|K| {np.abs(ref_k).max():.3e})")
        assert err_k <= TOL
        # density-only route on the same D (cross-route identity on the device)
        vkd = ri.vk_dm([d], 1, 1.0)[0]
        assert maxabs(vkd, ref_k) <= TOL


@pytest.mark.parametrize("nbf,naux,nw,ns", [
    (1, 1, (1, 1), 1), (5, 3, (2, 1), 2), (17, 29, (3, 3), 1), (130, 77, (17, 5), 2),
    (48, 226, (10, 9), 2), (129, 40, (33, 31), 1),
])
def test_ragged_shapes(nbf, naux, nw, ns):
    say(f"ragged nbf={nbf} naux={naux} nw={nw} ns={ns}")
    rng = np.random.default_rng(nbf * 1000 + naux)
    npair = nbf * (nbf + 1) // 2
    b = rng.standard_normal((npair, naux)) / np.sqrt(naux)
    nstate = max(nw) + 2
    cs, occs, dms = [], [], []
    for s in range(ns):
        c = rng.standard_normal((nbf, nstate)) / np.sqrt(nbf)
        occ = rng.uniform(0.2, 2.0, nstate)
        if nw[s] > 2:
            occ[1] = 0.0          # empty orbital inside the occupied block
            occ[2] = -0.7         # signed weight (density-only route produces these)
        cs.append(c)
        occs.append(occ)
        d = rng.standard_normal((nbf, nbf))
        dms.append(d + d.T)
    homo = [nw[s] - 1 for s in range(ns)]
    with jk.DeviceRI(nbf, naux) as ri:
        ri.set_rimatr(b)
        assert maxabs(ri.vj(dms, ns, 0.5), orc.vj_upper_rimatr(b, dms, ns, 0.5)) <= TOL
        assert maxabs(ri.vj(dms, ns, 1.0, full_dm=True),
                      orc.vj_upper_ri3fn(orc.ri3fn_from_rimatr(b, nbf), dms, ns, 1.0)) <= TOL
        vk = ri.vk_occ(cs, homo, occs, ns, -1.0)
        ref = orc.vk_upper_occ(b, True, cs, homo, occs, ns, -1.0)
        assert maxabs(vk, ref) <= TOL


def test_density_only_route_indefinite_d():
    nbf, naux = 40, 90
    say(f"density-only indefinite D nbf={nbf} naux={naux}")
    rng = np.random.default_rng(44)
    npair = nbf * (nbf + 1) // 2
    b = rng.standard_normal((npair, naux)) / np.sqrt(naux)
    d = rng.standard_normal((nbf, nbf))
    d = 0.5 * (d + d.T)
    c = np.linalg.qr(rng.standard_normal((nbf, 7)))[0]
    dp = 2.0 * c @ c.T
    with jk.DeviceRI(nbf, naux) as ri:
        ri.set_rimatr(b)
        for dm in (d, dp):
            vk = ri.vk_dm([dm, dp], 2, 1.0)
            ref = orc.vk_upper_dm(b, True, [dm, dp], 2, 1.0)
            assert maxabs(vk, ref) <= TOL


def test_density_only_device_route():
    """rjk_vk_upper_dm_dev: the density-only route on device memory equals the host entry point
    bit for bit and matches the oracle, for a symmetric D and for an asymmetric one (general route)."""
    import torch
    nbf, naux = 40, 90
    say(f"density-only device route nbf={nbf} naux={naux}")
    rng = np.random.default_rng(45)
    npair = nbf * (nbf + 1) // 2
    b = rng.standard_normal((npair, naux)) / np.sqrt(naux)
    d = rng.standard_normal((nbf, nbf))
    d = 0.5 * (d + d.T)
    skew = 1e-3 * rng.standard_normal((nbf, nbf))
    skew = skew - skew.T
    dev = torch.device("cuda:0")
    with jk.DeviceRI(nbf, naux) as ri:
        ri.set_rimatr(b)
        host = ri.vk_dm([d], 1, 0.5)
        out = torch.empty((1, npair), dtype=torch.float64, device=dev)
        host_skew = ri.vk_dm([d + skew], 1, 0.5)
        for dm, want in ((d, host), (d + skew, host_skew)):
            dm_dev = torch.from_numpy(np.ascontiguousarray(dm.T)).to(dev)   # column-major
            torch.cuda.synchronize()
            ri.vk_dm_dev(dm_dev.data_ptr(), 1, 0.5, out.data_ptr())
            ri.synchronize()
            assert np.array_equal(out.cpu().numpy(), want)
        assert maxabs(host, orc.vk_upper_dm(b, True, [d], 1, 0.5)) <= TOL
        assert maxabs(host_skew, orc.vk_upper_dm(b, True, [d + skew], 1, 0.5)) <= TOL
        assert maxabs(host_skew, host) > 1e-6          # the skew part is not dropped
        # two spins in one call (spin-major device buffers), FP64 and three BF16 planes
        d2 = rng.standard_normal((nbf, nbf))
        d2 = d2 @ d2.T / nbf
        ref2 = orc.vk_upper_dm(b, True, [d, d2], 2, -1.0)
        both = torch.from_numpy(np.ascontiguousarray(np.stack([d.T, d2.T]))).to(dev)
        out2 = torch.empty((2, npair), dtype=torch.float64, device=dev)
        torch.cuda.synchronize()
        ri.vk_dm_dev(both.data_ptr(), 2, -1.0, out2.data_ptr())
        ri.synchronize()
        assert maxabs(out2.cpu().numpy(), ref2) <= TOL
        ri.set_rimatr(b, jk.STORE_FP64 | jk.STORE_FP32 | jk.STORE_BF16)
        ri.vk_dm_dev(both.data_ptr(), 2, -1.0, out2.data_ptr(), prec=jk.PREC_BF16X3)
        ri.synchronize()
        k3 = out2.cpu().numpy()
        for s_ in range(2):
            assert np.max(np.abs(k3[s_] - ref2[s_])) / np.abs(ref2[s_]).max() < 2e-6


def test_workspace_chunking_matches_single_pass():
    nbf, naux, nocc = 200, 333, 21
    say(f"chunked K nbf={nbf} naux={naux}")
    rng = np.random.default_rng(5)
    npair = nbf * (nbf + 1) // 2
    b = rng.standard_normal((npair, naux)) / np.sqrt(naux)
    c = np.linalg.qr(rng.standard_normal((nbf, nocc)))[0]
    occ = np.full(nocc, 2.0)
    with jk.DeviceRI(nbf, naux) as ri:
        ri.set_rimatr(b)
        one = ri.vk_occ([c], [nocc - 1], [occ], 1, 1.0)
        ri.set_workspace_limit(8 << 20)   # forces many P chunks
        many = ri.vk_occ([c], [nocc - 1], [occ], 1, 1.0)
        ref = orc.vk_upper_occ(b, True, [c], [nocc - 1], [occ], 1, 1.0)
        assert maxabs(one, ref) <= TOL and maxabs(many, ref) <= TOL


def test_fp32_copy_j_error_is_stated():
    nbf, naux = 96, 452
    say(f"FP32 J nbf={nbf} naux={naux}")
    b = orc.synth_rimatr(9, nbf, naux)
    c = orc.random_orthonormal(4, nbf, 20)
    d = 2.0 * c @ c.T
    with jk.DeviceRI(nbf, naux) as ri:
        ri.fill_synthetic(9, jk.STORE_FP64 | jk.STORE_FP32)
        j32 = ri.vj([d], 1, 1.0, prec=jk.PREC_FP32)[0]
        ref = orc.vj_upper_rimatr(b, [d], 1, 1.0)[0]
        rel = maxabs(j32, ref) / np.abs(ref).max()
        say(f"  FP32-B J relative max error {rel:.2e}")
        assert rel < 5e-6


def test_fp32_k_error_is_stated():
    """FP32 B copy, FP32 C', FP32 MFMA with FP32 accumulation per split-K slot,
    FP64 partial tiles: relative max error vs the FP64 oracle is stated."""
    nbf, naux, nocc = 130, 333, 17
    say(f"FP32 K nbf={nbf} naux={naux} nocc={nocc}")
    b = orc.synth_rimatr(21, nbf, naux)
    c = orc.random_orthonormal(5, nbf, nocc)
    occ = np.full(nocc, 2.0)
    with jk.DeviceRI(nbf, naux) as ri:
        ri.fill_synthetic(21, jk.STORE_FP64 | jk.STORE_FP32)
        k32 = ri.vk_occ([c], [nocc - 1], [occ], 1, -0.5, prec=jk.PREC_FP32)[0]
        k64 = ri.vk_occ([c], [nocc - 1], [occ], 1, -0.5)[0]
        ref = orc.vk_upper_occ(b, True, [c], [nocc - 1], [occ], 1, -0.5)[0]
        assert maxabs(k64, ref) <= TOL
        rel = maxabs(k32, ref) / np.abs(ref).max()
        say(f"  FP32 K relative max error {rel:.2e}")
        assert rel < 2e-5
        d = (c * occ) @ c.T
        kd32 = ri.vk_dm([d], 1, -0.5, prec=jk.PREC_FP32)[0]
        assert maxabs(kd32, ref) / np.abs(ref).max() < 2e-5


def test_error_paths():
    say("error paths")
    nbf, naux = 6, 4
    with jk.DeviceRI(nbf, naux) as ri:
        d = np.eye(nbf)
        with pytest.raises(jk.RJKError) as e:
            ri.vj([d], 1, 1.0)
        assert e.value.code == 4  # B not set
        ri.set_rimatr(np.ones((nbf * (nbf + 1) // 2, naux)))
        with pytest.raises(jk.RJKError) as e:
            ri.vj([d, d, d], 3, 1.0)
        assert e.value.code == 1
        bad = d.copy()
        bad[1, 2] = np.nan
        with pytest.raises(jk.RJKError) as e:
            ri.vj([bad], 1, 1.0)
        assert e.value.code == 5
        with pytest.raises(jk.RJKError) as e:
            ri.vj([d], 1, 1.0, prec=jk.PREC_FP32)
        assert e.value.code == 4  # FP32 copy not stored
        with pytest.raises(jk.RJKError) as e:
            ri.vk_occ([np.eye(nbf)], [nbf], [np.ones(nbf)], 1, 1.0)
        assert e.value.code == 1  # homo >= num_state
        asym = np.eye(nbf)
        asym[0, 1] = 1.0
        assert np.isfinite(ri.vk_dm([asym], 1, 1.0)).all()   # general D: no error
        with pytest.raises(jk.RJKError) as e:
            ri.vk_dm([bad], 1, 1.0)
        assert e.value.code == 5
        # the context still works after errors
        assert np.isfinite(ri.vj([d], 1, 1.0)).all()


def test_h2o64_shape_rank3_closed_form():
    """(H2O)64 dimensions: B = U Vᵀ (rank 3) makes J and K closed-form:
    γ = V (Uᵀ w), J = U (VᵀV)(Uᵀ w), K = Σ_ab (VᵀV)_ab U_a D U_b."""
    import torch
    nbf, naux, nocc = 1536, 7232, 320
    npair = nbf * (nbf + 1) // 2
    say(f"(H2O)64 shape rank-3 nbf={nbf} naux={naux} nocc={nocc}")
    rng = np.random.default_rng(64)
    U = rng.standard_normal((npair, 3)) / np.sqrt(nbf)
    V = rng.standard_normal((naux, 3)) * np.sqrt(1.0 / naux)
    c = orc.random_orthonormal(2, nbf, nocc)
    d = c @ c.T
    dev = torch.device("cuda:0")
    with jk.DeviceRI(nbf, naux) as ri:
        bt = torch.from_numpy(V).to(dev) @ torch.from_numpy(U).to(dev).T    # (naux, npair) = [P][p]
        bt = bt.contiguous()
        torch.cuda.synchronize()
        ri.set_rimatr_dev(bt.data_ptr())
        del bt
        torch.cuda.empty_cache()
        i, j = upper_idx(nbf)
        w = np.where(i == j, 1.0, 2.0) * d[i, j]
        G = V.T @ V
        j_ref = U @ (G @ (U.T @ w))
        vj = ri.vj([d], 1, 1.0)[0]
        err = maxabs(vj, j_ref)
        say(f"  J max-abs err {err:.3e} (max|J| {np.abs(j_ref).max():.3e})")
        assert err <= TOL
        Ua = [unpack(U[:, a], nbf) for a in range(3)]
        k_ref = np.zeros((nbf, nbf))
        for a in range(3):
            for bb in range(3):
                k_ref += G[a, bb] * (Ua[a] @ d @ Ua[bb])
        vk = ri.vk_occ([c], [nocc - 1], [np.ones(nocc)], 1, 1.0)[0]
        errk = maxabs(vk, k_ref[i, j])
        say(f"  K max-abs err {errk:.3e} (max|K| {np.abs(k_ref).max():.3e})")
        assert errk <= TOL

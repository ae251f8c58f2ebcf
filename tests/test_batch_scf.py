"""CPU tests of the batched SCF driver's per-molecule loop (rest_amd/batch_scf.py) and of the
batch wave planner (rjk_batch_plan_waves / rjk_batch_device_bytes, host arithmetic only).

The loop is driven with numpy J/K of the oracle's B and compared with the oracle's restatement of
REST's scf() (oracle/scf_oracle.py) from the same starting orbitals: same iteration count, energy
within 1e-10 Eh; with start_check_oscillation lowered so that the energy-rise mixing fires, both
take the same path.
"""
import numpy as np
import pytest
import scipy.linalg as sla

from oracle import scf_oracle as so
from rest_amd import batch_scf as bs
from rest_amd import jk, ri
from rest_amd import scf as rs
from rest_amd import workload as wl
from tests.test_integrals import numpy_jk


def drive(loop, jkf):
    co = next(loop)
    try:
        while True:
            co = loop.send(jkf(2.0 * co @ co.T))
    except StopIteration as ex:
        return ex.value


@pytest.fixture(scope="module")
def water2():
    m = ri.Molecule.water_cluster(wl.water_cluster(3, 2))
    S, T, V = m.int1e()
    B, _ = so.rimatr_from_integrals(m.int3c2e_packed(), m.int2c2e(), 1e-8)
    return m, S, T + V, B, numpy_jk(B, m.nbf)


@pytest.mark.parametrize("check_osc,swap", [(20, False), (1, True)])
def test_loop_matches_oracle_scf(water2, check_osc, swap):
    """From the core orbitals, and from core orbitals with the HOMO and LUMO exchanged (a start whose
    energy rises on the way, so that with start_check_oscillation = 1 the mixing runs 3 times)."""
    m, S, H, B, jkf = water2
    nocc = m.nelectron // 2
    _, C0 = sla.eigh(H, S)
    if swap:
        C0[:, [nocc - 1, nocc]] = C0[:, [nocc, nocc - 1]]
    ctrl = rs.default_ctrl()
    ctrl.start_check_oscillation = check_osc
    E, conv, it, D, C, eps, energies = drive(bs._rhf_loop(S, H, m.nuclear_repulsion(), nocc, C0, ctrl), jkf)
    oc = so.ScfCtrl(start_check_oscillation=check_osc)
    r = so.rhf(S, H, B, m.nelectron, m.nuclear_repulsion(), oc, d_guess=2 * C0[:, :nocc] @ C0[:, :nocc].T, jk=jkf)
    assert conv and r.converged and it == r.iterations
    assert sum(r.osc) == (3 if swap else 0)
    assert abs(E - r.energy) <= 1e-10
    assert np.allclose(energies, r.energies, atol=1e-9, rtol=0)
    assert abs(np.sum(D * S) - m.nelectron) <= 1e-9


def test_loop_stops_at_max_scf_cycle(water2):
    m, S, H, B, jkf = water2
    nocc = m.nelectron // 2
    _, C0 = sla.eigh(H, S)
    ctrl = rs.default_ctrl()
    ctrl.max_scf_cycle = 1          # REST stops at num_iter >= max(max_scf_cycle, 2), unconverged
    E, conv, it, *_ = drive(bs._rhf_loop(S, H, m.nuclear_repulsion(), nocc, C0, ctrl), jkf)
    assert not conv and it == 2


def test_loop_asks_for_the_single_problem_path_when_mixing(water2):
    m, S, H, B, jkf = water2
    nocc = m.nelectron // 2
    _, C0 = sla.eigh(H, S)
    ctrl = rs.default_ctrl()
    ctrl.start_diis_cycle = 3       # linear mixing of densities first: not 2 C Cᵀ of nocc columns
    with pytest.raises(bs._NeedsMixing):
        drive(bs._rhf_loop(S, H, m.nuclear_repulsion(), nocc, C0, ctrl), jkf)


def test_wave_planner_covers_the_batched_workload_in_order():
    sizes = wl.batched_cluster_sizes(512)
    nbf, naux, nocc = [24 * n for n in sizes], [113 * n for n in sizes], [5 * n for n in sizes]
    for prec in (jk.PREC_FP64, jk.PREC_FP32, jk.PREC_BF16X2, jk.PREC_BF16X3):
        per = [jk.batch_device_bytes([a], [b], [c], prec) for a, b, c in zip(nbf, naux, nocc)]
        assert all(x > 0 for x in per)
        assert jk.batch_device_bytes(nbf, naux, nocc, prec) == sum(per)
        budget = 64 << 30
        waves = jk.batch_plan_waves(nbf, naux, nocc, prec, budget)
        assert waves[0][0] == 0 and waves[-1][1] == 512
        for (a0, a1), (b0, _) in zip(waves, waves[1:]):
            assert a1 == b0 and a1 > a0
            assert sum(per[a0:a1]) <= budget                 # fits
            assert sum(per[a0:a1 + 1]) > budget              # and the next problem would not
    # one problem above the budget gets a wave of its own
    waves = jk.batch_plan_waves([768, 48, 768], [3616, 226, 3616], [160, 10, 160], jk.PREC_FP32, 1 << 30)
    assert waves == [(0, 1), (1, 2), (2, 3)]
    # BF16 storage: B once in FP32, no plane copy
    assert jk.batch_device_bytes([768], [3616], [160], jk.PREC_BF16X2) < \
        jk.batch_device_bytes([768], [3616], [160], jk.PREC_FP64)

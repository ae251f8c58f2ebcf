"""GPU tests of the batched J/K (include/rest_batch.h): grouped persistent
kernels over ragged problems, replayed from a hipGraph.

Bars: FP64 batch vs the oracle (small problems) and vs the single-problem
device path (larger ones) ≤ 1e-11 · max(1, |ref|); FP32 batch vs FP64 batch
≤ 2e-6 relative to the largest element; graph replays (with eager device work
and synchronisations between them) bit-identical to the first run; results
after the autotuner's configuration change ≤ 1e-12 relative.
"""
import sys

import numpy as np
import pytest
import torch

from oracle import oracle as orc
from rest_amd import jk

pytestmark = pytest.mark.gpu

SHAPES = [(24, 113, 5), (48, 226, 10), (7, 40, 3), (130, 300, 17), (72, 339, 15), (200, 500, 40), (384, 1808, 80)]


def say(msg):
    print(f"[gpu-case] {msg}", flush=True)
    sys.stdout.flush()


def inputs(i, nbf, nocc):
    c = orc.random_orthonormal(1000 + i, nbf, nocc)
    occ = np.full(nocc, 2.0)
    occ[-1] = 1.5          # non-uniform weights
    d = (c * occ) @ c.T
    return d, c, occ


def make_batch(prec, seed=11):
    nbf, naux, nocc = zip(*SHAPES)
    b = jk.Batch(nbf, naux, nocc, prec)
    b.fill_synthetic(seed)
    ins = []
    for i, (n, _, o) in enumerate(SHAPES):
        d, c, occ = inputs(i, n, o)
        b.set_inputs(i, d, c, occ)
        ins.append((d, c, occ))
    return b, ins


def test_batch_fp64_matches_oracle_and_single_problem_path():
    say(f"batched FP64 J/K over {len(SHAPES)} ragged problems")
    b, ins = make_batch(jk.PREC_FP64)
    with b:
        b.run(1.0, 1.0)
        b.synchronize()
        for i, (n, naux, nocc) in enumerate(SHAPES):
            vj, vk = b.get_outputs(i)
            d, c, occ = ins[i]
            if n <= 130:
                B = orc.synth_rimatr(11 + i, n, naux)
                rj = orc.vj_upper_rimatr(B, [d], 1, 1.0)[0]
                rk = orc.vk_upper_occ(B, True, [c], [nocc - 1], [occ], 1, 1.0)[0]
                src = "oracle"
            else:
                with jk.DeviceRI(n, naux) as r:
                    r.fill_synthetic(11 + i)
                    rj = r.vj([d], 1)[0]
                    rk = r.vk_occ([c], [nocc - 1], [occ], 1)[0]
                src = "single-problem device path"
            ej = np.abs(vj - rj).max() / max(1.0, np.abs(rj).max())
            ek = np.abs(vk - rk).max() / max(1.0, np.abs(rk).max())
            say(f"  problem {i} nbf {n} naux {naux} nocc {nocc} vs {src}: J {ej:.1e} K {ek:.1e}")
            assert ej <= 1e-11 and ek <= 1e-11


def test_batch_fp32_close_to_fp64():
    say("batched FP32 vs FP64")
    b64, _ = make_batch(jk.PREC_FP64)
    b32, _ = make_batch(jk.PREC_FP32)
    with b64, b32:
        b64.run(1.0, -0.5)
        b32.run(1.0, -0.5)
        b64.synchronize()
        b32.synchronize()
        for i in range(len(SHAPES)):
            j64, k64 = b64.get_outputs(i)
            j32, k32 = b32.get_outputs(i)
            ej = np.abs(j32 - j64).max() / np.abs(j64).max()
            ek = np.abs(k32 - k64).max() / np.abs(k64).max()
            say(f"  problem {i}: J {ej:.1e} K {ek:.1e}")
            assert ej <= 2e-6 and ek <= 2e-6


def test_graph_replays_and_autotuned_configuration():
    say("graph replays with eager work in between, then autotune")
    b, _ = make_batch(jk.PREC_FP64)
    with b:
        b.run(1.0, 1.0)
        b.synchronize()
        first = [b.get_outputs(i) for i in range(len(SHAPES))]
        x = torch.randn(512, 512, device="cuda", dtype=torch.float64)
        for rep in range(3):
            y = (x @ x).sum().item()          # eager device work and a synchronisation between replays
            b.run(1.0, 1.0)
            b.synchronize()
            for i in range(len(SHAPES)):
                vj, vk = b.get_outputs(i)
                assert np.array_equal(vj, first[i][0]) and np.array_equal(vk, first[i][1]), (rep, i)
        ms = b.autotune(reps=2)
        cut, waves = b.config()
        say(f"  autotuned: {ms:.3f} ms per batch, khalf cut {cut}, {waves} work-groups per CU")
        for cut_, w_ in ((0, 1), (1 << 30, 4), (cut, waves)):
            b.set_config(cut_, w_)
            b.run(1.0, 1.0)
            b.synchronize()
            for i in range(len(SHAPES)):
                vj, vk = b.get_outputs(i)
                for a, r in ((vj, first[i][0]), (vk, first[i][1])):
                    assert np.abs(a - r).max() <= 1e-12 * max(1.0, np.abs(r).max())


def test_batch_with_real_b():
    from rest_amd import ri
    from rest_amd import workload as wl
    say("batch of two real molecules ((H2O)2, (H2O)3) vs DeviceRI")
    mols = [ri.Molecule.water_cluster(wl.water_cluster(s, n)) for s, n in ((3, 2), (5, 3))]
    Bs = []
    for m in mols:
        with jk.DeviceRI(m.nbf, m.naux) as r:
            m.build_rimatr(r)
            Bs.append(r.get_rimatr())
    nocc = [m.nelectron // 2 for m in mols]
    with jk.Batch([m.nbf for m in mols], [m.naux for m in mols], nocc, jk.PREC_FP64) as b:
        ins = []
        for i, m in enumerate(mols):
            b.set_rimatr(i, Bs[i])
            d, c, occ = inputs(i, m.nbf, nocc[i])
            b.set_inputs(i, d, c, occ)
            ins.append((d, c, occ))
        b.run(1.0, -0.5)
        b.synchronize()
        for i, m in enumerate(mols):
            vj, vk = b.get_outputs(i)
            with jk.DeviceRI(m.nbf, m.naux) as r:
                r.set_rimatr(Bs[i])
                d, c, occ = ins[i]
                rj = r.vj([d], 1)[0]
                rk = r.vk_occ([c], [nocc[i] - 1], [occ], 1, -0.5)[0]
            assert np.abs(vj - rj).max() <= 1e-11 * max(1.0, np.abs(rj).max())
            assert np.abs(vk - rk).max() <= 1e-11 * max(1.0, np.abs(rk).max())


def test_batch_bf16_planes_match_split_oracle():
    """BF16 batch K: the plane split of the FP32 copy of B, checked against the split-plane
    emulation of the same FP32 B (oracle/split_oracle.py) on the small problems, under both tile
    classes: the device within a quarter of the emulation's distance from FP64 (the rest is FP32
    accumulation) at one and two planes, within 2e-6 at three; BF16X3 within 2e-6 of the FP64 batch
    on every problem; J equal to the FP32 batch's."""
    from oracle import split_oracle as so
    say("batched BF16X1/X2/X3 vs the split-plane emulation and the FP64 batch")
    b64, ins = make_batch(jk.PREC_FP64)
    b32, _ = make_batch(jk.PREC_FP32)
    with b64, b32:
        b64.run(1.0, -0.5)
        b32.run(1.0, -0.5)
        b64.synchronize()
        b32.synchronize()
        ref = [b64.get_outputs(i) for i in range(len(SHAPES))]
        ref32 = [b32.get_outputs(i) for i in range(len(SHAPES))]
    for prec, S in ((jk.PREC_BF16X1, 1), (jk.PREC_BF16X2, 2), (jk.PREC_BF16X3, 3)):
        bb, _ = make_batch(prec)
        with bb:
            for cut, waves in ((1 << 30, 2), (0, 1)):    # every problem narrow, then wide where it can be
                bb.set_config(cut, waves)
                bb.run(1.0, -0.5)
                bb.synchronize()
                for i, (n, naux, nocc) in enumerate(SHAPES):
                    vj, vk = bb.get_outputs(i)
                    assert np.array_equal(vj, ref32[i][0]), (S, i)
                    ek = np.abs(vk - ref[i][1]).max() / np.abs(ref[i][1]).max()
                    line = f"  S={S} cut={cut} problem {i} nbf {n}: K vs FP64 {ek:.1e}"
                    if n <= 130 and cut == 0:
                        d, c, occ = ins[i]
                        B32 = orc.synth_rimatr(11 + i, n, naux).astype(np.float32).astype(np.float64)
                        em = so.vk_split(B32, c, occ, S, -0.5)
                        e_or = np.abs(em - ref[i][1]).max() / np.abs(ref[i][1]).max()
                        e_dev = np.abs(vk - em).max() / np.abs(ref[i][1]).max()
                        line += f", emulation vs FP64 {e_or:.1e}, device vs emulation {e_dev:.1e}"
                        if S < 3:   # FP32 accumulation (up to 4096 products per accumulator) is the gap
                            assert e_dev < 0.25 * e_or, (S, i)
                        else:
                            assert e_dev < 2e-6, i
                    say(line)
                    if S == 3:
                        assert ek <= 2e-6, (cut, i)
                    elif S == 2:
                        assert ek <= 1e-4, (cut, i)


def test_batch_device_upload_active_flags_and_profile():
    """B from a DeviceRI's device storage equals the host upload; a switched-off problem keeps its
    outputs while the others change; the staged profile matches the graph's results."""
    from rest_amd import ri
    from rest_amd import workload as wl
    say("batch: B from DeviceRI on the device, active flags, staged profile")
    mols = [ri.Molecule.water_cluster(wl.water_cluster(s, n)) for s, n in ((3, 2), (5, 3), (7, 4))]
    nocc = [m.nelectron // 2 for m in mols]
    for prec in (jk.PREC_FP64, jk.PREC_BF16X3):
        with jk.Batch([m.nbf for m in mols], [m.naux for m in mols], nocc, prec) as bd, \
                jk.Batch([m.nbf for m in mols], [m.naux for m in mols], nocc, prec) as bh:
            ins = []
            for i, m in enumerate(mols):
                with jk.DeviceRI(m.nbf, m.naux) as r:
                    m.build_rimatr(r)
                    bd.set_from_ri(i, r)
                    bh.set_rimatr(i, r.get_rimatr())
                d, c, occ = inputs(i, m.nbf, nocc[i])
                bd.set_inputs(i, d, c, occ)
                bh.set_inputs(i, d, c, occ)
                ins.append((d, c, occ))
            for b in (bd, bh):
                b.run(1.0, -0.5)
                b.synchronize()
            first = [bd.get_outputs(i) for i in range(len(mols))]
            for i in range(len(mols)):
                jh, kh = bh.get_outputs(i)
                assert np.array_equal(first[i][0], jh) and np.array_equal(first[i][1], kh), (prec, i)
            # switch problem 1 off, change every input: problem 1 keeps its outputs
            bd.set_active(1, False)
            for i, m in enumerate(mols):
                d, c, occ = inputs(i + 50, m.nbf, nocc[i])
                bd.set_inputs(i, d, c, occ)
            bd.run(1.0, -0.5)
            bd.synchronize()
            now = [bd.get_outputs(i) for i in range(len(mols))]
            assert np.array_equal(now[1][0], first[1][0]) and np.array_equal(now[1][1], first[1][1])
            assert not np.array_equal(now[0][1], first[0][1]) and not np.array_equal(now[2][1], first[2][1])
            bd.set_active(1, True)
            ms = bd.profile(1.0, -0.5)
            staged = [bd.get_outputs(i) for i in range(len(mols))]
            bd.run(1.0, -0.5)
            bd.synchronize()
            for i in range(len(mols)):
                g = bd.get_outputs(i)
                assert np.array_equal(g[0], staged[i][0]) and np.array_equal(g[1], staged[i][1]), (prec, i)
            say(f"  prec {prec}: stage ms " + ", ".join(f"{n} {t:.3f}" for n, t in zip(jk.BATCH_STAGES, ms)))
            assert len(ms) == jk.BATCH_NUM_STAGES and all(t >= 0 for t in ms) and sum(ms) > 0


def test_512_cluster_waves_fp32_and_bf16x3():
    """The 512-cluster workload ((H2O)n, n uniform in [2, 32], def2-SVP/JKFIT shapes, synthetic B)
    in waves planned by the library under a 160 GiB budget: every FP32 wave runs, and a spread
    sample of problems (the smallest, the largest, and others) matches the single-problem FP64 path
    to 2e-6 (FP32 batches, and BF16X3 batches of the waves that hold the sample); the smallest also
    matches the oracle."""
    from rest_amd import workload as wl
    sizes = wl.batched_cluster_sizes(512)
    nbf, naux, nocc = [24 * n for n in sizes], [113 * n for n in sizes], [5 * n for n in sizes]
    budget = 160 << 30
    smallest, largest = int(np.argmin(sizes)), int(np.argmax(sizes))
    sample = sorted({smallest, largest, 100, 257, 511})
    done_ref = {}
    for prec in (jk.PREC_FP32, jk.PREC_BF16X3):
        waves = jk.batch_plan_waves(nbf, naux, nocc, prec, budget)
        assert waves[0][0] == 0 and waves[-1][1] == 512
        assert all(a[1] == b[0] for a, b in zip(waves, waves[1:]))
        say(f"512 clusters, prec {prec}: {len(waves)} waves under {budget >> 30} GiB")
        got = {}
        for w, (i0, i1) in enumerate(waves):
            assert jk.batch_device_bytes(nbf[i0:i1], naux[i0:i1], nocc[i0:i1], prec) <= budget or i1 - i0 == 1
            if prec != jk.PREC_FP32 and not any(i0 <= i < i1 for i in sample):
                continue
            with jk.Batch(nbf[i0:i1], naux[i0:i1], nocc[i0:i1], prec) as b:
                b.fill_synthetic(7000 + i0)            # problem i gets seed 7000 + i
                for i in range(i0, i1):
                    d, c, occ = inputs(i, nbf[i], nocc[i])
                    b.set_inputs(i - i0, d, c, occ)
                b.run(1.0, -0.5)
                b.synchronize()
                for i in range(i0, i1):
                    if i in sample:
                        got[i] = b.get_outputs(i - i0)
                vj, vk = b.get_outputs(0)
                assert np.isfinite(vj).all() and np.isfinite(vk).all()
            say(f"  wave {w + 1}/{len(waves)}: problems {i0}..{i1 - 1}")
        for i in sample:
            if i in done_ref:
                rj, rk = done_ref[i]
            else:
                d, c, occ = inputs(i, nbf[i], nocc[i])
                with jk.DeviceRI(nbf[i], naux[i]) as r:
                    r.fill_synthetic(7000 + i)
                    rj = r.vj([d], 1)[0]
                    rk = r.vk_occ([c], [nocc[i] - 1], [occ], 1, -0.5)[0]
                done_ref[i] = (rj, rk)
            ej = np.abs(got[i][0] - rj).max() / np.abs(rj).max()
            ek = np.abs(got[i][1] - rk).max() / np.abs(rk).max()
            say(f"  problem {i} (n={sizes[i]}): J {ej:.1e} K {ek:.1e} vs the single-problem FP64 path")
            assert ej <= 2e-6 and ek <= 2e-6
            if i == smallest and prec == jk.PREC_FP32:
                d, c, occ = inputs(i, nbf[i], nocc[i])
                B = orc.synth_rimatr(7000 + i, nbf[i], naux[i])
                oj = orc.vj_upper_rimatr(B, [d], 1, 1.0)[0]
                ok = orc.vk_upper_occ(B, True, [c], [nocc[i] - 1], [occ], 1, -0.5)[0]
                assert np.abs(rj - oj).max() <= 1e-11 * max(1.0, np.abs(oj).max())
                assert np.abs(rk - ok).max() <= 1e-11 * max(1.0, np.abs(ok).max())


def test_batched_scf_matches_single_problem_scf():
    """REST's loop per molecule with the J/K of every iteration from one batch (FP64): energies of
    five water clusters within 1e-9 Eh of the single-problem device SCF from the same SAD guess."""
    from rest_amd import batch_scf as bs
    from rest_amd import ri
    from rest_amd import scf as rs
    from rest_amd import workload as wl
    mols = [ri.Molecule.water_cluster(wl.water_cluster(s, n)) for s, n in ((3, 2), (5, 3), (7, 4), (9, 6), (1, 1))]
    say("batched SCF of five water clusters vs the single-problem SCF")
    st = {}
    outs = bs.rhf_batch(mols, stats=st, budget_bytes=8 << 30)
    say(f"  {st}")
    refs = []
    for m, o in zip(mols, outs):
        ref, _ = rs.run_molecule(m, guess="sad")
        refs.append(ref)
        say(f"  nbf {m.nbf}: batched {o.energy:.10f} ({o.iterations} it, batched={o.batched}), "
            f"single {ref.energy:.10f} ({ref.iterations} it)")
        assert o.converged and ref.converged and o.batched
        assert abs(o.energy - ref.energy) <= 1e-9
        assert abs(np.trace(o.dm @ m.int1e()[0]) - m.nelectron) <= 1e-8
    # linear density mixing before DIIS needs more K columns than the batch holds: those molecules
    # leave the batch and finish on the single-problem SCF with the same controls
    ctrl = rs.default_ctrl()
    ctrl.start_diis_cycle = 3
    st = {}
    outs = bs.rhf_batch(mols[:2], ctrl=ctrl, stats=st, budget_bytes=8 << 30)
    say(f"  start_diis_cycle 3: {st['fallbacks']} fallbacks")
    assert st["fallbacks"] == 2
    for m, o, ref in zip(mols[:2], outs, refs[:2]):
        assert not o.batched and o.converged and abs(o.energy - ref.energy) <= 1e-7

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

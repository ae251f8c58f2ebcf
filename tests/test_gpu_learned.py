"""GPU tests of the learned guess's device network (include/rest_lg.h) against
the MLP oracle (oracle/mlp_oracle.py, pinned on the CPU by finite
differences), and of the guess pipeline's acceptance rule.

Bars: BF16-operand GEMM vs float64 products of the same BF16-rounded operands
≤ 1e-5 relative; loss ≤ 1e-5 relative; gradient ≤ 3e-3 in relative norm (FP32
accumulation and SiLU, occasional BF16 rounding flips of intermediates); Adam
steps ≤ 1e-6 relative to the oracle's update of the same gradient."""
import sys

import numpy as np
import pytest
import torch

from oracle import mlp_oracle as mo
from rest_amd import learned as lg

pytestmark = pytest.mark.gpu


def say(msg):
    print(f"[gpu-case] {msg}", flush=True)
    sys.stdout.flush()


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float32)).cuda()


@pytest.mark.parametrize("M,N,K,ta,tb", [(64, 64, 32, 0, 0), (70, 45, 50, 0, 0), (33, 130, 257, 1, 0),
                                         (300, 96, 77, 0, 1), (5, 7, 3, 1, 1)])
def test_bf16_gemm_matches_rounded_float64(M, N, K, ta, tb):
    say(f"rlg_gemm_bf16 M={M} N={N} K={K} transA={ta} transB={tb}")
    net = lg.Net(8, 8, 16, 1, 64)
    rng = np.random.default_rng(M * N + K)
    A = rng.standard_normal((K, M) if ta else (M, K)).astype(np.float32)
    B = rng.standard_normal((N, K) if tb else (K, N)).astype(np.float32)
    ref = (mo.bf16_round(A).T if ta else mo.bf16_round(A)) @ (mo.bf16_round(B).T if tb else mo.bf16_round(B))
    Ad, Bd = dev(A), dev(B)
    Cd = torch.full((M, N), np.nan, dtype=torch.float32, device="cuda")
    sam, sak = (1, M) if ta else (K, 1)
    sbk, sbn = (1, K) if tb else (N, 1)
    rc = net.lib.rlg_gemm_bf16(net.h, Ad.data_ptr(), sam, sak, Bd.data_ptr(), sbk, sbn, Cd.data_ptr(), N, M, N, K)
    assert rc == 0
    C = Cd.cpu().numpy()
    err = np.abs(C - ref).max() / np.abs(ref).max()
    say(f"  max rel err {err:.2e}")
    assert err <= 1e-5


def small_problem(seed=0, batch=600):
    rng = np.random.default_rng(seed)
    net = lg.Net(lg.N_IN, lg.N_OUT, 64, 3, 1024, seed=seed)
    p = net.get_params().astype(np.float64)
    p += rng.standard_normal(p.size) * 0.02       # non-zero last layer and biases
    net.set_params(p.astype(np.float32))
    X = rng.standard_normal((batch, lg.N_IN)).astype(np.float32)
    T = rng.standard_normal((batch, lg.N_OUT)).astype(np.float32) * 0.3
    M = (rng.random((batch, lg.N_OUT)) > 0.5).astype(np.float32)
    return net, X, T, M


def test_loss_and_gradient_match_oracle():
    say("rlg_loss_grad vs oracle (dims N_IN-64-64-64-196, batch 600)")
    net, X, T, M = small_problem()
    p = net.get_params()
    loss, g = net.loss_grad_dev(dev(X), dev(T), dev(M), want_grad=True)
    lo, go = mo.loss_grad(p, net.dims, X, T, M, bf16=True)
    gerr = np.linalg.norm(g - go) / np.linalg.norm(go)
    say(f"  loss {loss:.8e} oracle {lo:.8e}; gradient rel err {gerr:.2e}")
    assert abs(loss - lo) <= 1e-5 * abs(lo)
    assert gerr <= 3e-3
    # per layer, so that no layer hides behind a large one
    off = 0
    for din, dout in zip(net.dims[:-1], net.dims[1:]):
        for cnt in (din * dout, dout):
            a, b = g[off:off + cnt], go[off:off + cnt]
            assert np.linalg.norm(a - b) <= 5e-3 * np.linalg.norm(b) + 1e-12
            off += cnt


def test_adam_matches_oracle_over_three_steps():
    say("rlg_adam_step x3 vs oracle")
    net, X, T, M = small_problem(1)
    p = net.get_params().astype(np.float64)
    m = np.zeros_like(p)
    v = np.zeros_like(p)
    for step in range(1, 4):
        _, g = net.loss_grad_dev(dev(X), dev(T), dev(M), want_grad=True)
        net.adam(1e-3)
        p, m, v = mo.adam(p, g.astype(np.float64), m, v, step, 1e-3)
        pd = net.get_params()
        err = np.abs(pd - p).max() / np.abs(p).max()
        say(f"  step {step}: max rel diff {err:.2e}")
        assert err <= 1e-6
        p = pd.astype(np.float64)


def test_training_fits_a_smooth_target():
    """A learnable map (a fixed random two-layer function of the inputs) is fitted ≥ 20× below its variance."""
    say("training on a synthetic smooth target")
    rng = np.random.default_rng(5)
    X = rng.standard_normal((16384, lg.N_IN)).astype(np.float32)
    W1 = rng.standard_normal((lg.N_IN, 32)) / np.sqrt(lg.N_IN)
    W2 = rng.standard_normal((32, lg.N_OUT)) / np.sqrt(32)
    T = (np.tanh(X @ W1) @ W2).astype(np.float32)
    M = np.ones_like(T)
    net = lg.Net(lg.N_IN, lg.N_OUT, 256, 3, 4096, seed=1)
    hist = lg.train(net, X, T, M, epochs=40, batch=512, lr=2e-3, log=say)
    var = float(np.mean((T - T.mean(axis=0)) ** 2))
    say(f"  target variance {var:.3e}, final val {hist['val'][-1]:.3e}")
    assert hist["val"][-1] < var / 20
    Y = net.predict(X[:512])
    assert np.mean((Y - T[:512]) ** 2) < var / 20


def test_learned_guess_falls_back_to_sad_for_a_useless_network():
    """An untrained network's density is rejected (its guess energy is above SAD's) or, if kept, the
    SCF still converges to the SAD-started energy."""
    from rest_amd import jk, ri
    from rest_amd import scf as rs
    from rest_amd import workload as wl
    say("(H2O)2 learned guess with a random network")
    m = ri.Molecule.water_cluster(wl.water_cluster(3, 2))
    S, T_, V = m.int1e()
    H = T_ + V
    net = lg.Net(seed=3)
    with jk.DeviceRI(m.nbf, m.naux) as r:
        m.build_rimatr(r)
        g = lg.learned_guess(net, m, r, S, H)
        ctrl = rs.default_ctrl(scf_acc_etot=1e-11, scf_acc_rho=1e-9, scf_acc_eev=1e-8, max_scf_cycle=200)
        out = rs.rhf(r, S, H, m.nuclear_repulsion(), m.nelectron, None, ctrl, c_guess=g.c1)
        ref, _ = rs.run_molecule(m, guess="sad", ctrl=rs.default_ctrl(scf_acc_etot=1e-11, scf_acc_rho=1e-9,
                                                                        scf_acc_eev=1e-8, max_scf_cycle=200))
    say(f"  kind {g.kind}: E_learned {g.e_learned:.6f} E_sad {g.e_sad:.6f} electrons {g.electrons:.4f}; "
        f"SCF {out.energy:.10f} vs SAD-started {ref.energy:.10f}")
    assert (g.kind == "sad") == (not (g.e_learned < g.e_sad and abs(g.electrons - m.nelectron) < 0.5))
    assert out.converged and abs(out.energy - ref.energy) < 1e-9


@pytest.mark.parametrize("use_model,guard", [(True, "energy"), (False, "energy"), (True, "error")])
def test_learned_scf_matches_fp64_and_rejects(use_model, guard):
    """learned_scf: the SCF from the purified prediction, kept by the guard ("energy": its guess Fock
    energy below SAD's E(D1); "error": its guess build's max|F D S - S D F| below GUESS_ERR_MAX),
    else from SAD; on the mixed ladder (BF16X2, BF16X3, FP64 reference + increments).  The packaged
    model (kept) and an untrained network (rejected) both end at the FP64 SAD-started energy, and a
    forced rejection (a bar no guess meets) does too."""
    from pathlib import Path
    from rest_amd import guess as gs
    from rest_amd import jk, ri
    from rest_amd import scf as rs
    from rest_amd import workload as wl
    say(f"(H2O)8 learned_scf, {'packaged model' if use_model else 'untrained network'}, guard {guard}")
    m = ri.Molecule.water_cluster(wl.water_cluster(5, 8))
    S, T_, V = m.int1e()
    H = T_ + V
    model = Path(lg.__file__).resolve().parent / "models" / "learned_guess_h2o.npz"
    net = lg.Net.load(model) if use_model else lg.Net(seed=3)
    tight = dict(scf_acc_etot=1e-11, scf_acc_rho=1e-9, scf_acc_eev=1e-8, max_scf_cycle=200)
    with jk.DeviceRI(m.nbf, m.naux) as r:
        m.build_rimatr(r, 1e-8, jk.STORE_FP64 | jk.STORE_FP32 | jk.STORE_BF16)
        ref = rs.rhf(r, S, H, m.nuclear_repulsion(), m.nelectron, gs.sad_guess(m, r, S, H).d1,
                     rs.default_ctrl(**tight))
        ctrl = rs.mixed_ctrl(**tight)
        g = lg.learned_scf(net, m, r, S, H, ctrl, guard=guard)
        forced = lg.learned_scf(net, m, r, S, H, ctrl, guard=guard, guess_err_max=1e-12) if guard == "error" else None
    precs = [int(x["prec"]) for x in g.outcome.log]
    say(f"  kind {g.kind}: E_learned {g.e_learned:.6f} E_sad {g.e_sad} err {g.err_learned}; SCF "
        f"{g.outcome.energy:.10f} ({g.outcome.iterations} it, precisions {precs}, incremental "
        f"{g.outcome.incremental_iterations}) vs FP64 SAD-started {ref.energy:.10f}")
    if guard == "energy":
        assert (g.kind == "learned") == (g.e_learned < g.e_sad - g.band)
    else:
        assert g.e_sad is None and g.kind == "learned" and g.err_learned < lg.GUESS_ERR_MAX
        assert forced.kind == "sad" and forced.outcome.converged
        assert abs(forced.outcome.energy - ref.energy) <= 1e-9
    assert g.outcome.converged and abs(g.outcome.energy - ref.energy) <= 1e-9
    assert precs[-1] == jk.PREC_FP64 and g.outcome.incremental_iterations >= 1
    if use_model:
        assert g.kind == "learned"


def test_energy_guard_refines_an_inconclusive_estimate():
    """The energy guard on an estimate of SAD's E(D1) at one BF16 plane: with its error bound widened
    to 1% of the energy (so that the learned density falls inside the band), the guard recomputes
    E(D1) at FP64, keeps the prediction, and the SCF ends at the FP64 SAD-started energy; with the
    measured bound the prediction is kept as well (by the first estimate, or by FP64 if that was
    inconclusive)."""
    from pathlib import Path
    from rest_amd import guess as gs
    from rest_amd import jk, ri
    from rest_amd import scf as rs
    from rest_amd import workload as wl
    say("(H2O)8 energy guard at BF16X1, widened band")
    m = ri.Molecule.water_cluster(wl.water_cluster(5, 8))
    S, T_, V = m.int1e()
    H = T_ + V
    net = lg.Net.load(Path(lg.__file__).resolve().parent / "models" / "learned_guess_h2o.npz")
    with jk.DeviceRI(m.nbf, m.naux) as r:
        m.build_rimatr(r, 1e-8, jk.STORE_FP64 | jk.STORE_FP32 | jk.STORE_BF16)
        ref = rs.rhf(r, S, H, m.nuclear_repulsion(), m.nelectron, gs.sad_guess(m, r, S, H).d1, rs.default_ctrl())
        ctrl = rs.mixed_ctrl()
        wide = lg.learned_scf(net, m, r, S, H, ctrl, guard="energy", guard_prec=jk.PREC_BF16X1,
                              guard_band_rel={jk.PREC_BF16X1: 1e-2})
        plain = lg.learned_scf(net, m, r, S, H, ctrl, guard="energy", guard_prec=jk.PREC_BF16X1)
        e1_64 = rs.guess_step(r, S, H, m.nuclear_repulsion(), (m.nelectron,), sum(gs.sad_density(m)),
                              jk.PREC_FP64, want_e1=True, want_outputs=False, factors=gs.sad_factors(m))[4]
    say(f"  widened: kind {wide.kind}, refined {wide.refined}, E_learned {wide.e_learned:.6f}, E_sad {wide.e_sad:.6f} "
        f"(FP64 {e1_64:.6f}); measured band: kind {plain.kind}, refined {plain.refined}, band {plain.band:.3f}")
    assert wide.refined and wide.kind == "learned" and wide.e_sad == pytest.approx(e1_64, abs=1e-8)
    assert plain.kind == "learned" and plain.e_learned < plain.e_sad - plain.band
    if plain.refined:
        assert plain.e_sad == pytest.approx(e1_64, abs=1e-8)
    for g in (wide, plain):
        assert g.outcome.converged and abs(g.outcome.energy - ref.energy) <= 1e-8

"""Benchmarks of the MI355X RI-V J/K path.  Prints ONE JSON line.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--mode scf|jk] ...

--mode scf (default): the headline metric, SCF time to 1e-8 Eh for
  (H2O)64 def2-SVP / def2-SVP-JKFIT (water cluster seed 2): plain FP64 from
  REST's SAD guess vs mixed precision from the learned guess, on the same GPU
  and the same resident B.  Mixed (scf.mixed_ctrl): K on two BF16 planes of
  B (J on its FP32 copy) until the DIIS error max|FDS-SDF| is below 1e-4,
  then three planes until 1e-5, then one FP64 reference build and
  FP64-referenced BF16X2 increments, pseudo-diagonalising in between; the
  learned density is kept when its guess build's max|FDS-SDF| is below
  learned.GUESS_ERR_MAX (--guard error, default) or when its energy is below
  that of SAD's first orbital density (--guard energy).  One step = one
  plain-FP64 SCF plus one mixed+learned SCF, each including its guess;
  `value` is their time ratio (plain / mixed+learned).  The other guard, the
  FP64 tail without increments, both together and the mixed ladder from SAD
  are timed once each outside the timed region and reported under detail.
  Integrals, the B build and the atomic SAD densities are outside the timed
  region and reported separately.  --mixed-prec fp32 runs the mixed SCF on
  the FP32 copy of B instead.
--mode jk: one J + one K build (one SCF iteration's worth of the hot path)
  on synthetic B (seeded hash generator) at the (H2O)64 or (H2O)16 shape,
  D = C Cᵀ from the seeded orthonormal Gaussian C, occupied-route K.

`roofline` is for the kernel that dominates the timed region, from the
library's live HIP-event timing on the launch stream and the algorithmic work
of each launch; `cpu_baseline` times the oracle's reference-structured CPU
J/K on a bounded sample of the same workload.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0          # MI355X HBM3E
MFMA_PEAK_TFS = {"fp64": 78.6, "fp32": 157.3, "bf16": 2500.0}   # MI355X dense matrix peaks per instruction type
# MFMA products per algorithmic product on the split-BF16 planes (a + b < S)
BF16_PRODUCTS = {"bf16x1": 1, "bf16x2": 3, "bf16x3": 6}

PHASE_KERNEL = {"j_gamma": "k_jgamma", "j_acc": "k_jacc", "k_half": "k_khalf", "k_syrk": "k_ksyrk"}


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--mode", default="scf", choices=["scf", "jk", "batched", "cpu"])
    ap.add_argument("--clusters", type=int, default=512, help="batched mode: number of water clusters")
    ap.add_argument("--batch-prec", default="bf16x3", choices=["fp64", "fp32", "bf16x1", "bf16x2", "bf16x3"],
                    help="batched mode: J/K precision of the batch")
    ap.add_argument("--batch-budget-gib", type=float, default=150.0, help="batched mode: device memory per wave")
    ap.add_argument("--workload", default="h2o64", choices=["h2o64", "h2o16"])
    ap.add_argument("--model", default=str(ROOT / "rest_amd" / "models" / "learned_guess_h2o.npz"))
    ap.add_argument("--prec", default="fp64", choices=["fp64", "fp32", "bf16x3", "bf16x2", "bf16x1"],
                    help="jk mode: fp64 plain path; fp32: FP32 copy of B, FP32 MFMA K, FP64 accumulation of J; "
                         "bf16x3 / bf16x2 / bf16x1: K on three / two / one BF16 planes of B (6 / 3 / 1 MFMA products), "
                         "J on the FP32 copy")
    ap.add_argument("--mixed-prec", default="bf16x3", choices=["fp32", "bf16x2", "bf16x3"],
                    help="scf mode: J/K precision of the mixed SCF before its switch to FP64 (bf16x3: two planes "
                         "until max|e| < 1e-4, then three; bf16x2: two planes throughout)")
    ap.add_argument("--pseudo-diag-err", type=float, default=1e-2,
                    help="scf mode: low-precision iterations with max|e| below this pseudo-diagonalise (0: never)")
    ap.add_argument("--increment-prec", default="bf16x2", choices=["none", "fp32", "bf16x2", "bf16x3"],
                    help="scf mode: precision of the J/K increments on the FP64 reference build after the switch")
    ap.add_argument("--guard", default="error", choices=["error", "energy"],
                    help="scf mode: learned-guess acceptance (guess-build commutator, or energy against SAD's E(D1))")
    ap.add_argument("--guard-prec", default="bf16x1", choices=["fp64", "fp32", "bf16x1", "bf16x2", "bf16x3"],
                    help="scf mode: J/K precision of SAD's E(D1) for the energy rule")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--calibrate-copy-mib", type=int, default=0,
                    help="run one plain device copy of this many MiB first (counter calibration)")
    return ap.parse_args()


def cpu_threads(requested: int) -> tuple:
    """(threads used, nproc): every host core up to 16, one GPU's share of a multi-GPU host."""
    nproc = os.cpu_count() or 1
    return (requested or min(nproc, 16)), nproc


def cpu_medians(shape, seed, threads, reps=5):
    """Reference-structured CPU J, occupied-route K and density-only K over every auxiliary column
    of a synthetic B of `shape`: one warm-up, then the median of `reps` builds each (seconds)."""
    from oracle import oracle as orc
    from rest_amd import workload as wl
    orc.load_host_blas()
    b = orc.synth_rimatr(seed, shape.nbf, shape.naux)
    c = wl.random_orthonormal(seed, shape.nbf, shape.nocc)
    occ = np.full(shape.nocc, 2.0)
    d = (c * occ) @ c.T
    out = {}
    for name, fn in (("j", lambda: orc.cpu_baseline_vj(b, shape.nbf, d, threads)[0]),
                     ("k_occ", lambda: orc.cpu_baseline_vk(b, shape.nbf, c, occ, threads)[0]),
                     ("k_dm", lambda: orc.cpu_baseline_vk_dm(b, shape.nbf, d, threads)[0])):
        fn()
        out[name + "_s"] = float(np.median([fn() for _ in range(reps)]))
    return out


def cpu_baseline(shape, seed, c, d, threads):
    """Reference-structured CPU J/K (oracle): the headline shape sampled over a bounded number of
    auxiliary columns and scaled to naux (an estimate), plus full medians at the one-water and
    (H2O)16 shapes."""
    from oracle import oracle as orc
    from rest_amd import workload as wl
    blas = orc.load_host_blas()
    nthreads, nproc = cpu_threads(threads)
    ncol_j, ncol_k, ncol_dm, reps = 256, 128, 32, 3
    bj = orc.synth_rimatr(seed, shape.nbf, shape.naux, 0, ncol_j)
    tj = [orc.cpu_baseline_vj(bj, shape.nbf, d, nthreads)[0] for _ in range(reps)]
    del bj
    bk = orc.synth_rimatr(seed, shape.nbf, shape.naux, 0, ncol_k)
    occ = np.ones(c.shape[1])
    tk = [orc.cpu_baseline_vk(bk, shape.nbf, c, occ, nthreads)[0] for _ in range(reps)]
    bd = bk[:, :ncol_dm]
    tdm = [orc.cpu_baseline_vk_dm(bd, shape.nbf, d, nthreads)[0] for _ in range(reps)]
    del bk, bd
    t_j = float(np.median(tj)) * shape.naux / ncol_j
    t_k = float(np.median(tk)) * shape.naux / ncol_k
    t_dm = float(np.median(tdm)) * shape.naux / ncol_dm
    full = {s_.name: cpu_medians(s_, 1, nthreads) for s_ in (wl.H2O, wl.H2O_16)}
    # REST's RI-HF builds K through the density-only route when use_ri_symm is on (its default for
    # RI-V, src/scf_io/mod.rs:2027-2029): the baseline build is J + density-only K
    return {
        "value": 1.0 / (t_j + t_dm),
        "unit": "J/K builds/s",
        "route": "J + density-only K (vk_upper_with_rimatr_use_dm_only_sync); occupied-route K in k_seconds_scaled",
        "cores": nthreads,
        "nproc": nproc,
        "kind": "port",
        "sample": (f"oracle reference-structured J on {ncol_j}, occupied-route K on {ncol_k} and density-only K on "
                   f"{ncol_dm} of {shape.naux} aux columns (median of {reps}), scaled linearly to all columns "
                   f"(an estimate); per-P BLAS: {blas or 'built-in loops'} single-threaded inside {nthreads} OpenMP "
                   f"tasks ({nproc} cores reported by the OS)"),
        "j_seconds_scaled": t_j,
        "k_seconds_scaled": t_k,
        "k_dm_seconds_scaled": t_dm,
        "full_medians_s": full,
    }


def main_cpu(args):
    """CPU baseline protocol: reference-structured J, occupied-route K and density-only K (median of
    5 after one warm-up, every auxiliary column) at the one-water and (H2O)16 shapes, the (H2O)64
    shape as a scaled sample, and the whole RI-HF SCF on the CPU (REST's loop restated in
    oracle/scf_oracle.py, J and density-only K from the same reference-structured routines, FP64) for
    the single water and (H2O)16 seed 1 from the SAD density."""
    from oracle import oracle as orc
    from oracle import scf_oracle as so
    from rest_amd import guess as gs
    from rest_amd import ri
    from rest_amd import workload as wl
    blas = orc.load_host_blas()
    nthreads, nproc = cpu_threads(args.cpu_threads)
    out = {"mode": "cpu", "threads": nthreads, "nproc": nproc, "blas": blas or "built-in loops",
           "builds": {}, "scf": {}}
    for shape in (wl.H2O, wl.H2O_16):
        out["builds"][shape.name] = cpu_medians(shape, 1, nthreads)
        print(f"[bench] cpu builds {shape.name}: {out['builds'][shape.name]}", file=sys.stderr, flush=True)
    c3 = wl.H2O_64
    c = wl.random_orthonormal(2, c3.nbf, c3.nocc)
    cb = cpu_baseline(c3, 2, c, c @ c.T, nthreads)
    out["builds"][c3.name + " (scaled estimate)"] = {"j_s": cb["j_seconds_scaled"], "k_occ_s": cb["k_seconds_scaled"],
                                                     "k_dm_s": cb["k_dm_seconds_scaled"], "sample": cb["sample"]}
    mono = ri.Molecule.from_angstrom(["O", "H", "H"], np.array([[0.0, 0.0, 0.0], [0.7572, 0.5865, 0.0],
                                                                [-0.7572, 0.5865, 0.0]]))
    for name, mol in (("H2O", mono), ("(H2O)16 seed 1", ri.Molecule.water_cluster(wl.water_cluster(1, 16)))):
        t0 = time.perf_counter()
        S, T, V = mol.int1e()
        B, kept = so.rimatr_from_integrals(mol.int3c2e_packed(), mol.int2c2e(), 1e-8)
        t_setup = time.perf_counter() - t0
        da, db = gs.sad_density(mol)
        jk_s = [0.0]

        def jk(D, B=B, n=mol.nbf):
            a = time.perf_counter()
            _, vj = orc.cpu_baseline_vj(B, n, D, nthreads)
            _, vk = orc.cpu_baseline_vk_dm(B, n, D, nthreads)
            jk_s[0] += time.perf_counter() - a
            iu = np.triu_indices(n)
            order = np.lexsort((iu[0], iu[1]))
            J, K = np.zeros((n, n)), np.zeros((n, n))
            J[iu[0][order], iu[1][order]] = vj
            K[iu[0][order], iu[1][order]] = vk
            return J + np.triu(J, 1).T, K + np.triu(K, 1).T

        t1 = time.perf_counter()
        r = so.rhf(S, T + V, B, mol.nelectron, mol.nuclear_repulsion(), so.ScfCtrl(), d_guess=da + db, jk=jk)
        t_scf = time.perf_counter() - t1
        out["scf"][name] = {"nbf": mol.nbf, "naux": mol.naux, "energy": r.energy, "converged": bool(r.converged),
                            "iterations": r.iterations, "scf_s": t_scf, "jk_s": jk_s[0],
                            "setup_s_integrals_and_b": t_setup}
        print(f"[bench] cpu scf {name}: {out['scf'][name]}", file=sys.stderr, flush=True)
    print(json.dumps(out))


def load_traffic(kernel: str):
    p = ROOT / "profiles" / f"traffic_{kernel}.json"  # written from rocprofv3 --pmc runs
    if p.exists():
        try:
            return json.loads(p.read_text()).get("traffic_bytes_per_launch")
        except Exception:
            return None
    return None


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU only (asked for --gpus {args.gpus}); "
              "the J/K path would shard over the auxiliary index P with one all-reduce of the packed "
              "J/K partials, which is not part of this build.", file=sys.stderr)
        sys.exit(2)
    if args.mode == "scf":
        return main_scf(args)
    if args.mode == "batched":
        return main_batched(args)
    if args.mode == "cpu":
        return main_cpu(args)
    return main_jk(args)


def roofline_from_detail(detail: dict):
    """Dominant kernel of the timed region and its achieved rate from the library's per-launch work.
    BF16-plane launches execute 3 (BF16X2) or 6 (BF16X3) MFMA products per algorithmic product: their
    rate is the executed MFMA rate against the BF16 peak."""
    best = None
    for (prec, phase), (ms, n, work) in detail.items():
        if phase in PHASE_KERNEL and n and (best is None or ms > best[2]):
            best = (prec, phase, ms, n, work)
    prec, phase, ms, n, work = best
    sec = ms / 1e3
    if phase.startswith("j_"):
        ach, peak, unit, bound = work / sec / 1e9, HBM_PEAK_GBS, "GB/s", "hbm"
    else:
        mult = BF16_PRODUCTS.get(prec, 1)
        ach, peak, unit, bound = mult * work / sec / 1e12, MFMA_PEAK_TFS["bf16" if mult > 1 else prec], "TFLOP/s", "mfma"
    if prec in BF16_PRODUCTS and not phase.startswith("j_"):
        kernel = {"k_half": "k_bf_khalf", "k_syrk": "k_bf_ksyrk"}[phase] + f"<{prec[-1]}>"
    else:
        kernel = PHASE_KERNEL[phase] + ("<double>" if prec == "fp64" else "<float>")
    return {"bound": bound, "achieved": ach, "peak": peak, "unit": unit, "frac": ach / peak,
            "traffic": load_traffic(f"{PHASE_KERNEL[phase]}_{prec}_scf"), "kernel": kernel,
            "launches": n, "avg_launch_ms": ms / n, "algorithmic_per_launch": work / n}


def main_scf(args):
    import torch
    from rest_amd import guess as gs
    from rest_amd import jk, ri
    from rest_amd import learned as lg
    from rest_amd import scf as rs
    from rest_amd import workload as wl

    n, seed = (64, 2) if args.workload == "h2o64" else (16, 1)
    model = Path(args.model)
    if not model.exists():
        raise SystemExit(f"bench.py: learned-guess model {model} is missing (train it with "
                         "`python -m rest_amd.train_guess --out {model}`)")
    t0 = time.perf_counter()
    mol = ri.Molecule.water_cluster(wl.water_cluster(seed, n))
    S, T, V = mol.int1e()
    H = T + V
    enuc = mol.nuclear_repulsion()
    t1 = time.perf_counter()
    r = jk.DeviceRI(mol.nbf, mol.naux)
    stream = torch.cuda.Stream(device=torch.device("cuda:0"))
    torch.cuda.set_stream(stream)
    r.set_stream(stream.cuda_stream)
    mixed_prec = jk.PREC_NAMES[args.mixed_prec]
    store = jk.STORE_FP64 | jk.STORE_FP32 | (jk.STORE_BF16 if args.mixed_prec.startswith("bf16") else 0)
    binfo = mol.build_rimatr(r, 1e-8, store)
    t2 = time.perf_counter()
    net = lg.Net.load(model)
    net.set_stream(stream.cuda_stream)
    gs.sad_density(mol)            # atomic SCFs (cached per element)
    t3 = time.perf_counter()

    def plain():
        da, db = gs.sad_density(mol)
        e0, _, c1, _ = rs.guess_step(r, S, H, enuc, (mol.nelectron,), da + db, jk.PREC_FP64,
                                     factors=gs.sad_factors(mol))
        out = rs.rhf(r, S, H, enuc, mol.nelectron, None, rs.default_ctrl(jk_prec=jk.PREC_FP64), c_guess=c1)
        return out, {"guess": "sad", "e_guess": e0}

    guard_prec = jk.PREC_NAMES[args.guard_prec]

    def mixed(guard=args.guard, increment_prec=args.increment_prec):
        ctrl = rs.mixed_ctrl(jk_prec=mixed_prec, pseudo_diag_err=args.pseudo_diag_err,
                             jk_prec_increment=-1 if increment_prec == "none" else jk.PREC_NAMES[increment_prec])
        if not args.mixed_prec.startswith("bf16"):
            ctrl.jk_prec_start = -1                    # FP32 throughout the low-precision phase
        g = lg.learned_scf(net, mol, r, S, H, ctrl, guard=guard, guard_prec=guard_prec)
        return g.outcome, {"guess": g.kind, "guard": guard, "e_learned": g.e_learned, "e_sad": g.e_sad,
                           "guard_band_eh": g.band, "guard_refined_at_fp64": g.refined,
                           "err_learned": g.err_learned, "seconds": g.seconds,
                           "incremental_iterations": g.outcome.incremental_iterations,
                           "precisions": [int(x["prec"]) for x in g.outcome.log],
                           "diag_ms": [round(float(x["diag_ms"]), 2) for x in g.outcome.log],
                           "log": [{k: float(x[k]) for k in ("dE", "diis_max", "dm_err", "eev_err", "jk_ms", "iter_ms",
                                                             "dref")} for x in g.outcome.log]}

    # the target's rules read literally: the learned guess kept only if its first energy lies below
    # SAD's E(D1) (computed first), FP64 J/K after the DIIS switch (one full FP64 reference build,
    # then FP64-referenced increments, checked against full FP64 builds by the inc_check tests)
    literal_guard = "energy"
    for _ in range(args.warmup):
        plain()
        mixed()
        mixed(guard=literal_guard)
    r.profile_enable(True)
    r.profile_read_detail()   # reset
    tp, tm, tl, runs = [], [], [], []
    torch.cuda.synchronize()
    tw0 = time.perf_counter()
    for _ in range(args.steps):
        torch.cuda.synchronize()
        a = time.perf_counter()
        op, ip = plain()
        torch.cuda.synchronize()
        b = time.perf_counter()
        om, im = mixed()
        torch.cuda.synchronize()
        c = time.perf_counter()
        ol, il = mixed(guard=literal_guard)
        torch.cuda.synchronize()
        d = time.perf_counter()
        tp.append(b - a)
        tm.append(c - b)
        tl.append(d - c)
        runs.append((op, ip, om, im, ol, il))
    wall = time.perf_counter() - tw0
    detail_prof = r.profile_read_detail()
    r.profile_enable(False)
    op, ip, om, im, ol, il = runs[-1]
    t_plain, t_mixed, t_literal = float(np.median(tp)), float(np.median(tm)), float(np.median(tl))
    # the energy guard's own error: SAD's E(D1) at the guard's precision against FP64
    da, db = gs.sad_density(mol)
    e1_guard = rs.guess_step(r, S, H, enuc, (mol.nelectron,), da + db, guard_prec, want_e1=True,
                             want_outputs=False, factors=gs.sad_factors(mol))[4]
    e1_fp64 = rs.guess_step(r, S, H, enuc, (mol.nelectron,), da + db, jk.PREC_FP64, want_e1=True,
                            want_outputs=False, factors=gs.sad_factors(mol))[4]
    roofline = roofline_from_detail(detail_prof)
    # outside the timed region: the same mixed ladder from SAD (REST's guess step inside rscf_rhf),
    # to separate what the learned guess and what the ladder contribute
    torch.cuda.synchronize()
    a = time.perf_counter()
    da, db = gs.sad_density(mol)
    oms = rs.rhf(r, S, H, enuc, mol.nelectron, da + db, rs.mixed_ctrl(
        jk_prec=mixed_prec, pseudo_diag_err=args.pseudo_diag_err,
        jk_prec_increment=-1 if args.increment_prec == "none" else jk.PREC_NAMES[args.increment_prec]))
    torch.cuda.synchronize()
    mixed_sad = {"scf_s": time.perf_counter() - a, "iterations": oms.iterations, "energy": oms.energy,
                 "converged": oms.converged, "precisions": [int(x["prec"]) for x in oms.log]}
    # outside the timed region: the other acceptance rule and the tail without increments, each
    # timed on its second run, as ratios to the timed plain median
    variants = {}
    for name, kw in (("fp64_tail_without_increments", {"increment_prec": "none"}),
                     ("guard_energy_and_fp64_tail_without_increments", {"guard": "energy", "increment_prec": "none"})):
        for _ in range(2):
            torch.cuda.synchronize()
            a = time.perf_counter()
            ov, iv = mixed(**kw)
            torch.cuda.synchronize()
            tv = time.perf_counter() - a
        variants[name] = {"scf_s": tv, "ratio": t_plain / tv, "iterations": ov.iterations,
                          "fp64_iterations": ov.fp64_iterations, "energy": ov.energy, "dE_vs_plain": ov.energy - op.energy,
                          "guard": iv["guard"], "guess": iv["guess"], "sad_energy_s": iv["seconds"]["sad_energy"]}

    cpu = None
    if not args.no_cpu_baseline:
        shape = wl.H2O_64 if n == 64 else wl.H2O_16
        c_occ = wl.random_orthonormal(seed, shape.nbf, shape.nocc)
        cb = cpu_baseline(shape, seed, c_occ, c_occ @ c_occ.T, args.cpu_threads)
        per_build = cb["j_seconds_scaled"] + cb["k_dm_seconds_scaled"]
        cpu = {"value": per_build * (op.iterations + 1), "unit": "s (plain FP64 SCF, J/K only)",
               "cores": cb["cores"], "nproc": cb["nproc"], "kind": cb["kind"], "route": cb["route"],
               "j_seconds_scaled": cb["j_seconds_scaled"], "k_dm_seconds_scaled": cb["k_dm_seconds_scaled"],
               "k_occ_seconds_scaled": cb["k_seconds_scaled"], "full_medians_s": cb["full_medians_s"],
               "sample": cb["sample"] + f"; times the {op.iterations + 1} J/K builds of the plain FP64 SCF "
                                        "(diagonalisation and DIIS not included)",
               "gpu_plain_scf_s": t_plain}
    out = {
        "metric": "SCF time to 1e-8 Eh, (H2O)64 def2-SVP: mixed+learned vs plain FP64, same GPU"
                  if n == 64 else "SCF time to 1e-8 Eh, (H2O)16 def2-SVP: mixed+learned vs plain FP64, same GPU",
        "value": t_plain / t_literal,
        "unit": "x (plain FP64 SCF seconds / mixed+learned SCF seconds; learned guess kept by the energy rule "
                "against SAD's E(D1))",
        "value_commutator_guard": t_plain / t_mixed,
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * (t_plain + t_mixed + t_literal),
        "ms_per_step_literal": 1e3 * (t_plain + t_literal),
        "wall_s_timed_region": wall,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64 (plain); " + ("f32" if args.mixed_prec == "fp32" else "split-bf16 K (2, then 3 planes) + f32 J")
                 + " -> one f64 J/K reference build"
                 + ("" if args.increment_prec == "none" else f" + {args.increment_prec} J/K increments on it")
                 + ", f64 elsewhere (mixed)",
        "data": "synthetic (seeded water cluster geometry; integrals from this build's engine)",
        "config": {"workload": f"(H2O){n} seed {seed}, def2-SVP / def2-SVP-JKFIT, nbf={mol.nbf} naux={mol.naux} "
                               f"nocc={mol.nelectron // 2}; RI-HF to |dE| <= 1e-8 Eh, ||dD||/nbf <= 1e-5, "
                               "||d eps|| <= 1e-5 (REST defaults); FP64+FP32"
                               + ("+BF16-plane" if store & jk.STORE_BF16 else "") + " B resident in HBM",
                   "mixed_prec": args.mixed_prec, "increment_prec": args.increment_prec,
                   "pseudo_diag_err": args.pseudo_diag_err,
                   "guard": f"energy rule at {args.guard_prec} (value); {args.guard} (value_commutator_guard)",
                   "nbf": mol.nbf, "naux": mol.naux, "nocc": mol.nelectron // 2},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "detail": {
            "plain_fp64": {"scf_s": t_plain, "iterations": op.iterations, "energy": op.energy,
                           "converged": op.converged, "jk_s": op.jk_seconds, "diag_s": op.diag_seconds, **ip},
            "mixed_learned": {"scf_s": t_mixed, "iterations": om.iterations, "fp64_iterations": om.fp64_iterations,
                              "energy": om.energy, "converged": om.converged, "jk_s": om.jk_seconds,
                              "diag_s": om.diag_seconds, **im},
            "mixed_learned_energy_rule": {"scf_s": t_literal, "iterations": ol.iterations,
                                          "fp64_iterations": ol.fp64_iterations,
                                          "incremental_iterations": ol.incremental_iterations,
                                          "reference_builds": ol.reference_builds, "energy": ol.energy,
                                          "converged": ol.converged, "jk_s": ol.jk_seconds, "diag_s": ol.diag_seconds,
                                          **il,
                                          "guard_margin_eh": (il["e_sad"] - il["e_learned"]) if il["e_sad"] else None,
                                          "guard_margin_vs_fp64_e1_eh": e1_fp64 - il["e_learned"],
                                          "guard_prec": args.guard_prec, "guard_e1": e1_guard, "guard_e1_fp64": e1_fp64,
                                          "guard_e1_error_eh": e1_guard - e1_fp64},
            "dE_mixed_minus_plain": om.energy - op.energy,
            "dE_energy_rule_minus_plain": ol.energy - op.energy,
            "all_steps_plain_s": tp, "all_steps_mixed_s": tm, "all_steps_energy_rule_s": tl,
            "mixed_from_sad_untimed": mixed_sad,
            "variants_untimed": variants,
            "setup_s": {"int1e": t1 - t0, "b_build": t2 - t1, "b_build_parts": binfo, "sad_atoms_and_model": t3 - t2},
            "kernels": {f"{p}:{ph}": {"ms": ms, "launches": k, "work": w} for (p, ph), (ms, k, w) in
                        detail_prof.items() if k},
            "model": getattr(net, "meta", {}),
        },
    }
    print(json.dumps(out))
    net.close()
    r.close()


def main_batched(args):
    """Ragged J/K of many water clusters ((H2O)n, n uniform in [2, 32], workload.batched_cluster_sizes)
    in grouped persistent launches replayed from a hipGraph, in device-memory waves planned by the
    library (rjk_batch_plan_waves); autotuned tiles per wave.  A step is one J and one K of every
    cluster; the value is clusters per second over the summed replay times of all waves."""
    import torch
    from rest_amd import jk
    from rest_amd import workload as wl

    sizes = wl.batched_cluster_sizes(args.clusters)
    nbf = [24 * n for n in sizes]
    naux = [113 * n for n in sizes]
    nocc = [5 * n for n in sizes]
    prec = jk.PREC_NAMES[args.batch_prec]
    waves = jk.batch_plan_waves(nbf, naux, nocc, prec, int(args.batch_budget_gib * (1 << 30)))
    stream = torch.cuda.Stream(device=torch.device("cuda:0"))
    torch.cuda.set_stream(stream)
    es = 8 if prec == jk.PREC_FP64 else 4
    t_total, tuned, stage_ms = 0.0, [], np.zeros(jk.BATCH_NUM_STAGES)
    work = {"j": 0.0, "half": 0.0, "syrk": 0.0}
    for wi, (i0, i1) in enumerate(waves):
        b = jk.Batch(nbf[i0:i1], naux[i0:i1], nocc[i0:i1], prec)
        b.set_stream(stream.cuda_stream)
        b.fill_synthetic(1000 + i0)
        for i in range(i0, i1):
            c = wl.random_orthonormal(i + 1, nbf[i], nocc[i])
            b.set_inputs(i - i0, 2.0 * c @ c.T, c, np.full(nocc[i], 2.0))
        tuned.append((b.autotune(reps=1), b.config()))
        for _ in range(max(1, args.warmup)):
            b.run(1.0, -0.5)
        b.synchronize()
        ev0 = torch.cuda.Event(enable_timing=True)
        ev1 = torch.cuda.Event(enable_timing=True)
        ev0.record(stream)
        for _ in range(args.steps):
            b.run(1.0, -0.5)
        ev1.record(stream)
        torch.cuda.synchronize()
        t_total += ev0.elapsed_time(ev1) / 1e3 / args.steps
        stage_ms += np.array(b.profile(1.0, -0.5))      # one staged launch sequence, events between stages
        for i in range(i0, i1):
            n_, o_, a_ = nbf[i], nocc[i], naux[i]
            npair = n_ * (n_ + 1) // 2
            work["j"] += 2.0 * npair * a_ * es + 8.0 * (n_ * n_ + npair) + 16.0 * a_
            work["half"] += 2.0 * n_ * n_ * o_ * a_
            work["syrk"] += n_ * (n_ + 1.0) * o_ * a_
        vj, vk = b.get_outputs(0)
        assert np.isfinite(vj).all() and np.isfinite(vk).all()
        b.close()
        print(f"[bench] wave {wi + 1}/{len(waves)}: clusters {i0}..{i1 - 1}", file=sys.stderr, flush=True)
    st = dict(zip(jk.BATCH_STAGES, stage_ms.tolist()))
    j_ms = st["gamma"] + st["j_acc"]
    half_ms = st["k_half_narrow"] + st["k_half_wide"]
    syrk_ms = st["k_syrk_narrow"] + st["k_syrk_wide"]
    kmult = BF16_PRODUCTS.get(args.batch_prec, 1)
    kpeak = MFMA_PEAK_TFS["bf16" if args.batch_prec.startswith("bf16") else args.batch_prec]
    cand = {"k_half": (half_ms, "mfma", work["half"]), "k_syrk": (syrk_ms, "mfma", work["syrk"]),
            "j": (j_ms, "hbm", work["j"])}
    dom = max(cand, key=lambda k: cand[k][0])
    ms, bound, amount = cand[dom]
    if bound == "hbm":
        ach, peak, unit = amount / (ms / 1e3) / 1e9, HBM_PEAK_GBS, "GB/s"
    else:
        ach, peak, unit = kmult * amount / (ms / 1e3) / 1e12, kpeak, "TFLOP/s"
    out = {
        "metric": f"batched RI-V J/K builds/s, {args.clusters} water clusters (H2O)n, n in [2, 32], "
                  f"def2-SVP/JKFIT shapes, {args.batch_prec}",
        "value": args.clusters / t_total, "unit": "J/K builds/s (one J + one K per cluster)", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": 1e3 * t_total, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None,
        "dtype": {"fp64": "f64", "fp32": "f32 (B, MFMA) with f64 J and split-K sums"}.get(
            args.batch_prec, f"{args.batch_prec}: B in f32, K on split-bf16 MFMA with f32 accumulation, J f32 B / f64 sums"),
        "data": "synthetic B (seeded hash generator), D = 2 C C^T from seeded orthonormal C",
        "config": {"workload": f"{args.clusters} clusters in {len(waves)} waves of grouped persistent launches "
                               "(hipGraph replay), autotuned tiles", "clusters": args.clusters,
                   "waves": len(waves), "budget_gib": args.batch_budget_gib,
                   "device_bytes_total": jk.batch_device_bytes(nbf, naux, nocc, prec)},
        "roofline": {"bound": bound, "achieved": ach, "peak": peak, "unit": unit, "frac": ach / peak,
                     "traffic": None, "kernel": f"{dom} stage (grouped, all waves)",
                     "algorithmic_total": amount, "stage_ms_total": ms, "mfma_products_per_product": kmult},
        "cpu_baseline": None,
        "detail": {"autotune": [{"best_ms": t, "khalf_cut": c[0], "waves_per_cu": c[1]} for t, c in tuned],
                   "stage_ms_summed_over_waves": st,
                   "j_GBps": work["j"] / (j_ms / 1e3) / 1e9 if j_ms else None,
                   "k_half_TFLOPs_algorithmic": work["half"] / (half_ms / 1e3) / 1e12 if half_ms else None,
                   "k_syrk_TFLOPs_algorithmic": work["syrk"] / (syrk_ms / 1e3) / 1e12 if syrk_ms else None,
                   "staged_ms_total": float(stage_ms.sum()), "graph_ms_total": 1e3 * t_total},
    }
    print(json.dumps(out))


def main_jk(args):
    import torch
    from rest_amd import jk
    from rest_amd import workload as wl

    shape = wl.H2O_64 if args.workload == "h2o64" else wl.H2O_16
    seed = 2 if args.workload == "h2o64" else 1
    dev = torch.device("cuda:0")
    nbf, naux, nocc = shape.nbf, shape.naux, shape.nocc

    c = wl.random_orthonormal(seed, nbf, nocc)
    d = c @ c.T
    prec = jk.PREC_NAMES[args.prec]
    bf = args.prec.startswith("bf16")
    ri = jk.DeviceRI(nbf, naux)
    ri.fill_synthetic(seed, jk.STORE_FP64 | (jk.STORE_FP32 if prec != jk.PREC_FP64 else 0)
                      | (jk.STORE_BF16 if bf else 0))
    # a non-default torch stream (default priority) shared by the library and
    # the timing events: the null stream would hand the library back its own
    stream = torch.cuda.Stream(device=dev)
    torch.cuda.set_stream(stream)
    ri.set_stream(stream.cuda_stream)

    d_dev = torch.from_numpy(np.ascontiguousarray(d.T)).to(dev)            # column-major D
    c_dev = torch.from_numpy(np.ascontiguousarray(c.T)).to(dev)            # column-major C
    occ_dev = torch.ones(nocc, dtype=torch.float64, device=dev)
    vj_dev = torch.empty(shape.npair, dtype=torch.float64, device=dev)
    vk_dev = torch.empty(shape.npair, dtype=torch.float64, device=dev)

    if args.calibrate_copy_mib:
        nbytes = args.calibrate_copy_mib << 20
        src = torch.ones(nbytes // 8, dtype=torch.float64, device=dev)
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        ri.copy_dev(dst.data_ptr(), src.data_ptr(), nbytes)
        ri.synchronize()
        assert bool((dst == 1).all())
        del src, dst

    def step():
        ri.vj_dev(d_dev.data_ptr(), 1, 1.0, vj_dev.data_ptr(), prec=prec)
        ri.vk_occ_dev(c_dev.data_ptr(), nocc, [nocc - 1], occ_dev.data_ptr(), 1, 1.0, vk_dev.data_ptr(), prec=prec)

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    ri.profile_enable(True)
    ri.profile_read()  # reset
    ev0 = torch.cuda.Event(enable_timing=True)
    ev1 = torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ev0.record(stream)
    for _ in range(args.steps):
        step()
    ev1.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms_total = ev0.elapsed_time(ev1)
    phases = ri.profile_read()
    ri.profile_enable(False)
    assert np.isfinite(vj_dev.cpu().numpy()).all() and np.isfinite(vk_dev.cpu().numpy()).all()
    errors = None
    if prec != jk.PREC_FP64:   # J and K of the same step at FP64 (outside the timed region)
        vj_lo, vk_lo = vj_dev.cpu().numpy(), vk_dev.cpu().numpy()
        ri.vj_dev(d_dev.data_ptr(), 1, 1.0, vj_dev.data_ptr(), prec=jk.PREC_FP64)
        ri.vk_occ_dev(c_dev.data_ptr(), nocc, [nocc - 1], occ_dev.data_ptr(), 1, 1.0, vk_dev.data_ptr(),
                      prec=jk.PREC_FP64)
        vj64, vk64 = vj_dev.cpu().numpy(), vk_dev.cpu().numpy()
        errors = {"j_max_rel_err_vs_fp64": float(np.abs(vj_lo - vj64).max() / np.abs(vj64).max()),
                  "k_max_rel_err_vs_fp64": float(np.abs(vk_lo - vk64).max() / np.abs(vk64).max())}

    ms_step = ms_total / args.steps
    kf = wl.k_flops(shape)
    s_b = 8 if args.prec == "fp64" else 4
    jbytes = wl.jk_bytes(shape, s_b)
    per_kernel = {}
    for name, (ms, n) in phases.items():
        if n:
            per_kernel[name] = {"ms_per_step": ms / args.steps, "launches_per_step": n / args.steps,
                                "avg_launch_ms": ms / n}
    # algorithmic work per step of each kernel
    work = {"j_gamma": ("hbm", shape.npair * naux * s_b + 8 * shape.npair),
            "j_acc": ("hbm", shape.npair * naux * s_b + 8 * shape.npair),
            "k_half": ("mfma", kf["half"]), "k_syrk": ("mfma", kf["syrk"])}
    dom = max((k for k in per_kernel if k in work), key=lambda k: per_kernel[k]["ms_per_step"])
    bound, amount = work[dom]
    pk = per_kernel[dom]
    per_launch = amount / pk["launches_per_step"]
    sec = pk["avg_launch_ms"] / 1e3
    kpeak = MFMA_PEAK_TFS["bf16" if bf else args.prec]
    kmult = BF16_PRODUCTS.get(args.prec, 1)       # executed MFMA products per algorithmic product
    if bound == "hbm":
        ach, peak, unit = per_launch / sec / 1e9, HBM_PEAK_GBS, "GB/s"
    else:
        ach, peak, unit = kmult * per_launch / sec / 1e12, kpeak, "TFLOP/s"
    kname = PHASE_KERNEL[dom]
    if bf and dom.startswith("k_"):
        kname = kname.replace("k_", "k_bf_", 1) + f"<{args.prec[-1]}>"
    roofline = {"bound": bound, "achieved": ach, "peak": peak, "unit": unit, "frac": ach / peak,
                "traffic": load_traffic(kname), "kernel": kname, "algorithmic_per_launch": per_launch}
    # secondary figures: J against HBM, K (both stages) against FP64 MFMA
    j_ms = sum(per_kernel.get(k, {}).get("ms_per_step", 0.0) for k in ("j_gamma", "j_acc"))
    k_ms = sum(per_kernel.get(k, {}).get("ms_per_step", 0.0) for k in ("k_half", "k_syrk"))
    detail = {
        "j_build_hbm_GBps": jbytes / (j_ms / 1e3) / 1e9 if j_ms else None,
        "j_build_hbm_frac": (jbytes / (j_ms / 1e3) / 1e9) / HBM_PEAK_GBS if j_ms else None,
        "k_build_TFLOPs": kf["total"] / (k_ms / 1e3) / 1e12 if k_ms else None,
        "k_build_mfma_frac": (kmult * kf["total"] / (k_ms / 1e3) / 1e12) / kpeak if k_ms else None,
        "per_kernel": per_kernel,
    }

    cpu = None
    if not args.no_cpu_baseline and args.prec == "fp64":
        cpu = cpu_baseline(shape, seed, c, d, args.cpu_threads)

    out = {
        "metric": f"RI-V J/K builds/s (one J + one K), {shape.name} shape, "
                  + {"fp64": "plain FP64", "fp32": "FP32 B + FP32 MFMA", "bf16x3": "K on 3 BF16 planes, FP32 J",
                     "bf16x2": "K on 2 BF16 planes, FP32 J", "bf16x1": "K on 1 BF16 plane, FP32 J"}[args.prec],
        "value": args.steps / (ms_total / 1e3),
        "unit": "J/K builds/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "wall_s_timed_region": wall,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": {"fp64": "f64", "fp32": "f32"}.get(args.prec, "split-bf16 K, f32 J"),
        "data": "synthetic",
        "config": {"workload": f"{shape.name}: nbf={nbf} naux={naux} npair={shape.npair} nocc={nocc}, "
                               f"packed FP64 B resident in HBM (seed {seed}), D=C C^T, K occupied route",
                   "nbf": nbf, "naux": naux, "nocc": nocc, "nspin": 1},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "detail": {**detail, "errors_vs_fp64": errors},
    }
    print(json.dumps(out))
    ri.close()


if __name__ == "__main__":
    main()

"""Train the learned initial guess on densities this SCF produces, then compare
it with SAD on held-out water clusters.

    python -m rest_amd.train_guess --clusters 4096 --workers 4 --budget-gib 45 --setup-threads 2 --epochs 60 \
        --out rest_amd/models/learned_guess_h2o.npz
    python -m rest_amd.train_guess --eval-only --model rest_amd/models/learned_guess_h2o.npz \\
        --test-n 64 --test-seeds 100 101 ... --report test.json

Training clusters: learned.training_cluster(i), i < clusters: (H2O)n with n
uniform in [2, 24], monomer O–H ±0.05 Å and H–O–H ±5°, def2-SVP /
def2-SVP-JKFIT, converged from the SAD guess.  The default generator
(--generator batched) is the batched SCF (batch_scf.rhf_batch): REST's loop
per cluster with the FP64 J/K of all clusters of a memory wave in one grouped
launch per iteration, replayed from a hipGraph.  --generator single converges
one cluster at a time with the single-problem device SCF.  With --workers W the
clusters are split into W contiguous ranges, each converged by its own process
on the same GPU (the per-cluster set-up — one-electron integrals on the host,
the B build, the SAD guess step — runs in every process at once; each batched
worker keeps its waves within --budget-gib).  Generation stops at the
--gen-seconds deadline (before the next chunk of clusters, or, for a
single-generator worker, the next cluster) and the training uses the clusters
done by then.
Held-out test: (H2O)n clusters from workload.water_cluster with seeds not used
in training (--test-n, --test-seeds); for each, the FP64 SCF from SAD and the
SCF from the learned guess (learned.learned_scf with the --guard acceptance
rule) at FP64 and, with --mixed-eval or --mixed-only, on the mixed-precision
ladder (scf.mixed_ctrl), with iterations, guess energies and errors of both
guesses (the learned prediction's and SAD's first orbital density's) and final
energies.
"""
from __future__ import annotations

import argparse
import json
import os
import subprocess
import sys
import tempfile
import time
from pathlib import Path

import numpy as np


def log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def evaluate(net, n: int, seeds, prec_mixed: bool = False, guard: str = "error") -> list:
    from . import guess as gs
    from . import jk, ri
    from . import learned as lg
    from . import scf as rs
    from . import workload as wl
    rows = []
    for seed in seeds:
        m = ri.Molecule.water_cluster(wl.water_cluster(seed, n))
        S, T, V = m.int1e()
        H = T + V
        store = jk.STORE_FP64 | (jk.STORE_FP32 | jk.STORE_BF16 if prec_mixed else 0)
        with jk.DeviceRI(m.nbf, m.naux) as r:
            m.build_rimatr(r, 1e-8, store)
            t0 = time.perf_counter()
            da, db = gs.sad_density(m)
            e0, _, c1, _ = rs.guess_step(r, S, H, m.nuclear_repulsion(), (m.nelectron,), da + db)
            o_sad = rs.rhf(r, S, H, m.nuclear_repulsion(), m.nelectron, None, rs.default_ctrl(), c_guess=c1)
            t1 = time.perf_counter()
            ctrl = rs.mixed_ctrl() if prec_mixed else rs.default_ctrl()
            g = lg.learned_scf(net, m, r, S, H, ctrl, guard=guard)
            t2 = time.perf_counter()
        o_l = g.outcome
        row = {"n": n, "seed": seed, "sad_iterations": o_sad.iterations, "sad_energy": o_sad.energy,
               "sad_guess_energy": e0, "sad_seconds": t1 - t0, "learned_kind": g.kind, "guard": guard,
               "learned_guess_energy": g.e_learned, "sad_d1_energy": o_sad.guess_energy,
               "learned_guess_error": g.err_learned, "sad_d1_error": o_sad.guess_error,
               "learned_electrons": g.electrons,
               "learned_iterations": o_l.iterations, "learned_fp64_iterations": o_l.fp64_iterations,
               "learned_energy": o_l.energy, "learned_seconds": t2 - t1, "dE": o_l.energy - o_sad.energy,
               "learned_incremental_iterations": o_l.incremental_iterations,
               "converged": bool(o_sad.converged and o_l.converged), "mixed": prec_mixed}
        log(json.dumps(row))
        rows.append(row)
    return rows


def summarize(rows: list) -> dict:
    """Means and extremes of an evaluation's rows (per precision mode)."""
    out = {}
    for mode in sorted({bool(r["mixed"]) for r in rows}):
        rs_ = [r for r in rows if bool(r["mixed"]) == mode]
        col = lambda k: np.array([r[k] for r in rs_ if r.get(k) is not None], dtype=float)
        rng = lambda k: [float(col(k).min()), float(col(k).mean()), float(col(k).max())] if len(col(k)) else None
        out["mixed" if mode else "fp64"] = {
            "clusters": len(rs_), "kept": sum(r["learned_kind"] == "learned" for r in rs_),
            "converged": sum(bool(r["converged"]) for r in rs_),
            "sad_iterations_mean": float(col("sad_iterations").mean()),
            "learned_iterations_mean": float(col("learned_iterations").mean()),
            "max_abs_dE": float(np.abs(col("dE")).max()),
            "sad_seconds_mean": float(col("sad_seconds").mean()),
            "learned_seconds_mean": float(col("learned_seconds").mean()),
            "learned_guess_error_min_mean_max": rng("learned_guess_error"),
            "sad_d1_error_min_mean_max": rng("sad_d1_error"),
        }
        gap = np.array([r["learned_guess_energy"] - r["sad_d1_energy"] for r in rs_
                        if r.get("learned_guess_energy") is not None and r.get("sad_d1_energy") is not None])
        if len(gap):
            out["mixed" if mode else "fp64"]["learned_minus_sad_d1_energy_min_mean_max"] = \
                [float(gap.min()), float(gap.mean()), float(gap.max())]
    return out


def _generate_shard(a) -> int:
    """Worker: converge clusters [i0, i1) (single generator: until the deadline), save (X, T, M) and the
    generation statistics to --shard-out."""
    from . import jk
    from . import learned as lg
    i0, i1 = a.shard
    stats = {}
    say = lambda m: log(f"[worker {i0}] {m}")
    if a.generator == "batched":
        X, T, M = lg.make_training_data_batched(range(i0, i1), a.base_seed, log=say, stats=stats,
                                                prec=jk.PREC_NAMES[a.gen_prec],
                                                budget_bytes=int(a.budget_gib * (1 << 30)),
                                                setup_threads=a.setup_threads, deadline=a.deadline)
    else:
        X, T, M = lg.make_training_data(range(i0, i1), a.base_seed, log=say, deadline=a.deadline, stats=stats)
    np.savez(a.shard_out, X=X, T=T, M=M.astype(np.uint8), done=np.int64(stats.get("clusters_done", 0)),
             stats=json.dumps(stats))
    log(f"[worker {i0}] saved {stats.get('clusters_done', 0)} clusters, {len(X)} pairs")
    return 0


def _generate_parallel(a):
    """Spawn the workers (this process has not touched the GPU), wait, and join their shards."""
    tmp = Path(tempfile.mkdtemp(prefix="lg_shards_"))
    deadline = time.time() + a.gen_seconds
    env = dict(os.environ)
    env["OMP_NUM_THREADS"] = str(max(1, a.cpu_threads // a.workers))
    edges = np.linspace(0, a.clusters, a.workers + 1).astype(int)
    procs = []
    for w in range(a.workers):
        out = tmp / f"shard{w}.npz"
        cmd = [sys.executable, "-u", "-m", "rest_amd.train_guess", "--shard", str(edges[w]), str(edges[w + 1]),
               "--shard-out", str(out), "--deadline", repr(deadline), "--base-seed", str(a.base_seed),
               "--out", str(out), "--generator", a.generator, "--gen-prec", a.gen_prec,
               "--budget-gib", repr(a.budget_gib), "--setup-threads", str(a.setup_threads)]
        procs.append((subprocess.Popen(cmd, env=env), out))
    Xs, Ts, Ms, done, stats = [], [], [], 0, {}
    for p, out in procs:
        rc = p.wait()
        if rc != 0:
            raise RuntimeError(f"data worker for {out.name} exited with {rc}")
        with np.load(out, allow_pickle=False) as z:
            Xs.append(z["X"])
            Ts.append(z["T"])
            Ms.append(z["M"].astype(np.float32))
            done += int(z["done"])
            for k, v in json.loads(str(z["stats"])).items():
                stats[k] = stats.get(k, 0) + v if isinstance(v, (int, float)) else v
        out.unlink()
    tmp.rmdir()
    return np.concatenate(Xs), np.concatenate(Ts), np.concatenate(Ms), done, stats


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--clusters", type=int, default=256)
    ap.add_argument("--base-seed", type=int, default=4096)
    ap.add_argument("--generator", default="batched", choices=["batched", "single"],
                    help="training SCFs through the batched grouped J/K, or one cluster at a time")
    ap.add_argument("--gen-prec", default="fp64", choices=["fp64", "fp32", "bf16x3"],
                    help="batched generator: J/K precision of the training SCFs")
    ap.add_argument("--workers", type=int, default=1, help="data-generation processes sharing the GPU")
    ap.add_argument("--budget-gib", type=float, default=0.0,
                    help="batched generator: device memory per batch wave in each process (0: from free memory)")
    ap.add_argument("--setup-threads", type=int, default=16,
                    help="batched generator: clusters whose integrals and B are built at once per process")
    ap.add_argument("--cpu-threads", type=int, default=16, help="host threads split among the workers")
    ap.add_argument("--gen-seconds", type=float, default=1e9, help="data-generation deadline (seconds)")
    ap.add_argument("--epochs", type=int, default=40)
    ap.add_argument("--batch", type=int, default=4096)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--hidden", type=int, default=256)
    ap.add_argument("--layers", type=int, default=3)
    ap.add_argument("--out", type=Path, default=None)
    ap.add_argument("--eval-only", action="store_true", help="evaluate --model on the test clusters")
    ap.add_argument("--model", type=Path, default=None)
    ap.add_argument("--test-n", type=int, default=64)
    ap.add_argument("--test-seeds", type=int, nargs="*", default=[2])
    ap.add_argument("--mixed-eval", action="store_true", help="also evaluate the mixed-precision SCF")
    ap.add_argument("--mixed-only", action="store_true", help="evaluate the mixed-precision SCF only")
    ap.add_argument("--guard", default="error", choices=["error", "energy"], help="learned-guess acceptance rule")
    ap.add_argument("--report", type=Path, default=None)
    ap.add_argument("--merge", type=Path, nargs="+", default=None,
                    help="join the rows of these evaluation reports into --report, with a summary")
    ap.add_argument("--shard", type=int, nargs=2, default=None, help=argparse.SUPPRESS)
    ap.add_argument("--shard-out", type=Path, default=None, help=argparse.SUPPRESS)
    ap.add_argument("--deadline", type=float, default=None, help=argparse.SUPPRESS)
    a = ap.parse_args(argv)
    if a.shard is not None:
        return _generate_shard(a)

    if a.merge:
        parts = [json.loads(p.read_text()) for p in a.merge]
        rows = [r for p in parts for r in p["test"]]
        a.report.write_text(json.dumps({"model": parts[0].get("model"), "meta": parts[0].get("meta", {}),
                                        "guard": parts[0].get("guard"), "summary": summarize(rows), "test": rows},
                                       indent=1))
        return 0

    if a.eval_only:
        from . import learned as lg
        net = lg.Net.load(a.model)
        rows = [] if a.mixed_only else evaluate(net, a.test_n, a.test_seeds, guard=a.guard)
        if a.mixed_eval or a.mixed_only:
            rows += evaluate(net, a.test_n, a.test_seeds, prec_mixed=True, guard=a.guard)
        if a.report:
            a.report.write_text(json.dumps({"model": str(a.model), "meta": getattr(net, "meta", {}), "guard": a.guard,
                                            "summary": summarize(rows), "test": rows}, indent=1))
        return 0

    if a.out is None:
        ap.error("--out is required for training")
    t0 = time.time()
    log(f"generating training data: {a.clusters} clusters, generator {a.generator}")
    gen_stats = {}
    if a.workers > 1:
        X, T, M, done, gen_stats = _generate_parallel(a)
    elif a.generator == "batched":
        from . import learned as lg
        from . import jk
        X, T, M = lg.make_training_data_batched(range(a.clusters), a.base_seed, log=log, stats=gen_stats,
                                                prec=jk.PREC_NAMES[a.gen_prec],
                                                budget_bytes=int(a.budget_gib * (1 << 30)),
                                                setup_threads=a.setup_threads, deadline=t0 + a.gen_seconds)
        done = gen_stats["clusters_done"]
    else:
        from . import learned as lg
        stats = {}
        X, T, M = lg.make_training_data(range(a.clusters), a.base_seed, log=log,
                                        deadline=time.time() + a.gen_seconds, stats=stats)
        done = stats["clusters_done"]
    t_gen = time.time() - t0
    from . import learned as lg
    log(f"{done} clusters, {len(X)} pair samples in {t_gen:.0f}s; training {a.layers}x{a.hidden} SiLU MLP")
    net = lg.Net(lg.N_IN, lg.N_OUT, a.hidden, a.layers, max_batch=a.batch, seed=0)
    hist = lg.train(net, X, T, M, epochs=a.epochs, batch=a.batch, lr=a.lr, log=log)
    meta = {"clusters": int(done), "clusters_requested": a.clusters, "base_seed": a.base_seed,
            "samples": int(len(X)), "epochs": a.epochs, "batch": a.batch, "lr": a.lr, "generator": a.generator,
            "gen_prec": a.gen_prec,
            "workers": a.workers, "generation_stats": gen_stats,
            "train_loss": hist["train"][-1], "val_loss": hist["val"][-1], "generation_seconds": t_gen,
            "train_seconds": time.time() - t0 - t_gen}
    a.out.parent.mkdir(parents=True, exist_ok=True)
    net.save(a.out, meta)
    log(f"saved {a.out}: {json.dumps(meta)}")
    rows = []
    if a.test_seeds:
        rows = evaluate(net, a.test_n, a.test_seeds)
        if a.mixed_eval:
            rows += evaluate(net, a.test_n, a.test_seeds, prec_mixed=True)
    if a.report:
        a.report.write_text(json.dumps({"meta": meta, "history": hist, "test": rows}, indent=1))
    return 0


if __name__ == "__main__":
    sys.exit(main())

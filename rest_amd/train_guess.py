"""Train the learned initial guess on densities this SCF produces, then compare
it with SAD on held-out water clusters.

    python -m rest_amd.train_guess --clusters 256 --epochs 40 --out rest_amd/models/learned_guess_h2o.npz

Training clusters: learned.training_cluster(i), i < clusters: (H2O)n with n
uniform in [2, 24], monomer O–H ±0.05 Å and H–O–H ±5°, def2-SVP /
def2-SVP-JKFIT, converged with the FP64 device SCF from the SAD guess.
Held-out test: (H2O)n clusters from workload.water_cluster with seeds not used
in training (--test-n, --test-seeds); for each, SCF iterations and guess
energies with SAD and with the learned guess.
"""
from __future__ import annotations

import argparse
import json
import sys
import time
from pathlib import Path

import numpy as np


def log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def evaluate(net, n: int, seeds, prec_mixed: bool = False) -> list:
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
        store = jk.STORE_FP64 | (jk.STORE_FP32 if prec_mixed else 0)
        with jk.DeviceRI(m.nbf, m.naux) as r:
            m.build_rimatr(r, 1e-8, store)
            t0 = time.perf_counter()
            da, db = gs.sad_density(m)
            e0, _, c1, _ = rs.guess_step(r, S, H, m.nuclear_repulsion(), (m.nelectron,), da + db)
            o_sad = rs.rhf(r, S, H, m.nuclear_repulsion(), m.nelectron, None, rs.default_ctrl(), c_guess=c1)
            t1 = time.perf_counter()
            g = lg.learned_guess(net, m, r, S, H, jk.PREC_FP32 if prec_mixed else jk.PREC_FP64)
            ctrl = rs.default_ctrl(jk_prec=jk.PREC_FP32 if prec_mixed else jk.PREC_FP64)
            o_l = rs.rhf(r, S, H, m.nuclear_repulsion(), m.nelectron, None, ctrl, c_guess=g.c1)
            t2 = time.perf_counter()
        row = {"n": n, "seed": seed, "sad_iterations": o_sad.iterations, "sad_energy": o_sad.energy,
               "sad_guess_energy": e0, "sad_seconds": t1 - t0, "learned_kind": g.kind,
               "learned_guess_energy": g.e_learned, "learned_electrons": g.electrons,
               "learned_iterations": o_l.iterations, "learned_energy": o_l.energy, "learned_seconds": t2 - t1,
               "dE": o_l.energy - o_sad.energy, "mixed": prec_mixed}
        log(json.dumps(row))
        rows.append(row)
    return rows


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--clusters", type=int, default=256)
    ap.add_argument("--base-seed", type=int, default=4096)
    ap.add_argument("--epochs", type=int, default=40)
    ap.add_argument("--batch", type=int, default=4096)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--hidden", type=int, default=256)
    ap.add_argument("--layers", type=int, default=3)
    ap.add_argument("--out", type=Path, required=True)
    ap.add_argument("--test-n", type=int, default=64)
    ap.add_argument("--test-seeds", type=int, nargs="*", default=[2])
    ap.add_argument("--report", type=Path, default=None)
    a = ap.parse_args(argv)

    from . import learned as lg
    t0 = time.time()
    log(f"generating training data: {a.clusters} clusters")
    X, T, M = lg.make_training_data(range(a.clusters), a.base_seed, log=log)
    log(f"{len(X)} pair samples in {time.time() - t0:.0f}s; training {a.layers}x{a.hidden} SiLU MLP")
    net = lg.Net(lg.N_IN, lg.N_OUT, a.hidden, a.layers, max_batch=a.batch, seed=0)
    hist = lg.train(net, X, T, M, epochs=a.epochs, batch=a.batch, lr=a.lr, log=log)
    meta = {"clusters": a.clusters, "base_seed": a.base_seed, "samples": int(len(X)), "epochs": a.epochs,
            "batch": a.batch, "lr": a.lr, "train_loss": hist["train"][-1], "val_loss": hist["val"][-1],
            "train_seconds": time.time() - t0}
    a.out.parent.mkdir(parents=True, exist_ok=True)
    net.save(a.out, meta)
    log(f"saved {a.out}: {json.dumps(meta)}")
    rows = []
    if a.test_seeds:
        rows = evaluate(net, a.test_n, a.test_seeds) + evaluate(net, a.test_n, a.test_seeds, prec_mixed=True)
    if a.report:
        a.report.write_text(json.dumps({"meta": meta, "history": hist, "test": rows}, indent=1))
    return 0


if __name__ == "__main__":
    sys.exit(main())

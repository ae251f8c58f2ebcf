"""Write tests/golden/scf_energies.json: RI-HF energies from the oracle SCF loop
(oracle/scf_oracle.py, a restatement of REST's scf()) at tight thresholds,
with B built from this package's integral engine and V^{-1/2} by numpy.
These pin the device SCF driver (tests/test_gpu_scf.py).

Run: python tests/golden/make_scf_golden.py
"""
import json
import sys
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))

from oracle import scf_oracle as so          # noqa: E402
from rest_amd import ri, workload as wl       # noqa: E402
from tests.test_integrals import numpy_jk     # noqa: E402

TIGHT = so.ScfCtrl(scf_acc_etot=1e-11, scf_acc_rho=1e-9, scf_acc_eev=1e-8, max_scf_cycle=200)


def energy(m, ctrl=TIGHT):
    S, T, V = m.int1e()
    B, kept = so.rimatr_from_integrals(m.int3c2e_packed(), m.int2c2e(), 1e-8)
    r = so.rhf(S, T + V, B, m.nelectron, m.nuclear_repulsion(), ctrl, jk=numpy_jk(B, m.nbf))
    assert r.converged
    return r.energy, r.iterations


def main():
    out = {}
    lines = (ROOT / "tests/golden/pentacene.xyz").read_text().splitlines()[2:]
    sym = [l.split()[0] for l in lines]
    xyz = np.array([[float(x) for x in l.split()[1:]] for l in lines])
    # from the core guess pentacene needs ~60 iterations; REST's energy-rise check (which its own
    # 19-iteration run never reached) would steer this one to another stationary point
    no_check = so.ScfCtrl(**{**TIGHT.__dict__, "start_check_oscillation": 1000})
    e, it = energy(ri.Molecule.from_angstrom(sym, xyz, basis="STO-3G", auxbasis="def2-SV(P)-JKFIT"), no_check)
    out["pentacene_sto3g_def2svpjkfit"] = {"energy": e, "iterations": it}
    for seed, n in [(3, 2), (4, 4)]:
        e, it = energy(ri.Molecule.water_cluster(wl.water_cluster(seed, n)))
        out[f"water{n}_seed{seed}_def2svp"] = {"energy": e, "iterations": it}
    # the single-water workload (O at the origin, H at (±0.7572, 0.5865, 0) Å) and (H2O)16 seed 1
    mono = [l.split() for l in (ROOT / "tests/golden/h2o_monomer.xyz").read_text().splitlines()[2:]]
    e, it = energy(ri.Molecule.from_angstrom([l[0] for l in mono], np.array([[float(x) for x in l[1:]] for l in mono])))
    out["water1_monomer_def2svp"] = {"energy": e, "iterations": it}
    e, it = energy(ri.Molecule.water_cluster(wl.water_cluster(1, 16)))
    out["water16_seed1_def2svp"] = {"energy": e, "iterations": it}
    (ROOT / "tests/golden/scf_energies.json").write_text(json.dumps(out, indent=1))
    print(json.dumps(out, indent=1))


if __name__ == "__main__":
    main()

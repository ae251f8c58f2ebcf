"""Writes the benchmark geometries as .xyz fixtures (Ångström): the single water
monomer (O at the origin, H at (±0.7572, 0.5865, 0)), (H2O)16 seed 1 and
(H2O)64 seed 2 from rest_amd.workload.water_cluster.

    python tests/golden/make_xyz.py
"""
from pathlib import Path
import sys

import numpy as np

HERE = Path(__file__).resolve().parent
sys.path.insert(0, str(HERE.parents[1]))
from rest_amd import workload as wl  # noqa: E402


def write(path: Path, xyz: np.ndarray, title: str):
    n = len(xyz) // 3
    sym = ["O", "H", "H"] * n
    lines = [str(len(sym)), title] + [f"{s} {x:.12f} {y:.12f} {z:.12f}" for s, (x, y, z) in zip(sym, xyz)]
    path.write_text("\n".join(lines) + "\n")


if __name__ == "__main__":
    mono = np.array([[0.0, 0.0, 0.0], [0.7572, 0.5865, 0.0], [-0.7572, 0.5865, 0.0]])
    write(HERE / "h2o_monomer.xyz", mono, "H2O monomer: O at the origin, H at (+-0.7572, 0.5865, 0)")
    write(HERE / "h2o16_seed1.xyz", wl.water_cluster(1, 16), "(H2O)16, workload.water_cluster(1, 16)")
    write(HERE / "h2o64_seed2.xyz", wl.water_cluster(2, 64), "(H2O)64, workload.water_cluster(2, 64)")

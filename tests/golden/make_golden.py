"""Write the golden J/K fixtures (tests/golden/*.npz).

The expected values are computed here by brute force on unpacked, full
matrices in numpy long double — an algorithm independent of the C oracle's
packed "2·full − diag" contractions and per-P compensated sums — so the
fixtures pin the oracle, and through it the HIP library.

The reference (Rust, with un-vendored rest_tensors/rest_libcint) cannot run in
this image, so no value here comes from it; the formulas follow
src/scf_io/mod.rs:2428-2834 and are pinned independently by the closed-form
tests in tests/test_oracle.py.

Shape "h2o": nbf=24, naux=113 (one water in def2-SVP / def2-SVP-JKFIT),
nocc=5; synthetic B from the seeded hash generator (re-implemented below in
numpy, which also pins the C and HIP generators).

Run: python tests/golden/make_golden.py
"""
from __future__ import annotations

from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LD = np.longdouble
M64 = np.uint64(0xFFFFFFFFFFFFFFFF)


def mix64(z: np.ndarray) -> np.ndarray:
    z = z.astype(np.uint64)
    with np.errstate(over="ignore"):
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    return z


def synth_rimatr(seed: int, nbf: int, naux: int) -> np.ndarray:
    """Packed B [p, P] of the hash generator: (2u−1)·sqrt(3/naux)."""
    npair = nbf * (nbf + 1) // 2
    flat = np.arange(naux * npair, dtype=np.uint64)
    with np.errstate(over="ignore"):
        z = mix64(np.uint64(seed) + (flat + np.uint64(1)) * np.uint64(0x9E3779B97F4A7C15))
    u = (z >> np.uint64(11)).astype(np.float64) * 2.0 ** -53
    b = (2.0 * u - 1.0) * np.sqrt(3.0 / naux)
    return b.reshape(naux, npair).T.copy()


def upper_idx(nbf):
    j = np.concatenate([np.full(k + 1, k) for k in range(nbf)])
    i = np.concatenate([np.arange(k + 1) for k in range(nbf)])
    return i, j


def unpack_cols(b: np.ndarray, nbf: int) -> np.ndarray:
    """[p, P] -> [i, j, P], symmetric in (i, j)."""
    i, j = upper_idx(nbf)
    out = np.zeros((nbf, nbf, b.shape[1]), dtype=b.dtype)
    out[i, j, :] = b
    out[j, i, :] = b
    return out


def pack(m: np.ndarray) -> np.ndarray:
    i, j = upper_idx(m.shape[0])
    return m[i, j]


def vj_ref(bfull: np.ndarray, d_eff: np.ndarray) -> np.ndarray:
    """J = Σ_P B_P (Σ_kl B_kl,P d_eff_kl); returned full."""
    g = np.einsum("ijP,ij->P", bfull, d_eff)
    return np.einsum("ijP,P->ij", bfull, g), g


def vk_ref(bfull: np.ndarray, d: np.ndarray, transpose_right: bool) -> np.ndarray:
    """Σ_P B_P D B_P (or B_P D B_Pᵀ), full."""
    n, _, naux = bfull.shape
    k = np.zeros((n, n), dtype=bfull.dtype)
    for P in range(naux):
        bp = bfull[:, :, P]
        k += bp @ d @ (bp.T if transpose_right else bp)
    return k


def orthonormal(rng, nbf, nocc):
    q, r = np.linalg.qr(rng.standard_normal((nbf, nocc)))
    return q * np.sign(np.diag(r))


def main():
    nbf, naux, nocc = 24, 113, 5
    b = synth_rimatr(11, nbf, naux)
    rng = np.random.default_rng(20240607)
    c = orthonormal(rng, nbf, nocc)
    occ = np.full(nocc, 2.0)
    d = (c * occ) @ c.T
    d_asym = d + 0.01 * rng.standard_normal((nbf, nbf))
    ca = orthonormal(rng, nbf, nocc)
    cb = orthonormal(rng, nbf, nocc - 1)

    bl = unpack_cols(b.astype(LD), nbf)
    dl = d.astype(LD)
    dupper = np.triu(d_asym).astype(LD)
    dupper = dupper + np.triu(dupper, 1).T          # upper triangle of D, mirrored
    out = {}
    jfull, g = vj_ref(bl, np.triu(dl) + np.triu(dl, 1).T)
    out["gamma"] = g.astype(np.float64)
    out["vj"] = pack(jfull).astype(np.float64)
    j_up, _ = vj_ref(bl, dupper)
    out["vj_asym_rimatr"] = pack(j_up).astype(np.float64)
    j_full, _ = vj_ref(bl, d_asym.astype(LD))
    out["vj_asym_ri3fn"] = pack(j_full).astype(np.float64)
    out["vk_rhf"] = (-0.5 * pack(vk_ref(bl, dl, False))).astype(np.float64)
    out["vk_dm_asym"] = pack(vk_ref(bl, d_asym.astype(LD), False)).astype(np.float64)
    da = (ca @ ca.T).astype(LD)
    db = (cb @ cb.T).astype(LD)
    out["vk_uhf_a"] = (-1.0 * pack(vk_ref(bl, da, False))).astype(np.float64)
    out["vk_uhf_b"] = (-1.0 * pack(vk_ref(bl, db, False))).astype(np.float64)

    # a non-symmetric full B pins the N/T of the occupied route on ri3fn
    nb2, na2 = 12, 17
    b3n = rng.standard_normal((nb2, nb2, na2)) * 0.2
    c2 = orthonormal(rng, nb2, 3)
    occ2 = np.array([2.0, 2.0, 1.5])
    k2 = vk_ref(b3n.astype(LD), ((c2 * occ2) @ c2.T).astype(LD), True)
    out["nonsym_ri3fn"] = b3n
    out["nonsym_c"] = c2
    out["nonsym_occ"] = occ2
    out["nonsym_vk"] = pack(k2).astype(np.float64)

    # a handful of raw generator values (flat index -> value), seed 11, h2o shape
    idx = np.array([0, 1, 299, 300, 12345, naux * (nbf * (nbf + 1) // 2) - 1])
    out["synth_idx"] = idx
    out["synth_val"] = b.T.reshape(-1)[idx]

    np.savez_compressed(HERE / "jk_h2o_shape.npz", rimatr=b, c=c, occ=occ, dm=d, dm_asym=d_asym, ca=ca, cb=cb,
                        nbf=nbf, naux=naux, **out)
    print("wrote", HERE / "jk_h2o_shape.npz")


if __name__ == "__main__":
    main()

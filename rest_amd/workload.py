"""Seeded synthetic workloads of the RI-V J/K path (water clusters, orbitals).

PRNG: splitmix64-seeded xoshiro256**; uniforms are (x >> 11)·2^-53; normals
are Box–Muller cos branch from two uniforms.  The C oracle implements the same
generators independently (tests/test_workload.py checks they agree).

Water clusters: n waters on the first n sites (x fastest) of a k×k×k cubic
lattice (k^3 >= n) with 3.1 Å spacing, O jittered by ±0.2 Å per axis, each
monomer (O at the origin, H at (±0.7572, 0.5865, 0) Å) rotated about O by a
uniform unit quaternion (Shoemake); the cluster is redrawn if an O–O distance
is below 2.5 Å.  Shapes in def2-SVP / def2-SVP-JKFIT: nbf = 24n, naux = 113n,
nocc = 5n.
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np

_M = (1 << 64) - 1


def _splitmix64(state: int):
    state = (state + 0x9E3779B97F4A7C15) & _M
    z = state
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & _M
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & _M
    return state, z ^ (z >> 31)


class Xoshiro256:
    def __init__(self, seed: int):
        x = seed & _M
        s = []
        for _ in range(4):
            x, z = _splitmix64(x)
            s.append(z)
        self.s = s

    def next(self) -> int:
        s0, s1, s2, s3 = self.s
        r = (((s1 * 5) & _M) << 7 | ((s1 * 5) & _M) >> 57) & _M
        r = (r * 9) & _M
        t = (s1 << 17) & _M
        s2 ^= s0
        s3 ^= s1
        s1 ^= s2
        s0 ^= s3
        s2 ^= t
        s3 = ((s3 << 45) | (s3 >> 19)) & _M
        self.s = [s0, s1, s2, s3]
        return r

    def uniform(self) -> float:
        return (self.next() >> 11) * 2.0 ** -53

    def uniforms(self, n: int) -> np.ndarray:
        return np.array([self.uniform() for _ in range(n)])

    def normals(self, n: int) -> np.ndarray:
        u = self.uniforms(2 * n).reshape(n, 2)
        return np.sqrt(-2.0 * np.log(1.0 - u[:, 0])) * np.cos(2.0 * math.pi * u[:, 1])


def water_cluster(seed: int, n: int, oh_jitter: float = 0.0, hoh_jitter_deg: float = 0.0) -> np.ndarray:
    """(3n, 3) coordinates in Å, atoms O, H, H per water."""
    r = Xoshiro256(seed)
    k = 1
    while k ** 3 < n:
        k += 1
    for _attempt in range(1, 1000):
        xyz = np.empty((3 * n, 3))
        for w in range(n):
            site = np.array([w % k, (w // k) % k, w // (k * k)], dtype=float) * 3.1
            o = site + (2.0 * np.array([r.uniform(), r.uniform(), r.uniform()]) - 1.0) * 0.2
            u1, u2, u3 = r.uniform(), r.uniform(), r.uniform()
            qx = math.sqrt(1 - u1) * math.sin(2 * math.pi * u2)
            qy = math.sqrt(1 - u1) * math.cos(2 * math.pi * u2)
            qz = math.sqrt(u1) * math.sin(2 * math.pi * u3)
            qw = math.sqrt(u1) * math.cos(2 * math.pi * u3)
            R = np.array([
                [1 - 2 * (qy * qy + qz * qz), 2 * (qx * qy - qz * qw), 2 * (qx * qz + qy * qw)],
                [2 * (qx * qy + qz * qw), 1 - 2 * (qx * qx + qz * qz), 2 * (qy * qz - qx * qw)],
                [2 * (qx * qz - qy * qw), 2 * (qy * qz + qx * qw), 1 - 2 * (qx * qx + qy * qy)]])
            h1 = np.array([0.7572, 0.5865, 0.0])
            h2 = np.array([-0.7572, 0.5865, 0.0])
            if oh_jitter > 0 or hoh_jitter_deg > 0:
                r0 = math.hypot(0.7572, 0.5865)
                half = math.atan2(0.7572, 0.5865)
                r1 = r0 + (2 * r.uniform() - 1) * oh_jitter
                r2 = r0 + (2 * r.uniform() - 1) * oh_jitter
                half += 0.5 * (2 * r.uniform() - 1) * hoh_jitter_deg * math.pi / 180.0
                h1 = np.array([r1 * math.sin(half), r1 * math.cos(half), 0.0])
                h2 = np.array([-r2 * math.sin(half), r2 * math.cos(half), 0.0])
            xyz[3 * w] = o
            xyz[3 * w + 1] = o + R @ h1
            xyz[3 * w + 2] = o + R @ h2
        ox = xyz[0::3]
        d2 = ((ox[:, None, :] - ox[None, :, :]) ** 2).sum(-1)
        iu = np.triu_indices(n, 1)
        if n < 2 or d2[iu].min() >= 2.5 ** 2:
            return xyz
    raise RuntimeError("water cluster generator exhausted its attempts")


def batched_cluster_sizes(count: int, base: int = 0x5EED, lo: int = 2, hi: int = 32) -> list:
    """Water counts of the batched workload: cluster i has n = lo + floor(u (hi - lo + 1)), u the top
    53 bits of splitmix64(base ^ i) as a fraction (n uniform in [lo, hi])."""
    out = []
    for i in range(count):
        _, z = _splitmix64((base ^ i) & ((1 << 64) - 1))
        out.append(lo + (z >> 11) * (hi - lo + 1) // (1 << 53))
    return out


def random_orthonormal(seed: int, nbf: int, nocc: int) -> np.ndarray:
    """Seeded Gaussian nbf×nocc (column-major draw order), orthonormalised by
    Householder QR with diag(R) >= 0."""
    g = Xoshiro256(seed).normals(nbf * nocc).reshape(nocc, nbf).T
    q, rr = np.linalg.qr(g)
    sgn = np.where(np.diag(rr) < 0, -1.0, 1.0)
    return q * sgn


@dataclass(frozen=True)
class Shape:
    name: str
    nwater: int

    @property
    def nbf(self) -> int:
        return 24 * self.nwater

    @property
    def naux(self) -> int:
        return 113 * self.nwater

    @property
    def nocc(self) -> int:
        return 5 * self.nwater

    @property
    def npair(self) -> int:
        return self.nbf * (self.nbf + 1) // 2


H2O = Shape("H2O def2-SVP/def2-SVP-JKFIT", 1)
H2O_16 = Shape("(H2O)16 def2-SVP/def2-SVP-JKFIT", 16)
H2O_64 = Shape("(H2O)64 def2-SVP/def2-SVP-JKFIT", 64)


def jk_bytes(shape: Shape, s_b: int, nspin: int = 1) -> int:
    """Algorithmic HBM bytes of one J build: two passes over B plus D, J, γ."""
    return 2 * shape.npair * shape.naux * s_b + 8 * nspin * (shape.nbf ** 2 + shape.npair) + 16 * shape.naux


def k_flops(shape: Shape, nocc: int | None = None) -> dict:
    """Algorithmic FLOP of one K build (occupied route)."""
    o = shape.nocc if nocc is None else nocc
    half = 2 * shape.nbf ** 2 * o * shape.naux
    syrk = shape.nbf * (shape.nbf + 1) * o * shape.naux
    return {"half": half, "syrk": syrk, "total": half + syrk}

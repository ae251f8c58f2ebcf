"""CPU restatement of the split-BF16 exchange arithmetic (float64 numpy).

TEST INFRASTRUCTURE ONLY: imported by tests/, never by the rest_amd package.

The device K at RJK_PREC_BF16X2 / RJK_PREC_BF16X3 evaluates REST's occupied
route (vk_upper_with_ri_v_sync, src/scf_io/mod.rs:2758-2834)
    K = scale · Σ_P M_P diag(sign) M_Pᵀ,   M_P = B_P C',  C' = C·sqrt|occ|
with every operand split into S BF16 planes, x ≈ x_0 + … + x_{S-1},
x_k = bf16(x − x_0 − … − x_{k-1}) (round to nearest even after rounding the
remainder to float32), and each product restricted to the plane pairs (a, b)
with a + b < S.  M_P is formed in FP32 and split again before the second
product.  This module reproduces those roundings and plane selections exactly
and sums the selected products in float64, so a device result differs from it
only by FP32 accumulation, which is far below the dropped plane terms at S = 2.

Pins (tests/test_split_oracle.py): planes sum back to x within 2^-8S |x|;
values with at most 8 significant bits split into one exact plane; at S = 2, 3
the emulated K approaches the float64 K at the stated rates and S = 4 is closer
than S = 3.
"""
from __future__ import annotations

import numpy as np


def bf16(x) -> np.ndarray:
    """float64 → float32 (RNE) → BF16 (RNE), returned as float64."""
    u = np.ascontiguousarray(np.asarray(x, dtype=np.float64).astype(np.float32)).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def split_planes(x, nplanes: int) -> list:
    """[x_0, …, x_{S-1}] with x_k = bf16(x − Σ_{j<k} x_j)."""
    r = np.asarray(x, dtype=np.float64)
    out = []
    for _ in range(nplanes):
        h = bf16(r)
        out.append(h)
        r = r - h
    return out


def pairs(nplanes: int) -> list:
    """Plane pairs (a, b) with a + b < S, in the device's product order."""
    return [(a, s - a) for s in range(nplanes) for a in range(s + 1)]


def split_product(a_planes: list, b_planes: list, nplanes: int, fn) -> np.ndarray:
    """Σ_{a+b<S} fn(A_a, B_b) in float64."""
    return sum(fn(a_planes[i], b_planes[j]) for i, j in pairs(nplanes))


def unpack_sym(col: np.ndarray, nbf: int) -> np.ndarray:
    """Full symmetric matrix of one packed-upper B column (p = j(j+1)/2 + i)."""
    m = np.zeros((nbf, nbf))
    iu = np.triu_indices(nbf)
    order = np.lexsort((iu[0], iu[1]))          # sort by (j, i): packed order
    ii, jj = iu[0][order], iu[1][order]
    m[ii, jj] = col
    m[jj, ii] = col
    return m


def vk_split(rimatr: np.ndarray, c: np.ndarray, occ: np.ndarray, nplanes: int, scale: float) -> np.ndarray:
    """Packed-upper K of one spin from packed B (npair, naux), orbitals c (nbf, nw) and
    occupations occ (nw), in the split-plane arithmetic described above."""
    nbf = c.shape[0]
    cw = c * np.sqrt(np.abs(occ))[None, :]
    sign = np.where(occ < 0, -1.0, 1.0)
    cpl = split_planes(cw, nplanes)
    k = np.zeros((nbf, nbf))
    for P in range(rimatr.shape[1]):
        bpl = split_planes(unpack_sym(rimatr[:, P], nbf), nplanes)
        m = split_product(bpl, cpl, nplanes, lambda x, y: x @ y)           # (nbf, nw)
        mpl = split_planes(m.astype(np.float32).astype(np.float64), nplanes)
        k += split_product(mpl, mpl, nplanes, lambda x, y: (x * sign[None, :]) @ y.T)
    iu = np.triu_indices(nbf)
    order = np.lexsort((iu[0], iu[1]))
    return scale * k[iu[0][order], iu[1][order]]

"""CPU restatement of the learned guess's pair-block MLP (test infrastructure
only; include/rest_lg.h describes the device network it checks).

Forward, masked-MSE loss, backward and Adam in float64.  With bf16=True every
GEMM operand is first rounded to BF16 (round to nearest even), as the device
network does before its matrix-core products; biases, SiLU and the loss stay
exact.  Parameter layout: per layer W [in][out] row-major, then b [out].
Pinned by tests/test_mlp_oracle.py: central finite differences of the loss
(bf16=False) and the closed-form first Adam step.
"""
from __future__ import annotations

import numpy as np


def bf16_round(x) -> np.ndarray:
    """Round float32 values to the nearest BF16 (ties to even); returned as float64."""
    u = np.asarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def unpack(params: np.ndarray, dims) -> list:
    out, off = [], 0
    for din, dout in zip(dims[:-1], dims[1:]):
        W = params[off:off + din * dout].reshape(din, dout)
        off += din * dout
        b = params[off:off + dout]
        off += dout
        out.append((W, b))
    assert off == params.size
    return out


def _silu(z):
    return z / (1.0 + np.exp(-z))


def _dsilu(z):
    s = 1.0 / (1.0 + np.exp(-z))
    return s * (1.0 + z * (1.0 - s))


def _mm(a, b, bf16):
    if bf16:
        return bf16_round(a) @ bf16_round(b)
    return np.asarray(a, dtype=np.float64) @ np.asarray(b, dtype=np.float64)


def forward(params, dims, X, bf16: bool = True, keep: bool = False):
    layers = unpack(np.asarray(params, dtype=np.float64), dims)
    A = np.asarray(X, dtype=np.float64)
    acts, pres = [A], []
    for l, (W, b) in enumerate(layers):
        Z = _mm(A, W, bf16) + b
        A = _silu(Z) if l + 1 < len(layers) else Z
        pres.append(Z)
        acts.append(A)
    return (A, acts, pres) if keep else A


def loss_grad(params, dims, X, T, M, bf16: bool = True):
    """(loss, gradient) of L = Σ M (y − T)² / Σ M."""
    layers = unpack(np.asarray(params, dtype=np.float64), dims)
    Y, acts, pres = forward(params, dims, X, bf16, keep=True)
    M = np.asarray(M, dtype=np.float64)
    msum = M.sum()
    loss = float(np.sum(M * (Y - T) ** 2) / msum)
    G = 2.0 * M * (Y - T) / msum
    grads = [None] * len(layers)
    for l in range(len(layers) - 1, -1, -1):
        W, _ = layers[l]
        dZ = G * _dsilu(pres[l]) if l + 1 < len(layers) else G
        grads[l] = (_mm(acts[l].T, dZ, bf16), dZ.sum(axis=0))
        if l > 0:
            G = _mm(dZ, W.T, bf16)
    flat = np.concatenate([np.concatenate([gW.ravel(), gb]) for gW, gb in grads])
    return loss, flat


def adam(params, grad, m, v, step: int, lr: float, b1: float = 0.9, b2: float = 0.999, eps: float = 1e-8):
    """One Adam step (bias-corrected); returns (params, m, v) new arrays; `step` counts from 1."""
    m = b1 * m + (1 - b1) * grad
    v = b2 * v + (1 - b2) * grad * grad
    mh = m / (1 - b1 ** step)
    vh = v / (1 - b2 ** step)
    return params - lr * mh / (np.sqrt(vh) + eps), m, v

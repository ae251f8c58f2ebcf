"""The allocation-free J entry point, rjk_vj_upper_dev_noalloc, through the C ABI.

It runs the kernels of rjk_vj_upper_dev on a fixed device array instead of the
context's workspace, so its results must equal that entry point's bit for bit
(the same kernels on the same plan).  Two calls in a row on the one GPU, the
second with a smaller input and another precision, each checked against
rjk_vj_upper_dev and against the oracle (1e-11 max-abs for FP64); then the
size limit, which is checked before any device work.
"""
import sys

import numpy as np
import pytest

from oracle import oracle as orc
from rest_amd import jk

pytestmark = pytest.mark.gpu
TOL = 1e-11


def say(msg):
    print(f"[gpu-case] {msg}", flush=True)
    sys.stdout.flush()


def test_noalloc_j_equals_vj_dev_on_two_calls():
    import torch
    dev = torch.device("cuda:0")
    # (nbf, naux, spins, seed, full-D semantics, precision); the second call is smaller
    cases = [(384, 1808, 2, 31, False, jk.PREC_FP64), (96, 300, 1, 32, True, jk.PREC_FP32)]
    for nbf, naux, ns, seed, full, prec in cases:
        say(f"no-alloc J nbf={nbf} naux={naux} spins={ns} full_dm={full} prec={prec}")
        rng = np.random.default_rng(seed)
        npair = nbf * (nbf + 1) // 2
        dms = [rng.standard_normal((nbf, nbf)) for _ in range(ns)]
        b = orc.synth_rimatr(seed, nbf, naux)
        dm_dev = torch.from_numpy(np.ascontiguousarray(np.stack([d.T for d in dms]))).to(dev)   # column-major
        with jk.DeviceRI(nbf, naux) as ri:
            ri.fill_synthetic(seed, jk.STORE_FP64 | jk.STORE_FP32)
            want = torch.empty((ns, npair), dtype=torch.float64, device=dev)
            got = torch.full((ns, npair), float("nan"), dtype=torch.float64, device=dev)
            torch.cuda.synchronize()
            ri.vj_dev(dm_dev.data_ptr(), ns, 0.5, want.data_ptr(), prec=prec, full_dm=full)
            ri.vj_dev_noalloc(dm_dev.data_ptr(), ns, 0.5, got.data_ptr(), prec=prec, full_dm=full)
            ri.synchronize()
            w, g = want.cpu().numpy(), got.cpu().numpy()
        assert np.array_equal(g, w)
        if full:
            ref = np.stack(orc.vj_upper_ri3fn(orc.ri3fn_from_rimatr(b, nbf), dms, ns, 0.5)[:ns])
        else:
            ref = np.stack(orc.vj_upper_rimatr(b, dms, ns, 0.5)[:ns])
        err = float(np.max(np.abs(g - ref)))
        say(f"  max-abs vs oracle {err:.2e}")
        if prec == jk.PREC_FP64:
            assert err <= TOL
        else:
            assert err / np.abs(ref).max() < 5e-6


def test_noalloc_j_rejects_shapes_beyond_its_scratch():
    nbf, naux = 4096, 16
    say(f"no-alloc J size limit nbf={nbf} naux={naux}")
    with jk.DeviceRI(nbf, naux) as ri:
        with pytest.raises(jk.RJKError) as e:
            ri.vj_dev_noalloc(1, 2, 1.0, 1)    # rejected before any pointer is used
        assert e.value.code == 1
        assert "exceeds" in str(e.value)

"""GPU tests of the device three-centre integral kernels (rsi_int3c2e_packed_dev)
against the host McMurchie–Davidson engine, which the CPU suite pins to closed
forms, finite differences, rotation invariance and REST's pentacene energy.
Bar: max |device − host| ≤ 1e-12 · max(1, |host|) element-wise (same
algorithm, different summation order and Boys evaluation only)."""
import sys

import numpy as np
import pytest

from rest_amd import jk, ri
from rest_amd import workload as wl

pytestmark = pytest.mark.gpu


def say(msg):
    print(f"[gpu-case] {msg}", flush=True)
    sys.stdout.flush()


def compare(m, j0=0, j1=None):
    host = m.int3c2e_packed(j0, j1).T
    with jk.DeviceRI(m.nbf, m.naux) as r:
        dev = ri.int3c2e_packed_dev(r, m.orb, m.aux, j0, j1).cpu().numpy()
    err = np.abs(dev - host) / np.maximum(1.0, np.abs(host))
    say(f"  shape {host.shape} max|host| {np.abs(host).max():.3e} max rel err {err.max():.2e}")
    assert np.all(np.isfinite(dev))
    assert err.max() <= 1e-12
    return host


def test_water_cluster_all_classes_up_to_d_and_g():
    say("(H2O)2 def2-SVP / def2-SVP-JKFIT: orbital s,p,d x auxiliary s..g")
    m = ri.Molecule.water_cluster(wl.water_cluster(3, 2))
    assert m.orb.shell_l.max() == 2 and m.aux.shell_l.max() == 4
    compare(m)


def test_slab_on_shell_boundaries():
    m = ri.Molecule.water_cluster(wl.water_cluster(5, 2))
    cuts = m.orb.shell_boundaries()
    j0, j1 = int(cuts[3]), int(cuts[9])
    say(f"(H2O)2 slab [{j0}, {j1})")
    compare(m, j0, j1)


def test_concentric_and_distant_pairs():
    """One-centre pairs (the Gaussian product is exact) and a pair far enough apart that some
    primitive pairs fall under the screening cut-off."""
    say("H2 at 0.74 and 9 Angstrom, def2-SVP / def2-SVP-JKFIT")
    for d in (0.74, 9.0):
        m = ri.Molecule.from_angstrom(["H", "H"], np.array([[0, 0, 0], [0.1, 0.2, d]]))
        compare(m)


def test_rejects_f_orbital_shells_and_bad_ranges():
    xyz = np.array([[0.0, 0.0, 0.0]])
    orb = ri.Basis(["X"], xyz, "custom", [(0, 3, [0.8], [1.0])])
    aux = ri.Basis(["X"], xyz, "custom", [(0, 0, [1.0], [1.0])])
    m = ri.Molecule.water_cluster(wl.water_cluster(3, 1))
    with jk.DeviceRI(m.nbf, m.naux) as r:
        for args in ((orb, aux), (m.orb, m.aux, 4, m.nbf)):   # f shell; j0 inside the first p shell of O
            with pytest.raises(jk.RJKError) as e:
                ri.int3c2e_packed_dev(r, *args)
            assert e.value.code == 1

"""CrystalBatch (uneven random devices) against the reference's golden R/T and
against one-at-a-time Crystal solves."""
import numpy as np
import pytest

from conftest import golden
from khepri_amd import workloads as W

pytestmark = pytest.mark.gpu
TOL = 1e-9


def build(dv):
    from khepri_amd import Crystal
    cl = Crystal(dv["pw"])
    names = []
    for j, ly in enumerate(dv["layers"]):
        cl.add_layer_pixmap(f"L{j}", W.render_layer(ly), ly["depth"])
        names.append(f"L{j}")
    cl.set_device(names, False)
    cl.set_source(1 / dv["freq"], 1.0, 1.0, dv["theta"], dv["phi"])
    return cl


def test_batch_golden_devices(gpu):
    from khepri_amd.batch import CrystalBatch
    g = golden("random_devices.npz")
    devs = W.random_device_params()
    R, T = CrystalBatch([build(devs[i]) for i in g["index"]]).solve()
    assert np.abs(R - g["R"]).max() <= TOL and np.abs(T - g["T"]).max() <= TOL


def test_batch_golden_edge_devices(gpu):
    """Seeded devices at pw (11, 11) and (13, 13) and stacks of 6-8 layers
    (random_devices_edges.npz, the reference's golden R/T)."""
    from khepri_amd.batch import CrystalBatch
    g = golden("random_devices_edges.npz")
    devs = W.random_device_params()
    chosen = [devs[i] for i in g["index"]]
    assert any(dv["pw"] == (13, 13) for dv in chosen) and any(dv["pw"] == (11, 11) for dv in chosen)
    assert sum(len(dv["layers"]) >= 6 for dv in chosen) >= 3
    R, T = CrystalBatch([build(dv) for dv in chosen]).solve()
    assert np.abs(R - g["R"]).max() <= TOL and np.abs(T - g["T"]).max() <= TOL, \
        (np.abs(R - g["R"]).max(), np.abs(T - g["T"]).max())


def test_batch_matches_single_solves(gpu):
    """A mix of harmonic counts (5..13)^2, 1-8 layers, a uniform layer and
    non-vacuum media, batched vs solved one by one."""
    from khepri_amd import Crystal
    from khepri_amd.batch import CrystalBatch
    devs = W.random_device_params(count=40, seed=77)
    cls = [build(dv) for dv in devs if dv["pw"][0] <= 11][:14]
    extra = Crystal((5, 5), epsi=2.25, epse=1.44)
    extra.add_layer_pixmap("P", W.disc_pixmap(64, 6.0, 0.3, 1.0), 0.3)
    extra.add_layer_uniform("U", 3.0 - 0.1j, 0.2)
    extra.set_device(["P", "U", "P"], False)
    extra.set_source(1 / 0.7, 1.0, 0.3, 15.0, 40.0)
    cls.append(extra)
    Rb, Tb = CrystalBatch(cls).solve()
    for k, cl in enumerate(cls):
        cl.solve(rt_only=False)
        r, t = cl.poynting_flux_end()
        assert abs(Rb[k] - r) < 1e-10 and abs(Tb[k] - t) < 1e-10, k

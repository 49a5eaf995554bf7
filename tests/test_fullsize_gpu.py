"""Full-size runs of the disc_stack_bz, random-device, grad_grid and twisted
workloads, checked through properties
that hold at any size (the oracle is too slow at these sizes; the golden
subsets are compared elsewhere).

* disc_stack_bz (65 536 solves, pw (9, 9)): the disc stack is lossless (eps 12 and 1),
  so R + T = 1 within 1e-9 wherever the incident wave propagates, R and T lie
  in [0, 1], and R, T are non-finite exactly where |k_par| > k0 (evanescent
  incidence, khepri/alternative.py:241).
* twisted bilayer (8 angles x 32 frequencies, 625 extended harmonics): lossless, normal
  incidence: R + T = 1 within 1e-9.
* random devices (4 096 devices, pw (5..13)^2, 1-8 layers): one CrystalBatch over
  the whole workload; the devices with golden R/T (both golden files) are
  checked at their positions in the batch, the rest are finite.
* grad_grid (pw (11, 11), 4 x 64 x 64 real pixels, 16 frequencies): R + T = 1 for
  every real permittivity, so the adjoint gradient of R + T with respect to
  every pixel vanishes: |d(R + T)/d eps| <= 1e-7 max |dT/d eps|.
"""
import numpy as np
import pytest

from khepri_amd import workloads as W

pytestmark = pytest.mark.gpu


def test_disc_stack_bz_full_energy_and_evanescent_pattern(gpu):
    from khepri_amd import Crystal
    cfg = W.disc_stack_bz()
    cl = Crystal(cfg["pw"])
    cl.add_layer_uniform("S1", cfg["eps_gap"], cfg["depth_gap"])
    cl.add_layer_pixmap("Scyl", cfg["pixmap"], cfg["depth_disc"])
    cl.set_device(["Scyl", "S1", "Scyl"], False)
    cl.set_source(1 / cfg["freqs"], 1.0, 0.0, kp=cfg["kp"])
    cl.solve()
    R, T = cl.poynting_flux_end()
    assert R.shape == (65536,)
    k0 = 2 * np.pi * cfg["freqs"]
    evan = (cfg["kp"] ** 2).sum(1) > k0 ** 2
    fin = np.isfinite(R) & np.isfinite(T)
    assert np.array_equal(~fin, evan), (int((~fin).sum()), int(evan.sum()))
    assert 0.1 < evan.mean() < 0.2            # 13.9 % of the workload is evanescent
    assert np.abs(R[fin] + T[fin] - 1).max() <= 1e-9
    assert R[fin].min() >= -1e-12 and T[fin].min() >= -1e-12


def test_random_devices_full_batch_golden_positions(gpu):
    from conftest import golden
    from khepri_amd import Crystal
    from khepri_amd.batch import CrystalBatch
    devs = W.random_device_params()
    cls = []
    for dv in devs:
        c = Crystal(dv["pw"])
        names = []
        for j, ly in enumerate(dv["layers"]):
            c.add_layer_pixmap(f"L{j}", W.render_layer(ly), ly["depth"])
            names.append(f"L{j}")
        c.set_device(names, False)
        c.set_source(1 / dv["freq"], 1.0, 1.0, dv["theta"], dv["phi"])
        cls.append(c)
    R, T = CrystalBatch(cls).solve()
    assert R.shape == (4096,) and np.isfinite(R).all() and np.isfinite(T).all()
    for name in ("random_devices.npz", "random_devices_edges.npz"):
        g = golden(name)
        idx = g["index"]
        assert np.abs(R[idx] - g["R"]).max() <= 1e-9 and np.abs(T[idx] - g["T"]).max() <= 1e-9, name


def test_twisted_full_energy(gpu):
    from khepri_amd import Crystal, Expansion, Layer
    cfg = W.twisted_bilayer()
    for ang in cfg["angles_deg"]:
        e1, e2 = Expansion(cfg["pw"]), Expansion(cfg["pw"])
        e2.rotate(np.deg2rad(ang))
        etw = e1 + e2
        cl = Crystal.from_expansion(etw)
        cl.add_layer("L1", Layer.pixmap(e1, cfg["pixmap"], cfg["depth"]), extended=True)
        cl.add_layer("gap", Layer.uniform(etw, 1.0, cfg["gap"]))
        cl.add_layer("L2", Layer.pixmap(e2, cfg["pixmap"], cfg["depth"]), extended=True)
        cl.set_device(["L1", "gap", "L2"], False)
        cl.set_source(1 / cfg["freqs"], 1.0, 0.0)
        cl.solve(rt_only=True)
        R, T = cl.poynting_flux_end()
        assert R.shape == (32,)
        assert np.abs(R + T - 1).max() <= 1e-9, ang


def test_grad_grid_full_gradient_of_total_flux_vanishes(gpu):
    from khepri_amd import Crystal
    from khepri_amd.adjoint import flux_gradient
    cfg = W.grad_grid()
    cl = Crystal(cfg["pw"])
    names = []
    for j, pix in enumerate(cfg["pixmaps"]):
        cl.add_layer_pixmap(f"L{j}", pix, cfg["depth"])
        names.append(f"L{j}")
    cl.set_device(names, False)
    cl.set_source(1 / cfg["freqs"], 1.0, 0.0)
    R, T, gT = flux_gradient(cl, wT=1.0, wR=0.0)
    assert np.abs(R + T - 1).max() <= 1e-9
    _, _, gRT = flux_gradient(cl, wT=1.0, wR=1.0)
    scale = max(float(gT[nm].abs().max()) for nm in names)
    worst = max(float(gRT[nm].abs().max()) for nm in names)
    assert gT[names[0]].shape == (16, 64, 64)
    assert scale > 0 and worst <= 1e-7 * scale, (worst, scale)

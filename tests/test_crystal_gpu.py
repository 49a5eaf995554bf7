"""End-to-end parity of khepri_amd.Crystal on the GPU against the reference's
golden vectors (tests/golden) and its own known-answer test.

Tolerance: max |dR|, |dT| <= 1e-9 (FP64 path); NaN patterns of evanescent
incidence must match exactly."""
import numpy as np
import pytest

from conftest import golden
from khepri_amd import workloads as W

pytestmark = pytest.mark.gpu
TOL = 1e-9


def disc_crystal(pw, pixmap, fields=False, epsi=1, epse=1):
    from khepri_amd import Crystal
    cl = Crystal(pw, epsi=epsi, epse=epse)
    cl.add_layer_uniform("S1", 1, 1.1)
    cl.add_layer_pixmap("Scyl", pixmap, 0.55)
    cl.set_device(["Scyl", "S1", "Scyl"], [fields] * 3)
    return cl


def test_disc_sweep_batched(gpu):
    g = golden("disc_sweep.npz")
    cfg = W.disc_stack_sweep()
    cl = disc_crystal((5, 5), cfg["pixmap"])
    cl.set_source(1 / cfg["freqs"], 1.0, 0.0, 0.0, 0.0)
    cl.solve()
    R, T = cl.poynting_flux_end()
    assert R.shape == (151,)
    assert np.abs(R - g["R"]).max() <= TOL and np.abs(T - g["T"]).max() <= TOL


def test_disc_sweep_reference_protocol(gpu):
    """One scalar set_source / solve / poynting_flux_end at a time, as the
    reference README loop does."""
    g = golden("disc_sweep.npz")
    cfg = W.disc_stack_sweep()
    cl = disc_crystal((5, 5), cfg["pixmap"], fields=True)
    for i in [0, 40, 150]:
        cl.set_source(1 / cfg["freqs"][i], 1.0, 0, 0.0, 0.0)
        cl.solve()
        R, T = cl.poynting_flux_end()
        assert isinstance(R, (float, np.floating))
        assert abs(R - g["R"][i]) <= TOL and abs(T - g["T"][i]) <= TOL
    cl.set_source(2.0, 1.0, 0, 0.0, 0.0)
    cl.solve()
    np.testing.assert_allclose(cl.layers["Scyl"].S.cpu().numpy()[0], g["S_layer"], atol=1e-10)
    np.testing.assert_allclose(cl.Stot.cpu().numpy()[0], g["Stot"], atol=1e-10)


@pytest.mark.parametrize("mixed", ["0", "1"])
def test_disc_bz_batched_with_nan_pattern(gpu, monkeypatch, mixed):
    """The disc_stack_bz golden set; mixed = "1" runs the layer solves through the
    mixed-precision zgesv (FP32-MFMA triangular-solve products, one FP64
    residual refinement; KHX_LU_MIXED=1), held to the same 1e-9."""
    monkeypatch.setenv("KHX_LU_MIXED", mixed)
    g = golden("disc_bz.npz")
    cfg = W.disc_stack_bz()
    idx = g["index"]
    cl = disc_crystal((9, 9), cfg["pixmap"])
    cl.set_source(1 / cfg["freqs"][idx], 1.0, 0.0, kp=cfg["kp"][idx])
    cl.solve()
    R, T = cl.poynting_flux_end()
    np.testing.assert_array_equal(np.isnan(R), np.isnan(g["R"]))
    np.testing.assert_array_equal(np.isnan(T), np.isnan(g["T"]))
    np.testing.assert_array_equal(np.isinf(T), np.isinf(g["T"]))
    fin = np.isfinite(g["T"])
    assert fin.sum() >= 20
    assert np.abs(R[fin] - g["R"][fin]).max() <= TOL and np.abs(T[fin] - g["T"][fin]).max() <= TOL


def test_fresnel_thin_film(gpu):
    """Reference known-answer test test/integration/test_complex_eps.py:13-42,
    run batched over its 50 wavelengths."""
    from khepri_amd import Crystal
    c, e0, sigma, h = 299792458, 8.85418782e-12, 0.01e6, 1.2
    covera = c / 1e-6
    wls = np.linspace(0.7, 2.0)
    R = []
    for wl in wls:   # eps depends on the wavelength: one crystal per point, like the reference test
        omega = np.pi * 2 * covera / wl
        eps = 1.6 ** 2 - 1j * sigma / omega / e0
        cl = Crystal((1, 1))
        cl.add_layer_uniform("1", eps, h)
        cl.set_device(["1"])
        cl.set_source(wl, 1, 1)
        cl.solve()
        R.append(cl.poynting_flux_end()[0])
    omega = np.pi * 2 * covera / wls
    eps = 1.6 ** 2 + 1j * sigma / omega / e0
    n2 = np.conj(np.sqrt(eps))
    r12 = (1 - n2) / (1 + n2)
    r23 = (n2 - 1) / (n2 + 1)
    ph = np.exp(-2j * 2 * np.pi / wls * n2 * h)
    rf = np.abs((r12 + r23 * ph) / (1 + r12 * r23 * ph)) ** 2
    np.testing.assert_allclose(np.array(R), rf, rtol=1e-7)


def test_random_devices(gpu):
    from khepri_amd import Crystal
    g = golden("random_devices.npz")
    devs = W.random_device_params()
    for k, i in enumerate(g["index"]):
        dv = devs[i]
        cl = Crystal(dv["pw"])
        names = []
        for j, ly in enumerate(dv["layers"]):
            cl.add_layer_pixmap(f"L{j}", W.render_layer(ly), ly["depth"])
            names.append(f"L{j}")
        cl.set_device(names, False)
        cl.set_source(1 / dv["freq"], 1.0, 1.0, dv["theta"], dv["phi"])
        cl.solve()
        r, t = cl.poynting_flux_end()
        assert abs(r - g["R"][k]) <= TOL and abs(t - g["T"][k]) <= TOL, (i, r, t)
        if i == int(g["orders_index"]):
            (_, ro), (_, to) = cl.poynting_flux_end(only_total=False)
            np.testing.assert_allclose(ro, g["R_orders"], atol=1e-10)
            np.testing.assert_allclose(to, g["T_orders"], atol=1e-10)


@pytest.mark.parametrize("fuse", ["1", "0"])
def test_grad_grid_forward(gpu, monkeypatch, fuse):
    """Four distinct patterned layers: solved in one fused call (default for
    small batches) and one by one, both against the golden vectors."""
    from khepri_amd import Crystal
    monkeypatch.setenv("KHX_FUSE_LAYERS", fuse)
    g = golden("grad_grid.npz")
    cfg = W.grad_grid()
    cl = Crystal((11, 11))
    names = []
    for j, pix in enumerate(cfg["pixmaps"]):
        cl.add_layer_pixmap(f"G{j}", pix, cfg["depth"])
        names.append(f"G{j}")
    cl.set_device(names, False)
    cl.set_source(1 / cfg["freqs"][g["findex"]], 1.0, 0.0)
    cl.solve()
    R, T = cl.poynting_flux_end()
    assert np.abs(R - g["R"]).max() <= TOL and np.abs(T - g["T"]).max() <= TOL


def test_chunked_batch_equals_resident(gpu, monkeypatch):
    cfg = W.disc_stack_sweep()
    cl = disc_crystal((5, 5), cfg["pixmap"])
    f = cfg["freqs"][:20]
    cl.set_source(1 / f, 1.0, 0.0)
    cl.solve(rt_only=False)
    R1, T1 = cl.poynting_flux_end()
    monkeypatch.setenv("KHX_CHUNK", "7")
    cl.solve(rt_only=False)
    R2, T2 = cl.poynting_flux_end()
    np.testing.assert_array_equal(R1, R2)
    np.testing.assert_array_equal(T1, T2)
    cl.solve()   # streamed chunks default to the R/T-only chain
    R3, T3 = cl.poynting_flux_end()
    assert np.abs(R3 - R1).max() < 1e-11 and np.abs(T3 - T1).max() < 1e-11


@pytest.mark.parametrize("which", ["disc_sweep", "disc_bz", "grad_grid"])
def test_rt_only_chain_golden(gpu, which):
    from khepri_amd import Crystal
    if which == "disc_sweep":
        g = golden("disc_sweep.npz")
        cfg = W.disc_stack_sweep()
        cl = disc_crystal((5, 5), cfg["pixmap"])
        cl.set_source(1 / cfg["freqs"], 1.0, 0.0)
    elif which == "disc_bz":
        g = golden("disc_bz.npz")
        cfg = W.disc_stack_bz()
        cl = disc_crystal((9, 9), cfg["pixmap"])
        cl.set_source(1 / cfg["freqs"][g["index"]], 1.0, 0.0, kp=cfg["kp"][g["index"]])
    else:
        g = golden("grad_grid.npz")
        cfg = W.grad_grid()
        cl = Crystal((11, 11))
        names = []
        for j, pix in enumerate(cfg["pixmaps"]):
            cl.add_layer_pixmap(f"G{j}", pix, cfg["depth"])
            names.append(f"G{j}")
        cl.set_device(names, False)
        cl.set_source(1 / cfg["freqs"][g["findex"]], 1.0, 0.0)
    cl.solve(rt_only=True)
    assert cl.S is None
    R, T = cl.poynting_flux_end()
    np.testing.assert_array_equal(np.isnan(T), np.isnan(g["T"]))
    fin = np.isfinite(g["T"])
    assert np.abs(R[fin] - g["R"][fin]).max() <= TOL and np.abs(T[fin] - g["T"][fin]).max() <= TOL


def test_rt_only_random_devices_orders(gpu):
    from khepri_amd import Crystal
    g = golden("random_devices.npz")
    devs = W.random_device_params()
    i = int(g["orders_index"])
    k = list(g["index"]).index(i)
    dv = devs[i]
    cl = Crystal(dv["pw"])
    names = []
    for j, ly in enumerate(dv["layers"]):
        cl.add_layer_pixmap(f"L{j}", W.render_layer(ly), ly["depth"])
        names.append(f"L{j}")
    cl.set_device(names, False)
    cl.set_source(1 / dv["freq"], 1.0, 1.0, dv["theta"], dv["phi"])
    cl.solve(rt_only=True)
    (r, ro), (t, to) = cl.poynting_flux_end(only_total=False)
    assert abs(r - g["R"][k]) <= TOL and abs(t - g["T"][k]) <= TOL
    np.testing.assert_allclose(ro, g["R_orders"], atol=1e-10)
    np.testing.assert_allclose(to, g["T_orders"], atol=1e-10)


# ---------------------------------------------------------------- island layers
def _analytic_cases():
    import os
    import sys
    sys.path.insert(0, os.path.join(os.path.dirname(__file__), "golden"))
    from analytic_cases import ANALYTIC_CASES, ANALYTIC_SOURCES, GAP
    return ANALYTIC_CASES, ANALYTIC_SOURCES, GAP


def test_island_convmat_matches_oracle(gpu):
    from oracle import rcwa
    from khepri_amd import ops
    from khepri_amd.expansion import Expansion
    cases, _, _ = _analytic_cases()
    extra = [((7, 3), "square", [{"type": "rectangle", "params": [-0.5, -0.5, 0.5, 0.5], "epsilon": 9.0}], 1.0, 0.1),
             ((5, 5), "square", [], 2.5, 0.1),
             ((1, 5), "square", [{"type": "uniform", "params": [], "epsilon": 3.0},
                                 {"type": "disc", "params": [0.3, 0.0, 0.45], "epsilon": 1.0 + 2j}], 4.0, 0.1)]
    for pw, lattice, islands, eps_host, _ in list(cases) + extra:
        lat = rcwa.lattice_matrix(lattice)
        e = Expansion(pw, lat)
        C = ops.convmat_islands(islands, eps_host, pw, e.reciprocal, e.sigma, gpu).cpu().numpy()
        ref = rcwa.analytic_convolution_matrix(islands, eps_host, pw, lat)
        np.testing.assert_allclose(C, ref, atol=1e-13 * max(1.0, np.abs(ref).max()))


def test_island_layers_golden(gpu):
    """Island-layer stacks [A, gap, A] against the reference's R/T."""
    from khepri_amd import Crystal
    g = golden("analytic.npz")
    cases, sources, gap = _analytic_cases()
    for c, (pw, lattice, islands, eps_host, depth) in enumerate(cases):
        for s, (wl, te, tm, th, ph) in enumerate(sources):
            cl = Crystal(pw, lattice=lattice)
            cl.add_layer_analytical("A", islands, eps_host, depth)
            cl.add_layer_uniform("G", *gap)
            cl.set_device(["A", "G", "A"], False)
            cl.set_source(wl, te, tm, th, ph)
            cl.solve()
            r, t = cl.poynting_flux_end()
            assert abs(r - g[f"R{c}"][s]) <= TOL and abs(t - g[f"T{c}"][s]) <= TOL, (c, s, r, t)
        np.testing.assert_allclose(cl.layers["A"].convolution_matrices(gpu)[0].cpu().numpy(), g[f"C{c}"],
                                   atol=1e-13)


def test_island_layers_batched_and_mixed(gpu):
    """A batch of sources through an island + pixmap stack equals per-source
    solves, and CrystalBatch handles island layers."""
    from khepri_amd import Crystal
    from khepri_amd.batch import CrystalBatch
    cases, _, gap = _analytic_cases()
    pw, lattice, islands, eps_host, depth = cases[0]
    pix = W.disc_pixmap()
    wls = np.array([1.5, 1.8, 2.1, 2.4])

    def make(wl):
        cl = Crystal(pw, lattice=lattice)
        cl.add_layer_analytical("A", islands, eps_host, depth)
        cl.add_layer_pixmap("P", pix, 0.3)
        cl.add_layer_uniform("G", *gap)
        cl.set_device(["A", "G", "P"], False)
        cl.set_source(wl, 1.0, 0.0, 5.0, 10.0)
        return cl
    cl = make(wls)
    cl.solve()
    R, T = cl.poynting_flux_end()
    singles = []
    for k, wl in enumerate(wls):
        c1 = make(wl)
        c1.solve()
        r, t = c1.poynting_flux_end()
        singles.append((r, t))
        assert abs(r - R[k]) < 1e-11 and abs(t - T[k]) < 1e-11
    Rb, Tb = CrystalBatch([make(wl) for wl in wls]).solve()
    np.testing.assert_allclose(Rb, [s[0] for s in singles], atol=1e-11)
    np.testing.assert_allclose(Tb, [s[1] for s in singles], atol=1e-11)


def test_star_chain_reports_singular_junction(gpu):
    """A junction whose D = I - SB11 SA22 is exactly singular (SA22 = SB11 = I)
    is reported through the chain's status words and raises numpy's
    LinAlgError, as the reference's np.linalg.solve does
    (khepri/alternative.py:19-30); a regular pair reports nothing."""
    import torch
    from types import SimpleNamespace
    from khepri_amd import ops
    from khepri_amd.layer import stack_layers
    n, B = 6, 3
    eye = torch.eye(n, dtype=torch.complex128, device=gpu)
    SA = torch.zeros((B, 2, 2, n, n), dtype=torch.complex128, device=gpu)
    SB = torch.zeros_like(SA)
    SA[:, 0, 1] = eye
    SA[:, 1, 0] = eye
    SA[:, 1, 1] = eye          # SA22 = I
    SB[:, 0, 1] = eye
    SB[:, 1, 0] = eye
    SB[:, 0, 0] = 0.5 * eye    # SB11 = I/2: D = I/2, regular
    SB[1, 0, 0] = eye          # item 1: SB11 = I, D = 0
    infos = []
    stack_layers(None, [SimpleNamespace(S=SA), SimpleNamespace(S=SB)], [True, True], reverse=False, infos=infos)
    assert len(infos) == 1
    v = infos[0].cpu().numpy()
    assert v[0] == 0 and v[2] == 0 and v[1] != 0
    with pytest.raises(np.linalg.LinAlgError):
        ops.raise_on_info(infos[0], "star product", lu=True)
    ops.raise_on_info(infos[0][[0, 2]], "star product", lu=True)

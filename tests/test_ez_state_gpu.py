"""Project runner and state files on the GPU.

The runner's wiring (units, stacking order and repeats, media, angles,
polarisations, island parameters) is pinned by the normal-incidence
characteristic-matrix closed form and by the oracle Crystal on the same
structures; R and T within 1e-10.  A saved and reloaded solved Crystal
returns the same R / T bit for bit and solves again."""
import copy
import os

import numpy as np
import pytest

from oracle import crystal as OC
from test_ez_state import PROJECT

pytestmark = pytest.mark.gpu


def _abeles(ns, ds, wl, n_in, n_out):
    k0 = 2 * np.pi / wl
    M = np.eye(2, dtype=complex)
    for n, d in zip(ns, ds):
        dl = k0 * n * d
        M = M @ np.array([[np.cos(dl), 1j * np.sin(dl) / n], [1j * n * np.sin(dl), np.cos(dl)]])
    B, Cc = M @ np.array([1, n_out])
    r = (n_in * B - Cc) / (n_in * B + Cc)
    t = 2 * n_in / (n_in * B + Cc)
    return abs(r) ** 2, (n_out / n_in) * abs(t) ** 2


def test_uniform_stack_spectrum_closed_form(gpu, tmp_path):
    from khepri_amd import ez
    proj = copy.deepcopy(PROJECT)
    proj["layers"]["L"].pop("geometry")
    proj["lattice"]["pw"] = [1, 1]
    prm = proj["studies"]["spec"]["parameters"]
    prm["theta"] = {"start": 0.0, "end": 0.0, "steps": 1}
    prm["frequency"] = {"start": 0.3, "end": 1.7, "steps": 9}
    prm["output_data"] = os.path.join(tmp_path, "spec.csv")
    rows = ez.run(proj, "spec")
    assert rows.shape == (9, 7)
    stack = ez.flatten_stack(proj, "Dev")
    eps = {"H": 6.0, "L": 1.0}
    ns = [np.sqrt(eps[n]) for n in stack]
    ds = [proj["layers"][n]["depth"] for n in stack]
    for phi, theta, f, wl, R, T, A in rows:
        Rr, Tr = _abeles(ns, ds, 1 / f, 1.0, 1.5)
        assert abs(R - Rr) < 1e-10 and abs(T - Tr) < 1e-10 and abs(A - (1 - R - T)) < 1e-15
        assert abs(wl - 200e-9 / f) < 1e-22
    np.testing.assert_array_equal(np.loadtxt(prm["output_data"], delimiter=",", skiprows=1), rows)


def test_island_spectrum_vs_oracle(gpu):
    """Oblique incidence (two polar angles, one azimuth), mixed s / p, disc and
    rectangle islands: every row against the oracle Crystal."""
    from khepri_amd import ez
    proj = copy.deepcopy(PROJECT)
    prm = proj["studies"]["spec"]["parameters"]
    prm["phi"] = {"start": 30.0, "end": 30.0, "steps": 1}
    rows = ez.run(proj, "spec")
    s, p = ez.polarization(prm)
    stack = ez.flatten_stack(proj, "Dev")
    for phi, theta, f, wl, R, T, A in rows:
        cl = OC.Crystal((3, 3), epsi=1.0, epse=2.25)
        cl.add_layer_uniform("H", 6.0, 0.1)
        cl.add_layer_analytical("L", ez.islands(proj, proj["layers"]["L"]), 1.0, 0.3)
        cl.set_device(stack)
        cl.set_source(1 / f, s, p, theta, phi)
        cl.solve()
        Rr, Tr = cl.poynting_flux_end()
        assert abs(R - Rr) < 1e-10 and abs(T - Tr) < 1e-10


def test_state_round_trip_solved(gpu, tmp_path):
    import khepri_amd as K
    from khepri_amd import ez
    cl = ez.build_crystal(PROJECT, "Dev")
    cl.set_source(np.array([2.0, 2.7, 3.1]), 1.0, 0.0, 5.0, 0.0)
    cl.solve(rt_only=False)
    R, T = cl.poynting_flux_end()
    path = os.path.join(tmp_path, "solved.pkl")
    K.save(path, cl)
    c2 = K.load(path)
    assert c2.Stot.is_cuda
    R2, T2 = c2.poynting_flux_end()
    np.testing.assert_array_equal(R2, R)
    np.testing.assert_array_equal(T2, T)
    c2.set_source(np.array([2.0, 2.7, 3.1]), 1.0, 0.0, 5.0, 0.0)
    c2.solve()
    R3, T3 = c2.poynting_flux_end()
    assert np.abs(R3 - R).max() < 1e-13 and np.abs(T3 - T).max() < 1e-13


def test_runner_command_line(gpu, tmp_path):
    """python -m khepri_amd.ez -s STUDY -p PROJECT.json writes the study's CSV."""
    import json
    import subprocess
    import sys
    proj = copy.deepcopy(PROJECT)
    prm = proj["studies"]["spec"]["parameters"]
    prm["output_data"] = os.path.join(tmp_path, "cli.csv")
    path = os.path.join(tmp_path, "project.json")
    with open(path, "w") as f:
        json.dump(proj, f)
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    r = subprocess.run([sys.executable, "-m", "khepri_amd.ez", "-s", "spec", "-p", path], cwd=root,
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr[-2000:]
    rows = np.loadtxt(prm["output_data"], delimiter=",", skiprows=1)
    assert rows.shape == (8, 7)
    from khepri_amd import ez
    np.testing.assert_allclose(rows, ez.run(PROJECT, "spec"), rtol=0, atol=1e-13)

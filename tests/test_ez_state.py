"""Project runner (khepri_amd.ez, after khepri/ez.py) and state files
(khepri_amd.save / load, after khepri/__init__.py:12-19): the host logic,
no solves."""
import os

import numpy as np
import pytest

from khepri_amd import ez

PROJECT = {
    "lattice": {"lattice_size": 200e-9, "a1": [1, 0], "a2": [0, 1], "pw": [3, 3],
                "material_incident": "Air", "material_substrate": "Glass"},
    "materials": {"Air": {"type": "constant", "epsilon": 1.0, "color": "white"},
                  "Glass": {"type": "constant", "epsilon": 2.25, "color": "gray"},
                  "TiO2": {"type": "constant", "epsilon": 6.0, "color": "red"},
                  "Drude": {"type": "drude", "epsilon": 1.0}},
    "layers": {"H": {"material": "TiO2", "depth": 0.1},
               "L": {"material": "Air", "depth": 0.3,
                     "geometry": [{"shape": "disc", "center": [0.5, 0.5], "radius": 0.2, "material": "TiO2"},
                                  {"shape": "rectangle", "center": [0.1, 0.0], "size": [0.3, 0.6],
                                   "material": "Glass"}]},
               "X": {"material": "Drude", "depth": 0.1}},
    "stacking": {"P": {"layers": ["H", "L"], "repeat": 3},
                 "Dev": {"layers": ["P", "H"]},
                 "Nest": {"layers": ["Dev", "L"], "repeat": 2}},
    "studies": {"spec": {"type": "prefab_spectrum", "device": "Dev",
                         "parameters": {"frequency": {"start": 0.2, "end": 0.5, "steps": 4},
                                        "theta": {"start": 0.0, "end": 20.0, "steps": 2},
                                        "phi": {"start": 0.0, "end": 0.0, "steps": 1},
                                        "E0": {"s_real": 1.0, "s_imag": 0.0, "p_real": 0.0, "p_imag": 0.5}}},
                "view": {"type": "prefab_view_structure", "device": "Dev"}},
}


def test_flatten_stack_repeats_and_nesting():
    assert ez.flatten_stack(PROJECT, "H") == ["H"]
    assert ez.flatten_stack(PROJECT, "P") == ["H", "L"] * 3
    assert ez.flatten_stack(PROJECT, "Dev") == ["H", "L"] * 3 + ["H"]
    assert ez.flatten_stack(PROJECT, "Nest") == (["H", "L"] * 3 + ["H", "L"]) * 2
    with pytest.raises(KeyError):
        ez.flatten_stack(PROJECT, "nope")


def test_islands_and_materials():
    isl = ez.islands(PROJECT, PROJECT["layers"]["L"])
    assert isl == [{"type": "disc", "params": [0.5, 0.5, 0.2], "epsilon": 6.0},
                   {"type": "rectangle", "params": [0.1, 0.0, 0.3, 0.6], "epsilon": 2.25}]
    assert ez.islands(PROJECT, PROJECT["layers"]["H"]) == []
    assert ez.material_epsilon(PROJECT, "Glass") == 2.25
    with pytest.raises(NotImplementedError):
        ez.material_epsilon(PROJECT, "Drude")


def test_sweep_order_and_polarization():
    phi, theta, freq = ez.spectrum_sources(PROJECT["studies"]["spec"]["parameters"])
    assert len(freq) == 8
    np.testing.assert_array_equal(theta, [0.0] * 4 + [20.0] * 4)          # frequency fastest
    np.testing.assert_allclose(freq, np.tile(np.linspace(0.2, 0.5, 4), 2))
    assert (phi == 0).all()
    assert ez.polarization(PROJECT["studies"]["spec"]["parameters"]) == (1.0, 0.5j)


def test_rows_file_format(tmp_path):
    rows = np.arange(14, dtype=float).reshape(2, 7)
    path = os.path.join(tmp_path, "spectrum.csv")
    ez.write_rows(rows, path)
    with open(path) as f:
        lines = f.read().splitlines()
    assert lines[0] == "PHI,THETA,FREQUENCY,WAVELENGTH,R,T,A"
    np.testing.assert_array_equal(np.loadtxt(path, delimiter=",", skiprows=1), rows)


def test_view_study_not_provided():
    with pytest.raises(NotImplementedError):
        ez.run(PROJECT, "view")


def test_state_round_trip_unsolved(tmp_path):
    """An unsolved Crystal with every layer kind, its sources and a tensor
    survive save / load (tensors come back on the CPU here)."""
    import torch
    import khepri_amd as K
    cl = K.Crystal((3, 3), epsi=2.0)
    cl.add_layer_uniform("U", 4.0, 0.3)
    cl.add_layer_pixmap("P", np.full((16, 16), 3.0), 0.2)
    cl.add_layer_analytical("A", [{"type": "disc", "params": [0.5, 0.5, 0.2], "epsilon": 5.0}], 1.0, 0.1)
    cl.set_device(["U", "P", "A", "U"], [False, True, False, False])
    cl.set_source(np.array([1.0, 1.5]), 1.0, 0.5j, 10.0, 30.0)
    extra = torch.arange(4).to(torch.complex128)
    path = os.path.join(tmp_path, "state.pkl")
    K.save(path, {"crystal": cl, "extra": extra, "again": extra})
    d = K.load(path)
    c2 = d["crystal"]
    assert c2.pw == (3, 3) and c2.global_stacking == cl.global_stacking
    assert c2.stack_retain_mask == cl.stack_retain_mask
    np.testing.assert_array_equal(c2.layers["P"].epsilon, cl.layers["P"].epsilon)
    assert c2.layers["A"].epsilon == cl.layers["A"].epsilon
    np.testing.assert_array_equal(c2._kp, cl._kp)
    np.testing.assert_array_equal(c2._tm, cl._tm)
    assert torch.equal(d["extra"], extra) and d["again"] is d["extra"]


def test_set_source_rejects_ambiguous_kp_layouts():
    import numpy as np
    import pytest
    from khepri_amd import Crystal
    cl = Crystal((3, 3))
    kps = np.arange(6.0).reshape(2, 3) * 0.01     # (2, B): the transposed layout
    with pytest.raises(ValueError):
        cl.set_source(np.full(3, 1.5), 1.0, 0.0, kp=kps)
    with pytest.raises(ValueError):
        cl.set_source(1.5, 1.0, 0.0, kp=np.zeros((2, 3, 2)))
    cl.set_source(np.full(3, 1.5), 1.0, 0.0, kp=kps.T)   # (B, 2) is accepted
    assert cl.batch == 3 and np.allclose(cl.kp[:, 0], kps[0]) and np.allclose(cl.kp[:, 1], kps[1])
    cl.set_source(1.5, 1.0, 0.0, kp=(0.1, 0.2))           # one (kx, ky)
    assert cl.batch == 1


def test_deferred_status_records_then_raises():
    """ops.deferred_status: raise_on_info records the words (no host read)
    and check_status raises later, as the eager check would."""
    import torch
    from khepri_amd import ops
    sink = []
    good = torch.zeros(4, dtype=torch.int32)
    bad = torch.tensor([0, 0, 3, 0], dtype=torch.int32)
    with ops.deferred_status(sink):
        ops.raise_on_info(good, "ok", lu=True)
        ops.raise_on_info(bad, "lu", lu=True)
    assert len(sink) == 2
    with pytest.raises(np.linalg.LinAlgError):
        ops.check_status(sink)
    ops.check_status(sink[:1])
    with pytest.raises(np.linalg.LinAlgError):
        ops.raise_on_info(bad, "lu", lu=True)

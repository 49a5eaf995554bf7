"""JSON project runner (khepri/ez.py) on the GPU Crystal.

    python -m khepri_amd.ez -s STUDY -p PROJECT.json

A project file holds, as in the reference's launcher format:

* ``lattice``: ``lattice_size`` a (metres), lattice vectors ``a1``, ``a2`` in
  units of a, harmonics ``pw``, and the ``material_incident`` /
  ``material_substrate`` half spaces;
* ``materials``: name -> {"type": "constant", "epsilon": eps, "color": ...};
* ``layers``: name -> {"material": host, "depth": d (units of a),
  "geometry": [islands]}, an island being {"shape": "disc", "center": [x, y],
  "radius": r, "material": m} or {"shape": "rectangle", "center": [x0, y0],
  "size": [x1, y1], "material": m} (units of a);
* ``stacking``: name -> {"layers": [layer or stack names], "repeat": k}, the
  first name nearest the incident medium;
* ``studies``: name -> {"type": "prefab_spectrum", "device": stack or layer,
  "parameters": {"frequency" | "theta" | "phi": {"start", "end", "steps"},
  "E0": {"s_real", "s_imag", "p_real", "p_imag"}, "output_data": csv path,
  "output_figure": image path}}.

A spectrum sweeps every (phi, theta, frequency) of the three linspaces (phi
slowest, frequency fastest; angles in degrees; frequency in units of c/a, so
the wavelength is a / frequency) and writes the rows (PHI, THETA, FREQUENCY,
WAVELENGTH, R, T, A = 1 - R - T) like the reference.  The whole sweep is one
batched Crystal solve.  Layers without geometry are uniform layers; layers
with islands are analytical island layers, with the island parameters passed
as the reference passes them: (center, radius) for a disc and (center, size)
as the (lower-left, upper-right) corners of a rectangle.  s and p
polarisation are the TE and TM amplitudes of ``Crystal.set_source``.
``prefab_view_structure`` studies (3-D drawings) are not provided.
"""
from __future__ import annotations

import argparse
import json
import sys
from itertools import product

import numpy as np

HEADER = "PHI,THETA,FREQUENCY,WAVELENGTH,R,T,A"


def load_project(project):
    """A project dict from a path or a dict."""
    if isinstance(project, dict):
        return project
    with open(project, "r") as f:
        return json.load(f)


def material_epsilon(config, name):
    mat = config["materials"][name]
    kind = mat.get("type", "constant")
    if kind != "constant":
        raise NotImplementedError(f"material {name!r}: only constant materials are supported, got {kind!r}")
    return mat["epsilon"]


def flatten_stack(config, name):
    """Layer names of a stack or layer, first nearest the incident medium
    (nested stacks expanded, each repeated ``repeat`` times)."""
    stacking = config.get("stacking", {})
    if name in stacking:
        seq = []
        for sub in stacking[name]["layers"]:
            seq += flatten_stack(config, sub)
        return seq * int(stacking[name].get("repeat", 1))
    if name in config.get("layers", {}):
        return [name]
    raise KeyError(f"{name!r} is neither a stack nor a layer of the project")


def islands(config, layer):
    """Island descriptions of a layer for Crystal.add_layer_analytical."""
    out = []
    for isl in layer.get("geometry", []) or []:
        params = list(isl["center"])
        if "radius" in isl:
            params.append(isl["radius"])
        elif "size" in isl:
            params.extend(isl["size"])
        out.append({"type": isl["shape"], "params": params, "epsilon": material_epsilon(config, isl["material"])})
    return out


def spectrum_sources(params):
    """(phi, theta, frequency) columns of the sweep, phi slowest."""
    def lin(p):
        return np.linspace(p["start"], p["end"], p["steps"])
    cfg = np.array(list(product(lin(params["phi"]), lin(params["theta"]), lin(params["frequency"]))), dtype=float)
    return cfg.reshape(-1, 3).T


def polarization(params):
    e0 = params.get("E0", {})
    s = e0.get("s_real", 0.5) + 1j * e0.get("s_imag", 0.0)
    p = e0.get("p_real", 0.5) + 1j * e0.get("p_imag", 0.0)
    return s, p


def build_crystal(config, device_name):
    """Crystal (units of a) with the device's layers stacked."""
    from .crystal import Crystal
    lat = config["lattice"]
    cl = Crystal(tuple(lat["pw"]), lattice=np.array([lat["a1"], lat["a2"]], dtype=float),
                 epsi=material_epsilon(config, lat["material_incident"]),
                 epse=material_epsilon(config, lat["material_substrate"]))
    stack = flatten_stack(config, device_name)
    for name in dict.fromkeys(stack):
        layer = config["layers"][name]
        host = material_epsilon(config, layer["material"])
        isl = islands(config, layer)
        if isl:
            cl.add_layer_analytical(name, isl, host, layer["depth"])
        else:
            cl.add_layer_uniform(name, host, layer["depth"])
    cl.set_device(stack)
    return cl


def spectrum(config, params, device_name):
    """Rows (phi, theta, frequency, wavelength, R, T, A) of a spectrum study."""
    a = config["lattice"]["lattice_size"]
    phi, theta, freq = spectrum_sources(params)
    s, p = polarization(params)
    cl = build_crystal(config, device_name)
    cl.set_source(1.0 / freq, s, p, theta, phi)
    cl.solve()
    R, T = cl.poynting_flux_end()
    R, T = np.atleast_1d(np.asarray(R, dtype=float)), np.atleast_1d(np.asarray(T, dtype=float))
    return np.column_stack([phi, theta, freq, a / freq, R, T, 1.0 - R - T])


def write_rows(rows, path):
    np.savetxt(path, rows, header=HEADER, delimiter=",", comments="")


def run(project, study_name):
    """Run one study of a project; returns the spectrum rows."""
    config = load_project(project)
    study = config["studies"][study_name]
    kind = study["type"]
    if kind == "prefab_spectrum":
        params = study["parameters"]
        rows = spectrum(config, params, study["device"])
        if "output_data" in params:
            write_rows(rows, params["output_data"])
        if "output_figure" in params:
            import matplotlib
            matplotlib.use("Agg")
            import matplotlib.pyplot as plt
            fig, ax = plt.subplots()
            ax.plot(rows)
            fig.savefig(params["output_figure"])
            plt.close(fig)
        return rows
    if kind == "prefab_view_structure":
        raise NotImplementedError("prefab_view_structure (3-D drawing of the stack) is not provided")
    raise ValueError(f"unknown study type {kind!r}")


def main(argv=None):
    parser = argparse.ArgumentParser(prog="python -m khepri_amd.ez")
    parser.add_argument("-s", "--study", required=True)
    parser.add_argument("-p", "--project", required=True)
    args = parser.parse_args(argv)
    run(args.project, args.study)
    return 0


if __name__ == "__main__":
    sys.exit(main())

"""Field maps (fields_volume / fields_coords_xy) against the reference's golden
field volumes and the oracle; tolerance: relative field error <= 1e-8."""
import numpy as np
import pytest

from conftest import golden
from khepri_amd import workloads as W

pytestmark = pytest.mark.gpu


def rel_err(a, b):
    return np.abs(a - b).max() / np.abs(b).max()


def test_fields_patterned_oblique(gpu):
    from khepri_amd import Crystal, Drawing
    g = golden("fields_patterned.npz")
    d = Drawing((64, 64), 12)
    d.disc((0.1, -0.05), 0.3, 2.0)
    cl = Crystal((3, 3), epsi=1.0, epse=2.25)
    cl.add_layer_uniform("S1", 1, 1.1)
    cl.add_layer_pixmap("Scyl", d.canvas(), 0.55)
    cl.set_device(["Scyl", "S1", "Scyl"], [True] * 3)
    cl.set_source(1 / 0.52, 1.0, 0.5, 20.0, 35.0)
    cl.solve()
    E, H = cl.fields_volume(g["x"], g["y"], g["z"])
    assert E.shape == g["E"].shape
    assert rel_err(E, g["E"]) < 1e-8 and rel_err(H, g["H"]) < 1e-8
    R, T = cl.poynting_flux_end()
    assert abs(R - g["R"]) < 1e-9 and abs(T - g["T"]) < 1e-9
    # one depth through fields_coords_xy
    E1, H1 = cl.fields_coords_xy(g["x"], g["y"], float(g["z"][2]))
    assert rel_err(E1, g["E"][2]) < 1e-8 and rel_err(H1, g["H"][2]) < 1e-8


def test_fields_uniform_matches_oracle(gpu):
    """Reference test test/integration/test_pol.py:12-22 (uniform slab, one
    point), checked against the oracle on a small grid."""
    from khepri_amd import Crystal
    from oracle import crystal as OC
    x = y = np.array([[0.0, 0.3], [0.6, 0.9]])
    z = np.array([-0.2, 0.0, 0.15, 0.31, 0.5])
    out = []
    for C in (Crystal, OC.Crystal):
        cl = C((3, 3))
        cl.add_layer_uniform("U", 4, 0.3)
        cl.set_source(1.1, 1, 0)
        cl.set_device(["U"], True)
        cl.solve()
        out.append(cl.fields_volume(x, y, z))
    assert rel_err(out[0][0], out[1][0]) < 1e-10 and rel_err(out[0][1], out[1][1]) < 1e-10


def test_fields_twisted_small(gpu):
    from khepri_amd import Crystal, Expansion, Layer
    g = golden("twisted_small.npz")
    pix = W.disc_pixmap(128, 1.0, 0.3, 4.0)
    e1, e2 = Expansion((3, 3)), Expansion((3, 3))
    e2.rotate(float(g["field_angle"]))
    etw = e1 + e2
    cl = Crystal.from_expansion(etw)
    cl.add_layer("L1", Layer.pixmap(e1, pix, 0.4), extended=True)
    cl.add_layer("gap", Layer.uniform(etw, 1.0, 0.2))
    cl.add_layer("L2", Layer.pixmap(e2, pix, 0.4), extended=True)
    cl.set_device(["L1", "gap", "L2"], True)
    cl.set_source(1 / float(g["freq"]), 1.0, 0.0)
    cl.solve()
    E, H = cl.fields_volume(g["x"], g["y"], g["z"])
    assert rel_err(E, g["E"]) < 1e-8 and rel_err(H, g["H"]) < 1e-8


def _disc_stack(C, pw, pix):
    cl = C(pw)
    cl.add_layer_uniform("S1", 1.0, 1.1)
    cl.add_layer_pixmap("Scyl", pix, 0.55)
    cl.set_device(["Scyl", "S1", "Scyl"], [True] * 3)
    return cl


def test_brillouin_zone_loop_batched(gpu):
    """The reference's Brillouin-zone loop (examples/bzi/twisted_phc.py:126-147:
    set_source(wl, nan, nan, kp) -> solve -> amplitudes_from_fields ->
    fields_volume(incident_fields=F), one Bloch vector at a time) as ONE
    batched solve and ONE batched fields_volume over every Bloch vector,
    against the oracle running the reference's loop.  Relative field error
    <= 1e-8 per Bloch vector."""
    from functools import partial
    from khepri_amd import Crystal, beams as B
    from oracle import beams as OB, crystal as OC
    pw, wl = (3, 3), 1.6
    pix = W.disc_pixmap(64, 4.0, 0.3, 1.0)
    ns, bzs = 8, (2, 2)
    xs = np.arange(ns * bzs[0]) / ns
    X, Y = np.meshgrid(xs, xs)
    Zs = np.zeros_like(X)
    f = OB.shifted_rotated_fields(partial(OB.paraxial_gaussian_field_fn, beam_waist=1.0), X, Y, Zs, wl, 1.0, 1.0,
                                  -0.3, 0.1, 0.2, 0.0)
    fr = np.swapaxes(np.swapaxes(f[..., 0], 0, 2), 1, 3)
    kps = W.bz_grid(bzs).reshape(2, -1).T                      # (4, 2): one Bloch vector per row
    xg, yg = np.meshgrid(np.linspace(0, 1, 6), np.linspace(0, 1, 5))
    zs = [-0.1, 0.3, 1.0, 1.9, 2.5]
    cl = _disc_stack(Crystal, pw, pix)
    cl.set_source(np.full(len(kps), wl), np.nan, np.nan, kp=kps)
    cl.solve()
    F = np.stack([B.amplitudes_from_fields(fr, cl.expansion, wl, tuple(k), X, Y, bzs).reshape(-1) for k in kps])
    E, H = cl.fields_volume(xg, yg, zs, incident_fields=F)
    assert E.shape == (len(kps), len(zs), 3) + xg.shape and H.shape == E.shape
    oc = _disc_stack(OC.Crystal, pw, pix)
    for b, k in enumerate(kps):
        oc.set_source(wl, np.nan, np.nan, kp=tuple(k))
        oc.solve()
        Fo = OB.amplitudes_from_fields(fr, oc.expansion.g, pw, tuple(k), X, Y, bzs).reshape(-1)
        assert np.abs(Fo - F[b]).max() <= 1e-12 * np.abs(Fo).max()
        Eo, Ho = oc.fields_volume(xg, yg, zs, incident_fields=Fo)
        assert rel_err(E[b], Eo) < 1e-8 and rel_err(H[b], Ho) < 1e-8, b
    # the plane-wave sources themselves (incident_fields=None), batched over wavelengths
    freqs = np.array([0.5, 0.55, 0.6])
    cl = _disc_stack(Crystal, pw, pix)
    cl.set_source(1 / freqs, 1.0, 0.5, 20.0, 35.0)
    cl.solve()
    E, H = cl.fields_volume(xg, yg, zs)
    E1, H1 = cl.fields_coords_xy(xg, yg, 0.3)
    for b, fq in enumerate(freqs):
        oc.set_source(1 / fq, 1.0, 0.5, 20.0, 35.0)
        oc.solve()
        Eo, Ho = oc.fields_volume(xg, yg, zs)
        assert rel_err(E[b], Eo) < 1e-8 and rel_err(H[b], Ho) < 1e-8, b
        assert rel_err(E1[b], Eo[1]) < 1e-8 and rel_err(H1[b], Ho[1]) < 1e-8, b
    with pytest.raises(ValueError):
        cl.fields_volume(xg, yg, zs, incident_fields=np.zeros((2, 4 * 9)))

"""Extended-basis (twisted bilayer) R/T against the reference's golden vectors."""
import numpy as np
import pytest

from conftest import golden
from khepri_amd import workloads as W

pytestmark = pytest.mark.gpu
TOL = 1e-9


def twisted(pw, angle, pix, fields=False):
    from khepri_amd import Crystal, Expansion, Layer
    e1, e2 = Expansion(pw), Expansion(pw)
    e2.rotate(angle)
    etw = e1 + e2
    cl = Crystal.from_expansion(etw)
    cl.add_layer("L1", Layer.pixmap(e1, pix, 0.4), extended=True)
    cl.add_layer("gap", Layer.uniform(etw, 1.0, 0.2))
    cl.add_layer("L2", Layer.pixmap(e2, pix, 0.4), extended=True)
    cl.set_device(["L1", "gap", "L2"], fields)
    return cl


@pytest.mark.parametrize("rt_only", [False, True])
def test_twisted_small_rt(gpu, rt_only):
    g = golden("twisted_small.npz")
    pix = W.disc_pixmap(128, 1.0, 0.3, 4.0)
    for a, r0, t0 in zip(g["angles"], g["R"], g["T"]):
        cl = twisted((3, 3), float(a), pix)
        cl.set_source(1 / float(g["freq"]), 1.0, 0.0)
        cl.solve(rt_only=rt_only)
        r, t = cl.poynting_flux_end()
        assert abs(r - r0) <= TOL and abs(t - t0) <= TOL, (a, r, r0, t, t0)


def test_twisted_625_harmonics_rt(gpu):
    """Base pw (5,5): 625 extended harmonics, n = 1250; both seeded twist
    angles of the golden file."""
    g = golden("twisted.npz")
    cfg = W.twisted_bilayer()
    assert len(g["angle_index"]) == 2
    for ai, r0, t0 in zip(g["angle_index"], g["R"], g["T"]):
        cl = twisted((5, 5), np.deg2rad(cfg["angles_deg"][int(ai)]), cfg["pixmap"])
        cl.set_source(1 / float(g["freq"]), 1.0, 0.0)
        cl.solve(rt_only=True)
        r, t = cl.poynting_flux_end()
        assert abs(r - r0) <= TOL and abs(t - t0) <= TOL, (int(ai), r, r0, t, t0)


def test_twisted_625_harmonics_fields(gpu):
    """fields_volume of the 625-harmonic twisted bilayer (first seeded angle,
    middle frequency) on an 8 x 8 x 3 grid against the reference's golden
    E and H: relative error ||E - E_ref|| / ||E_ref|| <= 1e-8."""
    g = golden("twisted.npz")
    cfg = W.twisted_bilayer()
    cl = twisted((5, 5), np.deg2rad(cfg["angles_deg"][0]), cfg["pixmap"], fields=True)
    cl.set_source(1 / float(g["freq"]), 1.0, 0.0)
    cl.solve()
    E, H = cl.fields_volume(g["x"], g["y"], g["z"])
    for got, ref in ((E, g["E"]), (H, g["H"])):
        got = np.asarray(got)
        assert got.shape == ref.shape
        rel = np.linalg.norm(got - ref) / np.linalg.norm(ref)
        print(f"field rel. error {rel:.2e}")
        assert rel <= 1e-8


def _extended_over_crystal(mod, pw, angle, pix):
    e1, e2 = mod.Expansion(pw), mod.Expansion(pw)
    e2.rotate(angle)
    etw = e1 + e2
    base = mod.Crystal.from_expansion(e1)
    base.add_layer_pixmap("P", pix, 0.4)
    base.set_device(["P"], False)
    base.fields = False  # the reference's ExtendedLayer reads base.fields
    top = mod.Crystal.from_expansion(etw)
    top.add_layer("L1", mod.ExtendedLayer(etw, base))
    top.add_layer("gap", mod.Layer.uniform(etw, 1.0, 0.2))
    top.add_layer("L2", mod.Layer.pixmap(e2, pix, 0.4), extended=True)
    top.set_device(["L1", "gap", "L2"], False)
    return top


def test_extended_layer_over_crystal_base(gpu):
    """An ExtendedLayer whose base is a Crystal (khepri/extension.py:88-95: the
    base crystal's whole stack solved at kp + g for every g) against the oracle
    with the same branch; two frequencies in one batch."""
    import khepri_amd
    from oracle import crystal as OC
    pix = W.disc_pixmap(64, 1.0, 0.3, 4.0)
    freqs = np.array([0.52, 0.61])
    cl = _extended_over_crystal(khepri_amd, (3, 3), 0.31, pix)
    cl.set_source(1 / freqs, 1.0, 0.0)
    cl.solve()
    R, T = cl.poynting_flux_end()
    oc = _extended_over_crystal(OC, (3, 3), 0.31, pix)
    for b, fq in enumerate(freqs):
        oc.set_source(1 / fq, 1.0, 0.0)
        oc.solve()
        r, t = oc.poynting_flux_end()
        assert abs(R[b] - r) <= TOL and abs(T[b] - t) <= TOL, (b, R[b], r, T[b], t)


def test_extended_freespace_matches_oracle_layer(gpu):
    """extended_freespace (khepri/extension.py:114-126) against the W0, V0 the
    oracle's ExtendedLayer builds for its fields."""
    from khepri_amd import Expansion
    from khepri_amd.extension import extended_freespace
    from oracle import crystal as OC
    pw, angle, wl, kp = (3, 3), 0.2, 1.7, (0.11, -0.07)
    e1, e2 = Expansion(pw), Expansion(pw)
    e2.rotate(angle)
    W0, V0 = extended_freespace(e1, e2.g_vectors.T, wl, 1, kp)
    o1, o2 = OC.Expansion(pw), OC.Expansion(pw)
    o2.rotate(angle)
    el = OC.ExtendedLayer(o1 + o2, OC.Layer.uniform(o1, 1.0, 0.3))
    el.fields = True
    el.solve(np.array(kp), wl)
    assert np.abs(W0.cpu().numpy() - el.W0).max() < 1e-13
    assert np.abs(V0.cpu().numpy() - el.V0).max() < 1e-12 * np.abs(el.V0).max()

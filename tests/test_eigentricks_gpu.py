"""Band-structure tool on the GPU (khepri_amd.eigentricks, libkhx
khx_scattering_*) against the oracle (oracle/eigentricks.py, restating
khepri/eigentricks.py:5-55) and closed forms.

Tolerances: the split is bit-exact; Bloch factors within 1e-10 (on shell,
|w| = 1) and 1e-9 max(1, |w|) elsewhere; eigenvector columns are checked by
unit norm and by the pencil residual |Sl v - w Sr v| / ((|Sl| + |w| |Sr|)
|v|) <= 1e-12 (their phase is not fixed by the problem); determinants within
1e-10 relative."""
import numpy as np
import pytest

from oracle import eigentricks as OE

pytestmark = pytest.mark.gpu


def _rand_S(rng, B, ng, cond=True):
    S = rng.standard_normal((B, 4 * ng, 4 * ng)) + 1j * rng.standard_normal((B, 4 * ng, 4 * ng))
    S *= 0.5 / np.sqrt(4 * ng)
    if cond:
        S[:, 2 * ng:, 2 * ng:] += np.eye(2 * ng)
    return S


def _match_sets(w, ref, on_tol=1e-10, rel=1e-9):
    """Each Bloch factor of ref has a distinct partner in w within tolerance."""
    w, ref = list(np.asarray(w)), np.asarray(ref)
    for r in ref[np.argsort(-np.abs(np.abs(ref) - 1))]:   # hardest (far off shell) first
        d = np.abs(np.asarray(w) - r)
        k = int(np.argmin(d))
        tol = on_tol if abs(abs(r) - 1) < 1e-6 else rel * max(1.0, abs(r))
        assert d[k] <= tol, (r, w[k], d[k])
        w.pop(k)


def _residual(S, w, v):
    Sl, Sr = OE.scattering_splitlr(S)
    nl, nr = np.linalg.norm(Sl, 2), np.linalg.norm(Sr, 2)
    r = np.linalg.norm(Sl @ v - (Sr @ v) * w[None, :], axis=0)
    return (r / ((nl + np.abs(w) * nr) * np.linalg.norm(v, axis=0))).max()


def _stack_crystal(pw, layers, wl, kp):
    from khepri_amd import Crystal
    cl = Crystal(pw)
    names = []
    for i, (kind, eps, d) in enumerate(layers):
        name = f"L{i}"
        if kind == "uniform":
            cl.add_layer_uniform(name, eps, d)
        else:
            cl.add_layer_pixmap(name, eps, d)
        names.append(name)
    cl.set_device(names)
    wl = np.atleast_1d(np.asarray(wl, dtype=float))
    if kp is None:
        cl.set_source(wl, 1.0, 1.0, 0.0, 0.0)
    else:
        cl.set_source(wl, 1.0, 1.0, kp=np.asarray(kp, dtype=float))
    cl.solve(rt_only=False)
    return cl


def _disc(n=64, r=0.3, eps=8.9):
    x = (np.arange(n) + 0.5) / n - 0.5
    X, Y = np.meshgrid(x, x, indexing="ij")
    return np.where(X ** 2 + Y ** 2 < r ** 2, eps, 1.0)


@pytest.mark.parametrize("ng,B", [(1, 1), (2, 3), (25, 2)])
def test_splitlr_bitexact(gpu, ng, B):
    from khepri_amd import eigentricks as E
    S = _rand_S(np.random.default_rng(ng), B, ng)
    Sl, Sr = E.scattering_splitlr(S)
    for b in range(B):
        rl, rr = OE.scattering_splitlr(S[b])
        assert np.array_equal(Sl[b], rl) and np.array_equal(Sr[b], rr)
    Sl1, Sr1 = E.scattering_splitlr(S[0])
    assert Sl1.shape == S[0].shape and np.array_equal(Sl1, Sl[0]) and np.array_equal(Sr1, Sr[0])


@pytest.mark.parametrize("ng", [1, 3, 12])
def test_eigenvalues_random_vs_oracle(gpu, ng):
    from khepri_amd import eigentricks as E
    S = _rand_S(np.random.default_rng(100 + ng), 4, ng)
    w, v = E.scattering_eigenvalues(S)
    assert w.shape == (4, 4 * ng) and v.shape == (4, 4 * ng, 4 * ng)
    for b in range(4):
        wr, _ = OE.scattering_eigenvalues(S[b])
        _match_sets(w[b], wr)
        assert np.allclose(np.linalg.norm(v[b], axis=0), 1.0, atol=1e-12)
        assert _residual(S[b], w[b], v[b]) < 1e-12
    w1, v1 = E.scattering_eigenvalues(S[2])
    _match_sets(w1, w[2], on_tol=1e-12, rel=1e-12)


@pytest.mark.parametrize("wl", [[4.1, 2.3, 1.7, 1.1, 0.93]])
def test_bragg_band_from_crystal(gpu, wl):
    """Period of our own Crystal (pw 1 x 1, two uniform layers) at five
    wavelengths in one batch: w + 1/w = 2 cos(k1 d1) cos(k2 d2) - (n1/n2 +
    n2/n1) sin(k1 d1) sin(k2 d2), on shell exactly inside the bands."""
    from khepri_amd import eigentricks as E
    e1, d1, e2, d2 = 12.0, 0.2, 1.0, 0.8
    cl = _stack_crystal((1, 1), [("uniform", e1, d1), ("uniform", e2, d2)], wl, None)
    S = E.scattering_matrix(cl)
    assert S.shape == (5, 4, 4)
    w, v = E.scattering_eigenvalues(S)
    w = w.cpu().numpy()
    n1, n2 = np.sqrt(e1), np.sqrt(e2)
    for b, lam in enumerate(wl):
        k1, k2 = 2 * np.pi * n1 / lam, 2 * np.pi * n2 / lam
        rhs = 2 * np.cos(k1 * d1) * np.cos(k2 * d2) - (n1 / n2 + n2 / n1) * np.sin(k1 * d1) * np.sin(k2 * d2)
        assert np.allclose(w[b] + 1 / w[b], rhs, atol=1e-10, rtol=0)
        assert E.on_shell(w[b]).all() == (abs(rhs) <= 2)
        wr, _ = OE.scattering_eigenvalues(S[b].cpu().numpy())
        _match_sets(w[b], wr, on_tol=1e-8)       # double Bloch factors (two polarizations)


def test_patterned_period_vs_oracle(gpu):
    """Disc-pixmap + uniform period, pw 3 x 3, a (kx, ky) batch: Bloch factors
    against the oracle's QZ on the same S, pencil residuals, and the
    on-shell factors come in reciprocal pairs (w, 1/w = conj w) at kp = 0."""
    from khepri_amd import eigentricks as E
    kp = np.array([[0.0, 0.0], [0.4, 0.0], [0.3, 0.7], [1.1, -0.2]])
    cl = _stack_crystal((3, 3), [("pixmap", _disc(), 0.25), ("uniform", 2.25, 0.15)], [1.7] * 4, kp)
    S = E.scattering_matrix(cl).cpu().numpy()
    w, v = E.scattering_eigenvalues(S)
    for b in range(4):
        wr, _ = OE.scattering_eigenvalues(S[b])
        _match_sets(w[b], wr, on_tol=1e-9, rel=1e-8)
        assert _residual(S[b], w[b], v[b]) < 1e-11
    on = w[0][E.on_shell(w[0], tol=1e-8)]
    assert len(on) > 0
    _match_sets(1 / on, on, on_tol=1e-8)


def test_band_structure_slab(gpu):
    from khepri_amd import eigentricks as E
    cl = _stack_crystal((1, 1), [("uniform", 1.0, 1 + 1e-6)], [1.0], None)
    S = E.scattering_matrix(cl)
    kept = E.band_structure(S[0])
    assert len(kept) == 4 and np.allclose(kept, np.cos(2 * np.pi * 1e-6), atol=1e-13, rtol=0)
    ref = OE.band_structure(S[0].cpu().numpy())
    assert np.allclose(np.sort(kept), np.sort(ref), atol=1e-14, rtol=0)
    cl2 = _stack_crystal((1, 1), [("uniform", 1.0, 1 + 1e-4)], [1.0], None)
    assert len(E.band_structure(E.scattering_matrix(cl2)[0])) == 0


def test_det_vs_oracle(gpu):
    from khepri_amd import eigentricks as E
    S = _rand_S(np.random.default_rng(7), 5, 4)
    d = E.scattering_det(S)
    for b in range(5):
        r = OE.scattering_det(S[b])
        assert abs(d[b] - r) <= 1e-10 * abs(r)
    d0 = E.scattering_det(S[0])
    assert isinstance(d0, complex) and abs(d0 - d[0]) <= 1e-13 * abs(d[0])
    w, v, dd = E.scattering_eigenvalues(S[1], dos=True)
    assert abs(dd - d[1]) <= 1e-13 * abs(d[1])
    # singular backward transmission: numpy's inv raises LinAlgError
    Ss = S[:2].copy()
    Ss[1, 8:, 8:] = 0
    with pytest.raises(np.linalg.LinAlgError):
        E.scattering_det(Ss)


def test_nan_items(gpu):
    from khepri_amd import eigentricks as E
    S = _rand_S(np.random.default_rng(9), 3, 2)
    S[1, 3, 5] = np.nan
    assert E.scattering_eigenvalues(S[1]) is None
    assert E.band_structure(S[1]) is None
    w, v = E.scattering_eigenvalues(S)
    assert np.isnan(w[1]).all() and np.isnan(v[1]).all()
    for b in (0, 2):
        wr, _ = OE.scattering_eigenvalues(S[b])
        _match_sets(w[b], wr)
    res = E.band_structure(S)
    assert res[1] is None and res[0] is not None


def test_bloch_modes_81_harmonics(gpu):
    """pw 9 x 9 (4ng = 324, the size of the 81-harmonic solves) on a disc
    period: pencil residuals of every item, one item against the oracle."""
    from khepri_amd import eigentricks as E
    kp = np.array([[0.0, 0.0], [0.5, 0.1], [0.2, 1.3], [2.0, 0.4]])
    cl = _stack_crystal((9, 9), [("pixmap", _disc(), 0.2), ("uniform", 1.0, 0.1)], [1.3] * 4, kp)
    S = E.scattering_matrix(cl)
    assert S.shape == (4, 324, 324)
    w, v = E.scattering_eigenvalues(S)
    Sn, wn, vn = S.cpu().numpy(), w.cpu().numpy(), v.cpu().numpy()
    for b in range(4):
        assert _residual(Sn[b], wn[b], vn[b]) < 1e-11
    wr, _ = OE.scattering_eigenvalues(Sn[1])
    on = np.abs(np.abs(wr) - 1) < 1e-6
    _match_sets(wn[1][np.abs(np.abs(wn[1]) - 1) < 1e-6], wr[on], on_tol=1e-9)

"""Edge cases of the solve path on the GPU, each against what the oracle (the
reference's algorithm on the CPU) does with the same input:

* a pixmap layer of eps = 0 everywhere: C is singular, numpy's inv raises
  LinAlgError (khepri/tools.py:33-35 via khepri/alternative.py:71);
* a uniform layer of eps = 0 (Python float): the reference's P = (1/er) [...]
  raises ZeroDivisionError (khepri/alternative.py:138);
* te = tm = NaN: every R and T is NaN (no error);
* |k_par| = k0 exactly (grazing incidence): the incident medium's kz is 0,
  R and T are non-finite for that source, finite for the rest of the batch;
* a pixmap smaller than the harmonic span: IndexError;
* an eigensolver that does not converge (iteration cap lowered to 2 for the
  test through KHX_EIG_ITMAX, read at every launch): LinAlgError, and the next
  solve without the cap is correct again.
"""
import numpy as np
import pytest

from oracle import crystal as OC
from khepri_amd import workloads as W

pytestmark = pytest.mark.gpu
PIX = W.disc_pixmap(32, 4.0, 0.3, 1.0)


def _pair(pix=PIX, uni=1.0, pw=(3, 3)):
    """The same two-layer device as a khepri_amd and an oracle Crystal."""
    from khepri_amd import Crystal
    out = []
    for cls in (Crystal, OC.Crystal):
        c = cls(pw)
        c.add_layer_pixmap("P", pix, 0.4)
        c.add_layer_uniform("U", uni, 0.5)
        c.set_device(["P", "U"], False)
        out.append(c)
    return out


def _oracle_rt(c):
    with np.errstate(all="ignore"):
        c.solve()
        return c.poynting_flux_end()


def test_zero_permittivity_pixmap_raises(gpu):
    g, o = _pair(pix=np.zeros((32, 32)))
    for c in (g, o):
        c.set_source(1.6, 1.0, 0.0)
    with pytest.raises(np.linalg.LinAlgError):
        _oracle_rt(o)
    with pytest.raises(np.linalg.LinAlgError):
        g.solve()


@pytest.mark.parametrize("rt_only", [False, True])
def test_zero_permittivity_uniform_raises(gpu, rt_only):
    g, o = _pair(uni=0.0)
    for c in (g, o):
        c.set_source(1.6, 1.0, 0.0)
    with pytest.raises(ZeroDivisionError):
        _oracle_rt(o)
    with pytest.raises(ZeroDivisionError):
        g.solve(rt_only=rt_only)


@pytest.mark.parametrize("rt_only", [False, True])
def test_nan_polarisation_gives_nan(gpu, rt_only):
    g, o = _pair()
    for c in (g, o):
        c.set_source(1.6, np.nan, np.nan)
    ro, to = _oracle_rt(o)
    assert np.isnan(ro) and np.isnan(to)
    g.solve(rt_only=rt_only)
    r, t = g.poynting_flux_end()
    assert np.isnan(r) and np.isnan(t)


@pytest.mark.parametrize("rt_only", [False, True])
def test_grazing_incidence_pattern(gpu, rt_only):
    """A batch with |k_par| = k0 at items 1 (along x) and 3 (along y) between
    ordinary sources: the non-finite pattern and the finite values match the
    oracle source by source."""
    wl = 1.6
    k0 = 2 * np.pi / wl
    kp = np.array([[0.3, 0.1], [k0, 0.0], [0.0, 0.0], [0.0, k0], [0.5 * k0, -0.4 * k0]])
    g, o = _pair()
    g.set_source(np.full(5, wl), 1.0, 0.5, kp=kp)
    g.solve(rt_only=rt_only)
    R, T = g.poynting_flux_end()
    for b in range(5):
        _, oc = _pair()
        oc.set_source(wl, 1.0, 0.5, kp=tuple(kp[b]))
        ro, to = _oracle_rt(oc)
        assert np.isfinite(R[b]) == np.isfinite(ro) and np.isfinite(T[b]) == np.isfinite(to), b
        if np.isfinite(ro):
            assert abs(R[b] - ro) <= 1e-9 and abs(T[b] - to) <= 1e-9, b
    assert not np.isfinite(R[[1, 3]]).any() and np.isfinite(R[[0, 2, 4]]).all()


def test_pixmap_too_small_raises(gpu):
    g, o = _pair(pix=PIX[:4, :4].copy(), pw=(5, 5))
    for c in (g, o):
        c.set_source(1.6, 1.0, 0.0)
    with pytest.raises(IndexError):
        _oracle_rt(o)
    with pytest.raises(IndexError):
        g.solve()


@pytest.mark.parametrize("pw", [(3, 3), (7, 7)])
def test_eigensolver_nonconvergence_raises(gpu, monkeypatch, pw):
    """n = 18 runs the one-launch eigensolver, n = 98 the three-launch one."""
    g, o = _pair(pw=pw)
    for c in (g, o):
        c.set_source(1.6, 1.0, 0.0)
    monkeypatch.setenv("KHX_EIG_ITMAX", "2")
    with pytest.raises(np.linalg.LinAlgError):
        g.solve()
    monkeypatch.delenv("KHX_EIG_ITMAX")
    g.solve()
    r, t = g.poynting_flux_end()
    ro, to = _oracle_rt(o)
    assert abs(r - ro) <= 1e-9 and abs(t - to) <= 1e-9

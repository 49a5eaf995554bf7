"""Closed-form pins of the band-structure oracle (oracle/eigentricks.py,
restating khepri/eigentricks.py:5-55)."""
import cmath

import numpy as np
import pytest

from oracle import crystal as OC
from oracle import eigentricks as OE


def _stack_S(pw, layers, wl, kp=None):
    """Band-structure S of a stack of uniform layers (eps, depth) between eps = 1 gaps."""
    cl = OC.Crystal(pw)
    names = []
    for i, (eps, d) in enumerate(layers):
        cl.add_layer_uniform(f"L{i}", eps, d)
        names.append(f"L{i}")
    cl.set_device(names)
    if kp is None:
        cl.set_source(wl, 1.0, 1.0, 0.0, 0.0)
    else:
        cl.set_source(wl, 1.0, 1.0, kp=kp)
    cl.solve()
    return OE.from_blocks(cl.Stot), cl


def _match(a, b):
    """max over a of the distance to the nearest element of b (and back)."""
    a, b = np.asarray(a), np.asarray(b)
    d = np.abs(a[:, None] - b[None, :])
    return max(d.min(1).max(), d.min(0).max())


@pytest.mark.parametrize("kp", [(0.0, 0.0), (0.31, -0.17)])
def test_homogeneous_slab_bloch_factors(kp):
    """A vacuum slab of depth d (pw 3 x 3): the Bloch factors are exp(+- i kz d)
    for every harmonic and both polarizations, evanescent ones included."""
    wl, d = 1.3, 0.37
    S, cl = _stack_S((3, 3), [(1.0, d)], wl, kp=kp)
    w, v = OE.scattering_eigenvalues(S)
    k0 = 2 * np.pi / wl
    kx = kp[0] + cl.expansion.g[0].ravel()
    ky = kp[1] + cl.expansion.g[1].ravel()
    kz = np.array([cmath.sqrt(k0 ** 2 - a ** 2 - b ** 2) for a, b in zip(kx, ky)])
    expect = np.concatenate([np.exp(1j * kz * d)] * 2 + [np.exp(-1j * kz * d)] * 2)
    assert np.abs(expect).max() > 10          # evanescent harmonics are present
    assert _match(w / np.abs(w) ** 0.5, expect / np.abs(expect) ** 0.5) < 1e-9
    # unit-norm eigenvectors of the pencil
    Sl, Sr = OE.scattering_splitlr(S)
    assert np.allclose(np.linalg.norm(v, axis=0), 1.0, atol=1e-12)
    res = np.linalg.norm(Sl @ v - (Sr @ v) * w[None, :], axis=0) / (1 + np.abs(w))
    assert res.max() < 1e-10


@pytest.mark.parametrize("wl", [4.1, 2.3, 1.7])
def test_bragg_period_dispersion(wl):
    """Two-layer period at normal incidence (pw 1 x 1): every Bloch factor w obeys
    w + 1/w = 2 cos(k1 d1) cos(k2 d2) - (n1/n2 + n2/n1) sin(k1 d1) sin(k2 d2)."""
    e1, d1, e2, d2 = 12.0, 0.2, 1.0, 0.8
    S, _ = _stack_S((1, 1), [(e1, d1), (e2, d2)], wl)
    w, _ = OE.scattering_eigenvalues(S)
    n1, n2 = np.sqrt(e1), np.sqrt(e2)
    k1, k2 = 2 * np.pi * n1 / wl, 2 * np.pi * n2 / wl
    rhs = 2 * np.cos(k1 * d1) * np.cos(k2 * d2) - (n1 / n2 + n2 / n1) * np.sin(k1 * d1) * np.sin(k2 * d2)
    assert len(w) == 4
    assert np.allclose(w + 1 / w, rhs, atol=1e-10, rtol=0)
    on = OE.on_shell(w)
    assert on.all() == (abs(rhs) <= 2)        # propagating Bloch modes exactly inside a band


def test_band_structure_keeps_near_unit_real_parts():
    """band_structure keeps Re w in (0, 1) with |Re w| within 1e-10 of 1: a vacuum
    slab of depth (1 + 1e-6) wl gives w = exp(+- 2 pi i 1e-6), cos = 1 - 2e-11."""
    wl = 1.0
    S, _ = _stack_S((1, 1), [(1.0, wl * (1 + 1e-6))], wl)
    kept = OE.band_structure(S)
    assert len(kept) == 4
    assert np.allclose(kept, np.cos(2 * np.pi * 1e-6), atol=1e-13, rtol=0)
    S2, _ = _stack_S((1, 1), [(1.0, wl * (1 + 1e-4))], wl)
    assert len(OE.band_structure(S2)) == 0


def test_det_and_eigen_identities_random():
    """Schur-complement determinant = det(S); prod w = det(Spp)/det(Smm);
    sum w = trace(Sr^-1 Sl)."""
    rng = np.random.default_rng(5)
    ng = 3
    S = rng.standard_normal((4 * ng, 4 * ng)) + 1j * rng.standard_normal((4 * ng, 4 * ng))
    S[2 * ng:, 2 * ng:] += 4 * np.eye(2 * ng)
    d = OE.scattering_det(S)
    assert abs(d - np.linalg.det(S)) < 1e-11 * abs(d)
    w, v = OE.scattering_eigenvalues(S)
    Sl, Sr = OE.scattering_splitlr(S)
    p, m = slice(0, 2 * ng), slice(2 * ng, 4 * ng)
    pw_ = np.prod(w)
    ref = np.linalg.det(S[p, p]) / np.linalg.det(S[m, m])
    assert abs(pw_ - ref) < 1e-9 * abs(ref)
    assert abs(w.sum() - np.trace(np.linalg.solve(Sr, Sl))) < 1e-9 * np.abs(w).sum()
    w3, v3, d3 = OE.scattering_eigenvalues(S, dos=True)
    assert d3 == d and np.array_equal(w3, w)


def test_nan_gives_none():
    S = np.eye(8, dtype=complex)
    S[1, 6] = np.nan
    assert OE.scattering_eigenvalues(S) is None
    assert OE.band_structure(S) is None

"""CPU checks of the FP64 autograd reference (oracle/autograd.py): its forward
equals the numpy oracle, and its gradient equals central finite differences
of the numpy oracle."""
import numpy as np
import pytest

from oracle import autograd as AG
from oracle import crystal as OC


def _device(seed, pw, shape=(16, 16), nl=2):
    rng = np.random.default_rng(seed)
    pix = [rng.uniform(2.0, 12.0, size=shape) for _ in range(nl)]
    depths = list(rng.uniform(0.2, 0.5, size=nl))
    return pix, depths


def _oracle_T(pix, depths, pw, wl, theta=0.0, phi=0.0, gap=None):
    cl = OC.Crystal(pw)
    names = []
    for j, (p, d) in enumerate(zip(pix, depths)):
        cl.add_layer_pixmap(f"L{j}", p, d)
        names.append(f"L{j}")
        if gap is not None and j == 0:
            cl.add_layer_uniform("G", gap[0], gap[1])
            names.append("G")
    cl.set_device(names, False)
    cl.set_source(wl, 1.0, 0.0, theta, phi)
    cl.solve()
    return cl.poynting_flux_end()


def test_autograd_forward_matches_oracle():
    import torch
    pw = (3, 3)
    pix, depths = _device(1, pw)
    for wl, th, ph in [(1 / 0.55, 0.0, 0.0), (1 / 0.61, 17.0, 40.0)]:
        R0, T0 = _oracle_T(pix, depths, pw, wl, th, ph, gap=(2.25, 0.3))
        layers = [("pixmap", torch.tensor(pix[0]), depths[0]), ("uniform", 2.25, 0.3),
                  ("pixmap", torch.tensor(pix[1]), depths[1])]
        R, T = AG.fluxes(layers, pw, wl, theta=th, phi=ph)
        assert abs(float(R) - R0) < 1e-12 and abs(float(T) - T0) < 1e-12


@pytest.mark.parametrize("theta,phi", [(0.0, 0.0), (12.0, 25.0)])
def test_autograd_gradient_matches_finite_differences(theta, phi):
    import torch
    pw = (3, 3)
    pix, depths = _device(2, pw)
    wl = 1 / 0.57
    ts = [torch.tensor(p, requires_grad=True) for p in pix]
    _, T = AG.fluxes([("pixmap", t, d) for t, d in zip(ts, depths)], pw, wl, theta=theta, phi=phi)
    T.backward()
    grads = [t.grad.numpy() for t in ts]
    rng = np.random.default_rng(7)
    h = 1e-5
    for _ in range(6):
        li = int(rng.integers(0, 2))
        x, y = (int(v) for v in rng.integers(0, 16, size=2))
        vals = []
        for s in (+1, -1):
            pp = [p.copy() for p in pix]
            pp[li][x, y] += s * h
            vals.append(_oracle_T(pp, depths, pw, wl, theta, phi)[1])
        fd = (vals[0] - vals[1]) / (2 * h)
        scale = max(np.abs(grads[li]).max(), 1e-12)
        assert abs(fd - grads[li][x, y]) < 1e-6 * scale, (li, x, y, fd, grads[li][x, y])

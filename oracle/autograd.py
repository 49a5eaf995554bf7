"""FP64 CPU autograd reference for d(R, T)/d(pixel eps) -- TEST INFRASTRUCTURE ONLY.

The reference solver has no gradient.  This module restates its forward path
(same op sequence as oracle.rcwa, which follows khepri/tools.py:33-56,
khepri/alternative.py:19-30, 84-99, 130-195, 235-245 and khepri/layer.py:35-60)
in torch complex128 on the CPU, so that torch.autograd supplies the reference
gradient, including torch.linalg.eig's own eigenvector backward.  Only the
inputs that do not depend on eps (k vectors, free-space modes, incident
vector) come from oracle.rcwa.

Pinning: the forward agrees with oracle.crystal (which is pinned by the golden
vectors) and the gradient agrees with central finite differences of
oracle.crystal (tests/test_oracle.py).

Scope: devices of pixmap and uniform layers between eps = 1 half spaces
(identity half-space S-matrices), any incidence.
"""
from __future__ import annotations

from math import floor, prod

import numpy as np
import torch

from . import rcwa

CDT = torch.complex128


def _c(a):
    return torch.as_tensor(np.asarray(a, dtype=np.complex128))


def convolution_matrix(eps, pw):
    """fftshift(fft2(eps)) / (Nx Ny) gathered into the block-Toeplitz matrix
    (khepri/tools.py:33-56)."""
    P, Q = pw
    Nx, Ny = eps.shape
    F = torch.fft.fftshift(torch.fft.fft2(eps.to(CDT))) / prod(eps.shape)
    g0x, g0y = floor(Nx / 2), floor(Ny / 2)
    N = P * Q
    p = np.arange(N) % P
    q = np.arange(N) // P
    ix = torch.as_tensor(g0x + p[:, None] - p[None, :])
    iy = torch.as_tensor(g0y + q[:, None] - q[None, :])
    return F[ix, iy]


def patterned_layer_S(eps, depth, kx, ky, k0, pw):
    """Modes and S-matrix of a pixmap layer (khepri/alternative.py:158-195)."""
    C = convolution_matrix(eps, pw)
    N = C.shape[0]
    KX, KY = torch.diag(_c(kx)), torch.diag(_c(ky))
    I = torch.eye(N, dtype=CDT)
    CiKy = torch.linalg.solve(C, KY)
    CiKx = torch.linalg.solve(C, KX)
    Pm = torch.cat([torch.cat([KX @ CiKy, I - KX @ CiKx], 1), torch.cat([KY @ CiKy - I, -KY @ CiKx], 1)], 0)
    Qm = torch.cat([torch.cat([KX @ KY, C - KX @ KX], 1), torch.cat([KY @ KY - C, -KY @ KX], 1)], 0)
    mu, W = torch.linalg.eig(Pm @ Qm)
    lam = torch.sqrt(mu)
    V = (Qm @ W) / lam[None, :]
    W0, V0 = rcwa.free_space_eigenmodes(kx, ky)
    return build_scatmat(W, V, _c(W0), _c(V0), lam, depth, k0)


def uniform_layer_S(eps, depth, kx, ky, k0):
    """khepri/alternative.py:130-156 + 181-195 (no eps dependence tracked)."""
    W, V, lam = rcwa.uniform_layer_modes(kx, ky, eps)
    W0, V0 = rcwa.free_space_eigenmodes(kx, ky)
    return build_scatmat(_c(W), _c(V), _c(W0), _c(V0), _c(lam), depth, k0)


def build_scatmat(W, V, W0, V0, lam, depth, k0):
    """khepri/alternative.py:181-195."""
    t1 = torch.linalg.solve(W, W0)
    t2 = torch.linalg.solve(V, V0)
    A, B = t1 + t2, t1 - t2
    X = torch.diag(torch.exp(-lam * depth * k0))
    T = A - X @ B @ torch.linalg.solve(A, X) @ B
    S11 = torch.linalg.solve(T, X @ B @ torch.linalg.solve(A, X) @ A - B)
    S12 = torch.linalg.solve(T, X @ (A - B @ torch.linalg.solve(A, B)))
    return [[S11, S12], [S12, S11]]


def redheffer(SA, SB):
    """khepri/alternative.py:19-30."""
    n = SA[0][0].shape[0]
    I = torch.eye(n, dtype=CDT)
    D = I - SB[0][0] @ SA[1][1]
    F = I - SA[1][1] @ SB[0][0]
    S11 = SA[0][0] + SA[0][1] @ torch.linalg.solve(D, SB[0][0]) @ SA[1][0]
    S12 = SA[0][1] @ torch.linalg.solve(D, SB[0][1])
    S21 = SB[1][0] @ torch.linalg.solve(F, SA[1][0])
    S22 = SB[1][1] + SB[1][0] @ torch.linalg.solve(F, SA[1][1]) @ SB[0][1]
    return [[S11, S12], [S21, S22]]


def _flux(g, s, kp, wl, kzi_n, eps):
    """khepri/alternative.py:235-245 (sum over orders)."""
    kx, ky, kz = (_c(v) for v in rcwa.k_vectors(g, kp, wl, eps))
    N = len(kx)
    sx, sy = s[:N], s[N:]
    sz = -(kx * sx + ky * sy) / kz
    t = kz.real / kzi_n.real * (sx.abs() ** 2 + sy.abs() ** 2 + sz.abs() ** 2)
    return t.sum()


def fluxes(layers, pw, wavelength, te=1.0, tm=0.0, theta=0.0, phi=0.0, lattice="square"):
    """(R, T) torch scalars for layers = [("pixmap", eps_tensor, depth) |
    ("uniform", eps, depth)] between eps = 1 half spaces
    (khepri/crystal.py:131-206, 345-396)."""
    g = rcwa.g_vectors(pw, rcwa.lattice_matrix(lattice))
    kp = rcwa.compute_kplanar(1.0, wavelength, theta, phi)
    k0 = 2 * np.pi / wavelength
    kx, ky, _ = rcwa.k_vectors(g, kp, wavelength)
    S = None
    for kind, eps, depth in layers:
        if kind == "pixmap":
            Sl = patterned_layer_S(eps, depth, kx, ky, k0, pw)
        elif kind == "uniform":
            Sl = uniform_layer_S(eps, depth, kx, ky, k0)
        else:
            raise ValueError(kind)
        S = Sl if S is None else redheffer(S, Sl)
    kzi = np.conj(np.sqrt(complex(k0 ** 2 - kp[0] ** 2 - kp[1] ** 2)))
    inc = _c(rcwa.incident(pw, te, tm, (kp[0], kp[1], kzi)))
    kzi_n = _c(kzi / k0)
    R = _flux(g, S[0][0] @ inc, kp, wavelength, kzi_n, 1.0)
    T = _flux(g, S[1][0] @ inc, kp, wavelength, kzi_n, 1.0)
    return R, T


def transmission_gradient(pixmaps, depths, pw, wavelength, **kw):
    """T and dT/d eps for a stack of pixmap layers (list of (nx, ny) arrays)."""
    ts = [torch.tensor(np.asarray(p, dtype=np.float64), requires_grad=True) for p in pixmaps]
    _, T = fluxes([("pixmap", t, d) for t, d in zip(ts, depths)], pw, wavelength, **kw)
    T.backward()
    return float(T.detach()), [t.grad.numpy().copy() for t in ts]

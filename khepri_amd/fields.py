"""Field reconstruction in real space (mirrors khepri/crystal.py:208-343 and
khepri/fields.py).

For a layer, the mode amplitudes a = RI^-1 R0 [c+; c-] do not depend on z,
so every z-slice falling in that layer is one column of a single GEMM
RI [L(z_1) a, L(z_2) a, ...]; the inverse DFT to the (x, y) points is one
more GEMM against the phase matrix exp(i (kx x + ky y)).  All arithmetic runs
in libkhx; this module only groups the z values by layer.
"""
from __future__ import annotations

import numpy as np
import torch

from . import ops
from .layer import Layer


def locate_layer(self, z):
    """Layer containing depth z and the depth relative to it (khepri/crystal.py:208-232)."""
    assert z != np.nan
    layer_index = int(np.searchsorted(self.stack_positions, z)) - 1
    layer_name = self.global_stacking[layer_index]
    zr = z if z <= 0 else z - self.stack_positions[layer_index]
    return self.layers[layer_name], layer_index, zr


def _require_single_source(self):
    assert self.solved and self.S is not None, "Call solve first (fields need the resident S-matrix chain)."
    if self.batch != 1:
        raise ValueError("fields are computed for one source at a time (set_source with scalars)")
    if self.stacking_reverse_matrices is None:
        raise ValueError("fields need set_device(..., fields_mask=True) before solve()")


def get_source_as_field_vectors(self):
    """(E, H) Fourier vectors of the incident plane wave (khepri/crystal.py:281-295)."""
    kp, wl, te, tm = self._sources_dev
    e = ops.incident(self.expansion.N, kp, wl, te, tm, self.epsi)[0].cpu().numpy()
    return e, np.zeros_like(e)


def _free_space(self, kx, ky):
    """W0 = I, V0 of the gap medium (khepri/alternative.py:84-99), (1, n, n)."""
    k0d = torch.zeros(kx.shape[0], dtype=torch.float64, device=kx.device)
    _, (W0, V0, _) = ops.analytic_layer("uniform", kx, ky, 1.0, k0d, modes=True)
    return W0, V0


def _fourier_fields_many(self, zs, inc, kx, ky):
    """[sx, sy, sz, ux, uy, uz] Fourier fields, each (N, len(zs)), in z order."""
    dev = self._dev()
    N = self.expansion.N
    n = 2 * N
    k0 = 2 * np.pi / float(self._wl[0])
    inc = torch.as_tensor(np.asarray(inc, dtype=np.complex128), device=dev).reshape(1, 2 * n, 1)
    Sref = self.layers["Sref"]
    Rref = ops.eigenbasis(Sref.W, Sref.V)
    sol, info = ops.solve(Rref, inc)
    ops.raise_on_info(info, "incident-medium eigenbasis", lu=True)
    c1p = sol[:, :n, :]                                       # (1, n, 1)
    groups = {}
    for j, z in enumerate(zs):
        layer, li, zr = self.locate_layer(z)
        assert layer.fields, f"Layer at {z} did not store eigenspace."
        groups.setdefault(li, []).append((j, zr))
    out = torch.empty((6, N, len(zs)), dtype=ops.CDT, device=dev)
    W0V0 = None
    for li, items in groups.items():
        layer = self.layers[self.global_stacking[li]]
        Sl = self.stacking_matrices[li]
        Sr = self.stacking_reverse_matrices[li]
        # translate_mode_amplitudes (khepri/fields.py:18-27)
        F = ops.identity_minus(ops.matmul(Sl[:, 1, 1], Sr[:, 0, 0]))
        cdp, info = ops.solve(F, ops.matmul(Sl[:, 1, 0], c1p))
        ops.raise_on_info(info, "mode amplitudes", lu=True)
        cdm = ops.matmul(Sr[:, 0, 0], cdp)
        RI = ops.eigenbasis(layer.W, layer.V)
        if isinstance(layer, Layer):
            if W0V0 is None:
                W0V0 = _free_space(self, kx, ky)
            W0, V0 = W0V0
        else:
            W0, V0 = layer.W0, layer.V0
        R0 = ops.eigenbasis(W0, V0)
        amp = ops.matmul(R0, torch.cat([cdp, cdm], dim=1))
        a, info = ops.solve(RI, amp)                          # (1, 2n, 1)
        ops.raise_on_info(info, "layer eigenbasis", lu=True)
        cols = [j for j, _ in items]
        zbar = torch.as_tensor(np.array([k0 * (layer.depth - zr) for _, zr in items], dtype=np.float64),
                               device=dev)
        M = ops.mode_propagate(layer.L[0], a[0, :, 0], zbar)  # (2n, nz)
        Fz = ops.matmul(RI, M[None])[0]                       # [sx; sy; ux; uy]
        uz, w = ops.longitudinal(kx[0], ky[0], Fz)
        IC = layer.IC
        if isinstance(IC, torch.Tensor):
            sz = ops.matmul(IC, w[None], alpha=-1j)[0]
        else:
            coef = torch.full((1, w.shape[0]), complex(-1j * IC), dtype=ops.CDT, device=dev)
            sz = ops.diag_scale(w[None], coef, None)[0]
        idx = torch.as_tensor(cols, device=dev)
        comps = [Fz[0:N], Fz[N:2 * N], sz, Fz[2 * N:3 * N], Fz[3 * N:4 * N], uz]
        for c, v in enumerate(comps):
            out[c][:, idx] = v
    return out


def _kvecs(self, kp):
    dev = self._dev()
    kpt = torch.as_tensor(np.asarray(kp, dtype=np.complex128).reshape(1, 2), device=dev)
    wl = torch.as_tensor(self._wl[:1], device=dev)
    return ops.kvectors(self.expansion.g_device(dev), kpt, wl)


def fourier_fields(self, z, incident_fields):
    """Fourier fields at depth z (khepri/crystal.py:234-279)."""
    _require_single_source(self)
    kx, ky = _kvecs(self, self._kp[0])
    out = _fourier_fields_many(self, [z], incident_fields, kx, ky)
    return tuple(out[c, :, 0].cpu().numpy() for c in range(6))


def _real_space(self, four, x, y, kp):
    """6 x (N, nz) Fourier fields -> (nz, 6, *x.shape) real-space fields."""
    dev = self._dev()
    k0 = 2 * np.pi / float(self._wl[0])
    kx, ky = _kvecs(self, kp)
    xs = torch.as_tensor(np.ascontiguousarray(x, dtype=np.float64).ravel(), device=dev)
    ys = torch.as_tensor(np.ascontiguousarray(y, dtype=np.float64).ravel(), device=dev)
    E = ops.idft_phase(kx[0], ky[0], k0, xs, ys)           # (npts, N), absolute k = k0 K
    real = ops.matmul(E, four)                              # (6, npts, nz)
    arr = real.cpu().numpy().transpose(2, 0, 1)             # (nz, 6, npts)
    return arr.reshape((arr.shape[0], 6) + tuple(np.shape(x)))


def fields_coords_xy(self, x, y, z, incident_fields=None, kp=None, return_fourier=False):
    """E, H at the points (x, y) of depth z (khepri/crystal.py:297-333)."""
    x = np.array(x)
    y = np.array(y)
    assert x.shape == y.shape and (len(x.shape) == 2), "x and y must be 2D meshgrids"
    _require_single_source(self)
    if kp is None:
        kp = self._kp[0]
    if incident_fields is None:
        incident_fields = np.hstack(self.get_source_as_field_vectors())
    elif isinstance(incident_fields, tuple) and len(incident_fields) == 2:
        incident_fields = np.hstack(incident_fields)
    if isinstance(z, str):
        raise NotImplementedError("far-field evaluation is not provided by the reference either")
    kx, ky = _kvecs(self, self._kp[0])
    four = _fourier_fields_many(self, [z], incident_fields, kx, ky)
    if return_fourier:
        return tuple(four[c, :, 0].cpu().numpy() for c in range(6))
    f = _real_space(self, four, x, y, kp)[0]
    return np.split(f, 2, axis=0)


def fields_volume(self, x, y, z, incident_fields=None):
    """E, H on the (x, y) grid for every depth in z: (Nz, 3, *x.shape) each
    (khepri/crystal.py:335-343)."""
    _require_single_source(self)
    if incident_fields is None:
        incident_fields = self.get_source_as_field_vectors()
    inc = np.hstack(incident_fields)
    x = np.array(x)
    y = np.array(y)
    zs = [float(v) for v in np.atleast_1d(z)]
    kx, ky = _kvecs(self, self._kp[0])
    four = _fourier_fields_many(self, zs, inc, kx, ky)
    f = _real_space(self, four, x, y, self._kp[0])
    return f[:, :3], f[:, 3:]

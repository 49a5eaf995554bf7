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


def _require_fields(self):
    assert self.solved and self.S is not None, "Call solve first (fields need the resident S-matrix chain)."
    if self.stacking_reverse_matrices is None:
        raise ValueError("fields need set_device(..., fields_mask=True) before solve()")


def get_source_as_field_vectors(self):
    """(E, H) Fourier vectors of the incident plane wave (khepri/crystal.py:281-295);
    (B, 2N) each for a batch of sources."""
    kp, wl, te, tm = self._sources_dev
    e = ops.incident(self.expansion.N, kp, wl, te, tm, self.epsi).cpu().numpy()
    if self._scalar:
        e = e[0]
    return e, np.zeros_like(e)


def _incident_tensor(self, incident_fields):
    """The incident field supervector [E; H] of every source, (B, 2n) on the
    device: from the sources (None), an (E, H) pair, or an array / tensor of
    shape (2n,) (every source) or (B, 2n)."""
    dev = self._dev()
    n = 2 * self.expansion.N
    B = self.batch
    if incident_fields is None:
        kp, wl, te, tm = self._sources_dev
        e = ops.incident(self.expansion.N, kp, wl, te, tm, self.epsi)
        return torch.cat([e, torch.zeros_like(e)], dim=1)
    if isinstance(incident_fields, tuple) and len(incident_fields) == 2:
        incident_fields = np.concatenate([np.asarray(v) for v in incident_fields], axis=-1)
    inc = torch.as_tensor(incident_fields, device=dev).to(ops.CDT)
    if inc.shape[-1] != 2 * n or inc.dim() > 2:
        raise ValueError(f"incident fields must have shape (2n,) or (batch, 2n) with 2n = {2 * n}, got "
                         f"{tuple(inc.shape)}")
    if inc.dim() == 2 and inc.shape[0] not in (1, B):
        raise ValueError(f"incident fields for {inc.shape[0]} sources, the crystal has {B}")
    return inc.reshape(-1, 2 * n).expand(B, 2 * n).contiguous()


def _free_space(self, kx, ky):
    """W0 = I, V0 of the gap medium (khepri/alternative.py:84-99), (B, n, n)."""
    k0d = torch.zeros(kx.shape[0], dtype=torch.float64, device=kx.device)
    _, (W0, V0, _) = ops.analytic_layer("uniform", kx, ky, 1.0, k0d, modes=True)
    return W0, V0


def _fourier_fields_many(self, zs, inc, kx, ky):
    """[sx, sy, sz, ux, uy, uz] Fourier fields of every source at the depths
    zs: (B, N, 6, len(zs)), in z order.  inc (B, 2n) device tensor."""
    dev = self._dev()
    N = self.expansion.N
    n = 2 * N
    B = inc.shape[0]
    k0 = 2 * np.pi / self._wl
    Sref = self.layers["Sref"]
    Rref = ops.eigenbasis(Sref.W, Sref.V)
    sol, info = ops.solve(Rref, inc.reshape(B, 2 * n, 1), overwrite_a=True)   # Rref: a fresh eigenbasis
    ops.raise_on_info(info, "incident-medium eigenbasis", lu=True)
    c1p = sol[:, :n, :]                                       # (B, n, 1)
    groups = {}
    for j, z in enumerate(zs):
        layer, li, zr = self.locate_layer(z)
        assert layer.fields, f"Layer at {z} did not store eigenspace."
        groups.setdefault(li, []).append((j, zr))
    out = torch.empty((B, N, 6, len(zs)), dtype=ops.CDT, device=dev)
    W0V0 = None
    for li, items in groups.items():
        layer = self.layers[self.global_stacking[li]]
        Sl = self.stacking_matrices[li]
        Sr = self.stacking_reverse_matrices[li]
        # translate_mode_amplitudes (khepri/fields.py:18-27)
        P = ops.matmul(Sl[:, 1, 1], Sr[:, 0, 0])
        F = ops.identity_minus(P, out=P)
        cdp, info = ops.solve(F, ops.matmul(Sl[:, 1, 0], c1p), overwrite_a=True, overwrite_b=True)
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
        a, info = ops.solve(RI, amp, overwrite_b=True)        # (B, 2n, 1); RI is used again below
        ops.raise_on_info(info, "layer eigenbasis", lu=True)
        cols = [j for j, _ in items]
        if len(set(np.round(k0, 15))) == 1:                   # one wavelength: zbar shared by the batch
            zbar = torch.as_tensor(np.array([k0[0] * (layer.depth - zr) for _, zr in items], dtype=np.float64),
                                   device=dev)
            M = ops.mode_propagate_batch(layer.L, a[:, :, 0], zbar)            # (B, 2n, nz)
        else:                                                 # k0 differs per source: one item at a time
            M = torch.empty((B, 2 * n, len(items)), dtype=ops.CDT, device=dev)
            for b in range(B):
                zb = torch.as_tensor(np.array([k0[b] * (layer.depth - zr) for _, zr in items], dtype=np.float64),
                                     device=dev)
                M[b] = ops.mode_propagate_batch(layer.L[b:b + 1], a[b:b + 1, :, 0], zb)[0]
        Fz = ops.matmul(RI, M)                                # (B, 4N, nz): [sx; sy; ux; uy]
        uz, w = ops.longitudinal_batch(kx, ky, Fz)
        IC = layer.IC
        if isinstance(IC, torch.Tensor):
            sz = ops.matmul(IC, w, alpha=-1j)
        else:
            coef = torch.full((B, N), complex(-1j * IC), dtype=ops.CDT, device=dev)
            sz = ops.diag_scale(w, coef, None)
        idx = torch.as_tensor(cols, device=dev)
        comps = [Fz[:, 0:N], Fz[:, N:2 * N], sz, Fz[:, 2 * N:3 * N], Fz[:, 3 * N:4 * N], uz]
        for c, v in enumerate(comps):
            out[:, :, c, idx] = v
    return out


def _kvecs(self, kp=None):
    """Normalised (Kx, Ky) of every source, (B, N) each; kp overrides the
    sources' in-plane vectors ((2,) or (B, 2))."""
    dev = self._dev()
    kp_all, wl_all = self._sources_dev[0], self._sources_dev[1]
    if kp is not None:
        kp_all = torch.as_tensor(np.broadcast_to(np.asarray(kp, dtype=np.complex128).reshape(-1, 2),
                                                 (self.batch, 2)).copy(), device=dev)
    return ops.kvectors(self.expansion.g_device(dev), kp_all, wl_all)


def fourier_fields(self, z, incident_fields):
    """Fourier fields at depth z (khepri/crystal.py:234-279)."""
    _require_fields(self)
    kx, ky = _kvecs(self)
    out = _fourier_fields_many(self, [z], _incident_tensor(self, incident_fields), kx, ky)
    comps = tuple(out[:, :, c, 0].cpu().numpy() for c in range(6))
    return tuple(v[0] for v in comps) if self._scalar else comps


def _real_space(self, four, x, y, kp=None):
    """(B, N, 6, nz) Fourier fields -> (B, nz, 6, *x.shape) real-space fields."""
    dev = self._dev()
    B, N, _, nz = four.shape
    kx, ky = _kvecs(self, kp)
    xs = torch.as_tensor(np.ascontiguousarray(x, dtype=np.float64).ravel(), device=dev)
    ys = torch.as_tensor(np.ascontiguousarray(y, dtype=np.float64).ravel(), device=dev)
    E = ops.idft_phase_batch(kx, ky, self._sources_dev[1], xs, ys)   # (B, npts, N), absolute k = k0 K
    real = ops.matmul(E, four.reshape(B, N, 6 * nz))                 # (B, npts, 6 nz)
    arr = real.reshape(B, -1, 6, nz).permute(0, 3, 2, 1).cpu().numpy()  # (B, nz, 6, npts)
    return arr.reshape((B, nz, 6) + tuple(np.shape(x)))


def fields_coords_xy(self, x, y, z, incident_fields=None, kp=None, return_fourier=False):
    """E, H at the points (x, y) of depth z (khepri/crystal.py:297-333).  For
    a batch of sources every output gains a leading source axis; incident
    fields may be one (2n,) vector or one per source, (B, 2n)."""
    x = np.array(x)
    y = np.array(y)
    assert x.shape == y.shape and (len(x.shape) == 2), "x and y must be 2D meshgrids"
    _require_fields(self)
    if isinstance(z, str):
        raise NotImplementedError("far-field evaluation is not provided by the reference either")
    kx, ky = _kvecs(self)
    four = _fourier_fields_many(self, [z], _incident_tensor(self, incident_fields), kx, ky)
    if return_fourier:
        comps = tuple(four[:, :, c, 0].cpu().numpy() for c in range(6))
        return tuple(v[0] for v in comps) if self._scalar else comps
    f = _real_space(self, four, x, y, kp)[:, 0]                      # (B, 6, *x.shape)
    if self._scalar:
        return np.split(f[0], 2, axis=0)
    return np.split(f, 2, axis=1)


def fields_volume(self, x, y, z, incident_fields=None):
    """E, H on the (x, y) grid for every depth in z: (Nz, 3, *x.shape) each
    (khepri/crystal.py:335-343); (B, Nz, 3, *x.shape) for a batch of sources,
    all of them in one pass (incident fields (2n,) or (B, 2n))."""
    _require_fields(self)
    x = np.array(x)
    y = np.array(y)
    zs = [float(v) for v in np.atleast_1d(z)]
    kx, ky = _kvecs(self)
    four = _fourier_fields_many(self, zs, _incident_tensor(self, incident_fields), kx, ky)
    f = _real_space(self, four, x, y)
    if self._scalar:
        return f[0, :, :3], f[0, :, 3:]
    return f[:, :, :3], f[:, :, 3:]

"""Beam / Brillouin-zone integration front end (khepri/beams.py) on the GPU.

A finite beam is decomposed over Bloch waves: its fields are sampled on a
grid covering several unit cells (``shifted_rotated_fields``), their
Fourier amplitudes per Bloch vector are taken (``amplitudes_from_fields``)
and fed to ``Crystal.fields_volume(incident_fields=...)`` for every k-point
of the Brillouin-zone grid (``gen_bzi_grid``).

Field evaluation of the built-in paraxial beams (Gaussian, Laguerre-Gauss)
and the Bloch amplitudes run in libkhx (``khx_beam_fields``,
``khx_bloch_amplitudes``; ``points_dft`` binds the plain point DFT,
``khx_points_dft``);
the 3 x 3 beam rotation and the grid bookkeeping are host geometry.  A field
function written by the caller (any other callable) is evaluated by that
callable on host arrays, as in the reference, and only rotated here.
Signatures, argument meaning and output shapes follow the reference,
including the trailing length-1 axis of ``shifted_rotated_fields``.
"""
from __future__ import annotations

import functools

import numpy as np
import torch

from . import _lib
from ._lib import check
from .device import default_device


def rotation_matrix(polar_angle, azimuthal_angle, polarization_angle):
    """Beam frame -> lab frame: rotation by the polarization angle about the
    beam axis after R_z(azimuthal) R_y(polar).  khepri/beams.py:6-47."""
    cp, sp = np.cos(polar_angle), np.sin(polar_angle)
    ca, sa = np.cos(azimuthal_angle), np.sin(azimuthal_angle)
    cq, sq = np.cos(polarization_angle), np.sin(polarization_angle)
    ry = np.array([[cp, 0.0, sp], [0.0, 1.0, 0.0], [-sp, 0.0, cp]])
    rz = np.array([[ca, -sa, 0.0], [sa, ca, 0.0], [0.0, 0.0, 1.0]])
    u = np.array([ca * sp, sa * sp, cp])                  # beam axis
    cross = np.array([[0.0, -u[2], u[1]], [u[2], 0.0, -u[0]], [-u[1], u[0], 0.0]])
    rp = cq * np.eye(3) + (1 - cq) * np.outer(u, u) + sq * cross   # Rodrigues' formula
    return rp @ rz @ ry


def gen_bzi_grid(shape, a=1, reciproc=None):
    """Bloch vectors at the centres of a shape[0] x shape[1] grid of cells
    over the Brillouin zone, shape (2, *shape) (khepri/beams.py:193-207):
    fractional coordinates -1/2 + (i + 1/2)/N along each reciprocal vector
    (b1, b2; by default those of a square lattice of pitch a)."""
    basis = reciproc if reciproc is not None else ((2 * np.pi / a, 0), (0, 2 * np.pi / a))
    fractions = [np.arange(-0.5 + 0.5 / N, 0.5, 1 / N) for N in shape]
    f1, f2 = np.meshgrid(*fractions, indexing="ij")
    return np.stack([basis[0][c] * f1 + basis[1][c] * f2 for c in range(2)])


def _device():
    return default_device()


def _beam(kind, x, y, z, wl, geom, w0, er=1, p=0, l=0):
    """Run khx_beam_fields on host arrays; returns E, H as (3, *shape) numpy."""
    lib = _lib.load()
    dev = _device()
    x, y, z = np.broadcast_arrays(np.asarray(x, dtype=float), np.asarray(y, dtype=float),
                                  np.asarray(z, dtype=float))
    shape = x.shape
    xd, yd, zd = (torch.as_tensor(np.ascontiguousarray(v).ravel(), device=dev) for v in (x, y, z))
    n = xd.numel()
    g = torch.as_tensor(np.ascontiguousarray(geom, dtype=np.float64), device=dev)
    E = torch.empty((3, n), dtype=torch.complex128, device=dev)
    H = torch.empty_like(E)
    check(lib.khx_beam_fields(kind, xd.data_ptr(), yd.data_ptr(), zd.data_ptr(), n, float(wl), float(w0),
                              float(er), int(p), int(l), g.data_ptr(), E.data_ptr(), H.data_ptr(),
                              torch.cuda.current_stream().cuda_stream), "khx_beam_fields")
    return E.cpu().numpy().reshape(3, *shape), H.cpu().numpy().reshape(3, *shape)


_IDENTITY_GEOM = np.concatenate([np.eye(3).ravel(), np.eye(3).ravel(), np.zeros(3)])


def _paraxial_gaussian_field_fn(x, y, z, wl, beam_waist=1, er=1):
    """x-polarised paraxial Gaussian beam in its own frame.  khepri/beams.py:135-160."""
    E, H = _beam(0, x, y, z, wl, _IDENTITY_GEOM, beam_waist, er)
    return (E[0], E[1], E[2]), (H[0], H[1], H[2])


def _paraxial_laguerre_gaussian_field_fn(x, y, z, wl, w0=1, er=1, p=0, l=0):
    """x-polarised paraxial Laguerre-Gauss beam LG_pl in its own frame.  khepri/beams.py:104-133."""
    E, H = _beam(1, x, y, z, wl, _IDENTITY_GEOM, w0, er, p, l)
    return (E[0], E[1], E[2]), (H[0], H[1], H[2])


def _builtin(field_fn, kwargs):
    """(kind, parameters) when field_fn is a built-in beam (possibly a functools.partial)."""
    kw = dict(kwargs)
    fn = field_fn
    while isinstance(fn, functools.partial):
        kw = {**fn.keywords, **kw}
        if fn.args:
            return None
        fn = fn.func
    if fn is _paraxial_gaussian_field_fn and set(kw) <= {"beam_waist", "er"}:
        return 0, dict(w0=kw.get("beam_waist", 1), er=kw.get("er", 1))
    if fn is _paraxial_laguerre_gaussian_field_fn and set(kw) <= {"w0", "er", "p", "l"}:
        return 1, dict(w0=kw.get("w0", 1), er=kw.get("er", 1), p=kw.get("p", 0), l=kw.get("l", 0))
    return None


def shifted_rotated_fields(field_fn, x, y, z, wavelength, beam_origin_x, beam_origin_y, beam_origin_z,
                           polar_angle, azimuthal_angle, polarization_angle, **kwargs):
    """Fields of a beam defined in its own frame, placed at the origin o and
    rotated by rotation_matrix(...): field_fn is evaluated at M^-1 r - M^-1 o
    and E, H are rotated back by M.  Returns (2, 3, *x.shape, 1) like the
    reference.  khepri/beams.py:50-101."""
    x, y, z = np.broadcast_arrays(np.asarray(x, dtype=float), np.asarray(y, dtype=float),
                                  np.asarray(z, dtype=float))
    mat = rotation_matrix(polar_angle, azimuthal_angle, polarization_angle)
    minv = np.linalg.inv(mat)
    r0 = np.linalg.solve(mat, np.array([beam_origin_x, beam_origin_y, beam_origin_z], dtype=float))
    b = _builtin(field_fn, kwargs)
    if b is not None:
        kind, prm = b
        E, H = _beam(kind, x, y, z, wavelength, np.concatenate([mat.ravel(), minv.ravel(), r0]), **prm)
    else:   # a caller's own field function: evaluated by it on host arrays
        coords = np.stack([x, y, z], axis=-1)
        r = np.einsum("ij,...j->...i", minv, coords) - r0
        (exr, eyr, ezr), (hxr, hyr, hzr) = field_fn(r[..., 0], r[..., 1], r[..., 2], wavelength, **kwargs)
        Er = np.stack(np.broadcast_arrays(exr, eyr, ezr), axis=0)
        Hr = np.stack(np.broadcast_arrays(hxr, hyr, hzr), axis=0)
        E = np.einsum("ij,j...->i...", mat, Er)
        H = np.einsum("ij,j...->i...", mat, Hr)
    return np.asarray([E, H])[..., np.newaxis]


def amplitudes_from_fields(fields, e, wl, kp, x, y, bzs, a=1):
    """Fourier amplitudes (Ex, Ey, Hx, Hy) x harmonics of fields sampled on
    bzs[0] x bzs[1] unit cells (fields (rows, cols, 2, 3), x, y (rows, cols),
    rows = ns bzs[1], cols = ns bzs[0]), after removing the Bloch phase of kp;
    normalised by (number of cells x points per cell side) like the
    reference.  The cells of the fields and of the coordinates are paired in
    the orders the reference splits them.  The whole computation, phase
    removal included, is one libkhx launch (khx_bloch_amplitudes); numpy
    arrays or device tensors are accepted.  khepri/beams.py:164-190."""
    lib = _lib.load()
    dev = _device()

    def on_dev(v, dtype):
        if isinstance(v, torch.Tensor):
            return v.to(device=dev, dtype=dtype).contiguous()
        return torch.as_tensor(np.ascontiguousarray(v), dtype=dtype, device=dev)
    f = on_dev(fields, torch.complex128)
    xd, yd = on_dev(x, torch.float64), on_dev(y, torch.float64)
    rows, cols = xd.shape
    if tuple(f.shape) != (rows, cols, 2, 3) or tuple(yd.shape) != (rows, cols):
        raise ValueError(f"fields must be (rows, cols, 2, 3) and x, y (rows, cols); got {tuple(f.shape)}, "
                         f"{tuple(xd.shape)}, {tuple(yd.shape)}")
    if rows % bzs[1] or cols % bzs[0] or rows // bzs[1] != cols // bzs[0]:
        raise ValueError(f"{rows} x {cols} samples do not split into {bzs[0]} x {bzs[1]} square cells")
    ns = rows // bzs[1]
    kx, ky = (torch.as_tensor(np.ascontiguousarray(np.ravel(v), dtype=np.float64), device=dev) for v in e.g_vectors)
    nk = kx.numel()
    out = torch.empty((4, nk), dtype=torch.complex128, device=dev)
    check(lib.khx_bloch_amplitudes(f.data_ptr(), xd.data_ptr(), yd.data_ptr(), ns, int(bzs[0]), int(bzs[1]),
                                   float(np.real(kp[0])), float(np.real(kp[1])), kx.data_ptr(), ky.data_ptr(), nk,
                                   out.data_ptr(), torch.cuda.current_stream().cuda_stream), "khx_bloch_amplitudes")
    return out.cpu().numpy()


def points_dft(f, x, y, kx, ky):
    """sum_p f[..., p] exp(-i (kx x_p + ky y_p)) for each (kx, ky) on the
    device; f (..., npts) with x, y (npts,) shared.  khepri/fourier.py:94-104."""
    lib = _lib.load()
    dev = _device()
    f = np.asarray(f, dtype=np.complex128)
    lead = f.shape[:-1]
    fb = torch.as_tensor(np.ascontiguousarray(f.reshape(-1, f.shape[-1])), device=dev)
    batch, npts = fb.shape
    xb = torch.as_tensor(np.tile(np.asarray(x, dtype=np.float64).ravel(), batch), device=dev)
    yb = torch.as_tensor(np.tile(np.asarray(y, dtype=np.float64).ravel(), batch), device=dev)
    kxd = torch.as_tensor(np.ascontiguousarray(np.ravel(kx), dtype=np.float64), device=dev)
    kyd = torch.as_tensor(np.ascontiguousarray(np.ravel(ky), dtype=np.float64), device=dev)
    out = torch.empty((batch, kxd.numel()), dtype=torch.complex128, device=dev)
    check(lib.khx_points_dft(fb.data_ptr(), xb.data_ptr(), yb.data_ptr(), npts, kxd.data_ptr(), kyd.data_ptr(),
                             kxd.numel(), batch, 1.0, out.data_ptr(), torch.cuda.current_stream().cuda_stream),
          "khx_points_dft")
    return out.cpu().numpy().reshape(*lead, kxd.numel())

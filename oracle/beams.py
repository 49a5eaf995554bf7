"""numpy restatement of the reference's beam / Brillouin-zone front end
(khepri/beams.py) -- TEST ORACLE ONLY.

Each function names the reference function and file:line it restates.
Pinned by golden vectors written from the reference (tests/golden/beams.npz,
tests/golden/make_golden.py: gen_beams) and by closed forms
(tests/test_oracle.py): rotation matrices are orthonormal with the right
axes, the Gaussian beam reduces to its waist profile at z = 0, the LG beam
with p = l = 0 is the Gaussian beam, and the point DFT of a single plane
wave picks its harmonic.
"""
from __future__ import annotations

from math import prod

import numpy as np


def rotation_matrix(polar_angle, azimuthal_angle, polarization_angle):
    """R_p(polarization about k) R_z(azimuthal) R_y(polar).  khepri/beams.py:6-47."""
    cp, sp = np.cos(polar_angle), np.sin(polar_angle)
    ca, sa = np.cos(azimuthal_angle), np.sin(azimuthal_angle)
    cq, sq = np.cos(polarization_angle), np.sin(polarization_angle)
    ry = np.array([[cp, 0.0, sp], [0.0, 1.0, 0.0], [-sp, 0.0, cp]])
    rz = np.array([[ca, -sa, 0.0], [sa, ca, 0.0], [0.0, 0.0, 1.0]])
    ux, uy, uz = ca * sp, sa * sp, cp
    rp = np.array([
        [cq + ux ** 2 * (1 - cq), ux * uy * (1 - cq) - uz * sq, ux * uz * (1 - cq) + uy * sq],
        [uy * ux * (1 - cq) + uz * sq, cq + uy ** 2 * (1 - cq), uy * uz * (1 - cq) - ux * sq],
        [uz * ux * (1 - cq) - uy * sq, uz * uy * (1 - cq) + ux * sq, cq + uz ** 2 * (1 - cq)],
    ])
    return rp @ rz @ ry


def shifted_rotated_fields(field_fn, x, y, z, wavelength, ox, oy, oz, polar, azimuthal, polarization, **kw):
    """Evaluate field_fn in the beam frame (coordinates M^-1 (r - o)) and
    rotate E, H back (M E_r).  Returns (2, 3, *x.shape, 1).  khepri/beams.py:50-101."""
    mat = rotation_matrix(polar, azimuthal, polarization)
    coords = np.stack([x, y, z], axis=-1).reshape(*x.shape, 3, 1)
    r = np.linalg.solve(np.broadcast_to(mat, (*x.shape, 3, 3)), coords)[..., 0]
    r0 = np.linalg.solve(mat, np.array([ox, oy, oz], dtype=float))
    (exr, eyr, ezr), (hxr, hyr, hzr) = field_fn(r[..., 0] - r0[0], r[..., 1] - r0[1], r[..., 2] - r0[2],
                                                wavelength, **kw)
    E = np.stack(np.broadcast_arrays(exr, eyr, ezr), axis=-1)
    H = np.stack(np.broadcast_arrays(hxr, hyr, hzr), axis=-1)
    E = np.einsum("ij,...j->...i", mat, E)
    H = np.einsum("ij,...j->...i", mat, H)
    # the reference's squeezes leave a trailing axis of length 1 on every component
    return np.asarray([np.moveaxis(E, -1, 0), np.moveaxis(H, -1, 0)])[..., np.newaxis]


def paraxial_gaussian_field_fn(x, y, z, wl, beam_waist=1, er=1):
    """x-polarised paraxial Gaussian beam.  khepri/beams.py:135-160."""
    k = 2 * np.pi / wl
    z_r = np.pi * beam_waist ** 2 * np.sqrt(er) / wl
    w_z = beam_waist * np.sqrt(1 + (z / z_r) ** 2)
    r = np.sqrt(x ** 2 + y ** 2)
    ex = (beam_waist / w_z * np.exp(-(r ** 2) / w_z ** 2)
          * np.exp(1j * (k * z + k * r ** 2 / 2 * z / (z ** 2 + z_r ** 2) - np.arctan(z / z_r))))
    zero = np.zeros_like(ex)
    return (ex, zero, zero), (zero, ex / np.sqrt(er), zero)


def paraxial_laguerre_gaussian_field_fn(x, y, z, wl, w0=1, er=1, p=0, l=0):
    """x-polarised paraxial Laguerre-Gauss beam LG_pl.  khepri/beams.py:104-133."""
    from scipy.special import genlaguerre
    k = 2 * np.pi / wl
    z_r = np.pi * w0 ** 2 * np.sqrt(er) / wl
    N = abs(l) + 2 * p
    w_z = w0 * np.sqrt(1 + (z / z_r) ** 2)
    phi = np.arctan2(x, y)
    r = np.sqrt(x ** 2 + y ** 2)
    ex = (w0 / w_z * (r * np.sqrt(2) / w_z) ** abs(l) * np.exp(-(r ** 2) / w_z ** 2)
          * genlaguerre(p, abs(l))(2 * r ** 2 / w_z ** 2) * np.exp(-1j * l * phi)
          * np.exp(1j * (k * z + k * r ** 2 / 2 * z / (z ** 2 + z_r ** 2) - (N + 1) * np.arctan(z / z_r))))
    zero = np.zeros_like(ex)
    return (ex, zero, zero), (zero, ex / np.sqrt(er), zero)


def slow_dft(real_field, x, y, kx, ky):
    """sum_k f_k e^{-i (kx x_k + ky y_k)}.  khepri/fourier.py:94-104."""
    return np.exp(-1j * (np.outer(kx, x) + np.outer(ky, y))) @ real_field


def split(array, nrows, ncols):
    """khepri/misc.py:20-26."""
    r, h = array.shape
    return array.reshape(h // nrows, nrows, -1, ncols).swapaxes(1, 2).reshape(-1, nrows, ncols)


def amplitudes_from_fields(fields, g_vectors, pw, kp, x, y, bzs):
    """Fourier amplitudes (Ex, Ey, Hx, Hy) x harmonics of sampled fields
    (rows, cols, 2, 3) over bzs cells; the cells of F and of x, y are split
    in the orders the reference uses.  khepri/beams.py:164-190."""
    kxi, kyi = kp
    phase = np.exp(1j * (kxi * x + kyi * y))
    F = fields / phase[..., np.newaxis, np.newaxis]
    F = np.asarray(np.split(F, bzs[0], axis=1))
    F = np.asarray(np.split(F, bzs[1], axis=1))
    F = np.swapaxes(F, 0, 1)
    NS = F.shape[2]
    xs = np.asarray(np.split(x, bzs[0], axis=1))
    xs = np.asarray(np.split(xs, bzs[1], axis=1)).reshape(-1, NS, NS)
    ys = np.asarray(np.split(y, bzs[0], axis=1))
    ys = np.asarray(np.split(ys, bzs[1], axis=1)).reshape(-1, NS, NS)
    kx, ky = g_vectors
    F = F.reshape(-1, NS, NS, 2, 3)
    F = F[..., :2].reshape(-1, NS, NS, 4)
    out = np.empty((prod(bzs), 4, prod(pw)), dtype=np.complex128)
    for i in range(F.shape[0]):
        for j in range(4):
            out[i, j] = slow_dft(F[i, ..., j].flatten(), xs[i].flatten(), ys[i].flatten(),
                                 np.ravel(kx), np.ravel(ky)) / F.shape[0] / F.shape[1]
    return np.sum(out, 0)

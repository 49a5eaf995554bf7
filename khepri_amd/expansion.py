"""Plane-wave basis (mirrors khepri/expansion.py and khepri/tools.py helpers).

Host-side set-up only (O(N) numbers per expansion); the per-solve k vectors
are produced on the device by ``khx_kvectors``.
"""
from __future__ import annotations

import cmath
from math import prod

import numpy as np
import torch

TWOPI = 2 * np.pi


def generate_expansion_indices(pw):
    """Harmonic orders as a (2, P Q) integer array for pw = (P, Q), both odd:
    row 0 holds m in -(P-1)/2 .. (P-1)/2, row 1 holds n likewise, m varying
    fastest -- the ordering of khepri/expansion.py:4-16, which every matrix
    of the basis follows.  An even count fails the same assertion."""
    P, Q = pw
    assert P % 2 == 1 and Q % 2 == 1, f"plane-wave counts must be odd, got {pw}"
    m = np.arange(P) - P // 2
    n = np.arange(Q) - Q // 2
    return np.stack([np.tile(m, Q), np.repeat(n, P)])


def _cross(a1, a2):
    """z component of a1 x a2 (the signed cell area)."""
    return a1[0] * a2[1] - a1[1] * a2[0]


def unitcellarea(a1, a2):
    """Area of the cell spanned by a1, a2."""
    return abs(_cross(a1, a2))


def reciproc(a1, a2):
    """Reciprocal basis (b1, b2) with a_i . b_j = 2 pi delta_ij, as tuples
    (khepri/tools.py:64-69): b1 is a2 turned by -90 degrees, b2 is a1 turned
    by +90 degrees, both scaled by 2 pi / (a1 x a2)."""
    scale = TWOPI / _cross(a1, a2)
    b1 = (a2[1] * scale, -a2[0] * scale)
    b2 = (-a1[1] * scale, a1[0] * scale)
    return b1, b2


def rotation_matrix(theta):
    """Counter-clockwise rotation by theta radians."""
    c, s = np.cos(theta), np.sin(theta)
    return np.array([[c, -s], [s, c]])


def compute_kplanar(eps_inc, wavelength, theta_deg=0.0, phi_deg=0.0):
    """In-plane wave vector of a plane wave in the incidence medium, polar
    angle theta and azimuth phi in degrees (khepri/tools.py:17-21):
    n k0 sin(theta) (cos phi, sin phi), complex."""
    k_inc = cmath.sqrt(eps_inc) * TWOPI / wavelength
    theta, phi = np.deg2rad(theta_deg), np.deg2rad(phi_deg)
    return (k_inc * np.sin(theta)) * np.array([np.cos(phi), np.sin(phi)], dtype=complex)


def kz_from_kplanar(kx, ky, k0, epsilon):
    """Normal wave-vector component with the reference's branch
    (khepri/expansion.py:19-26): conj(sqrt(k0^2 conj(eps) - kx^2 - ky^2))."""
    kz2 = np.asarray(k0 ** 2 * np.conj(epsilon) - kx ** 2 - ky ** 2, dtype=complex)
    return np.conj(np.sqrt(kz2))


class Expansion:
    """Harmonics of a lattice (khepri/expansion.py:28-97).

    ``expansion_lhs`` / ``expansion_rhs`` are set on Minkowski sums (twisted
    bilayers); ``k_vectors`` returns host numpy arrays like the reference and
    keeps no shared mutable state, so expansions may be used from several
    threads.
    """

    def __init__(self, pw, lattice=None) -> None:
        self.pw = tuple(pw)
        if lattice is None:
            lattice = np.asarray([[1, 0], [0, 1]])
        lattice = np.asarray(lattice)
        self.lattice = lattice
        self.a = np.linalg.norm(lattice[0])
        self.reciprocal = np.asarray(reciproc(lattice[0], lattice[1]))
        self.expansion_indices = generate_expansion_indices(self.pw)
        self._g_vectors = np.sum(self.reciprocal[:, :, None] * self.expansion_indices[:, None], axis=0)
        self.sigma = unitcellarea(*lattice)
        self.expansion_lhs = self.expansion_rhs = None
        self._dev_cache = {}

    @property
    def N(self):
        return self._g_vectors.shape[1]

    @property
    def g_vectors(self):
        return self._g_vectors.copy()

    def rotate(self, angle_rad):
        """Rotates the G vectors only (khepri/expansion.py:52-53)."""
        self._g_vectors = rotation_matrix(angle_rad) @ self._g_vectors
        self._dev_cache = {}

    def __add__(self, rhs):
        """Minkowski sum: index i_self * N_rhs + i_rhs (khepri/expansion.py:55-73)."""
        g = (rhs._g_vectors[:, None, :] + self._g_vectors[:, :, None]).reshape(2, -1)
        e = Expansion((self.pw[0] ** 2, self.pw[1] ** 2))
        e._g_vectors = g
        e._base_g_vectors = np.stack((self._g_vectors, rhs._g_vectors))
        e.expansion_lhs = self
        e.expansion_rhs = rhs
        return e

    def k_vectors(self, k_parallel, wavelength, epsilon=1):
        k0 = TWOPI / wavelength
        kv = np.zeros((3, self.N), dtype=np.complex128)
        kv[0:2] = np.asarray(k_parallel)[:, None] + self._g_vectors
        kv[2] = kz_from_kplanar(kv[0], kv[1], k0, epsilon)
        return kv / k0

    def g_device(self, device):
        key = str(device)
        if key not in self._dev_cache:
            self._dev_cache[key] = torch.as_tensor(np.ascontiguousarray(self._g_vectors, dtype=np.float64),
                                                   device=device)
        return self._dev_cache[key]

    @property
    def nharm(self):
        return prod(self.pw) if self.expansion_lhs is None else self.N

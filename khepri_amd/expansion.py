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
    """(2, N) integer pairs (m, n), m fastest (khepri/expansion.py:4-16)."""
    assert pw[0] % 2 == 1
    assert pw[1] % 2 == 1
    M, N = (pw[0] - 1) // 2, (pw[1] - 1) // 2
    mm, nn = np.meshgrid(np.arange(-M, M + 1), np.arange(-N, N + 1))
    return np.vstack([mm.ravel(), nn.ravel()])


def reciproc(a1, a2):
    """khepri/tools.py:64-69."""
    coef = TWOPI / (a1[0] * a2[1] - a1[1] * a2[0])
    return (a2[1] * coef, -a2[0] * coef), (-a1[1] * coef, a1[0] * coef)


def unitcellarea(a1, a2):
    return abs(a1[0] * a2[1] - a1[1] * a2[0])


def rotation_matrix(theta):
    return np.array([[np.cos(theta), -np.sin(theta)], [np.sin(theta), np.cos(theta)]])


def compute_kplanar(eps_inc, wavelength, theta_deg=0.0, phi_deg=0.0):
    """In-plane wave vector from angles in degrees (khepri/tools.py:17-21)."""
    phi = np.deg2rad(phi_deg)
    theta = np.deg2rad(theta_deg)
    kp = np.array([np.cos(phi), np.sin(phi)], dtype=complex)
    return kp * cmath.sqrt(eps_inc) * TWOPI / wavelength * np.sin(theta)


def kz_from_kplanar(kx, ky, k0, epsilon):
    """khepri/expansion.py:19-26."""
    arg = k0 ** 2 * np.conj(epsilon) - kx ** 2 - ky ** 2
    return np.conj(np.sqrt(np.asarray(arg).astype(complex)))


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

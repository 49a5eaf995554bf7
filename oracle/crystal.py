"""Oracle restatement of the reference ``Crystal`` / ``Layer`` / ``ExtendedLayer``
objects (test oracle and CPU baseline only).

Follows khepri/crystal.py:42-400, khepri/layer.py:63-193 and
khepri/extension.py:66-126.  Layers are solved with oracle.rcwa.
"""
from __future__ import annotations

import cmath
from copy import copy

import numpy as np

from . import rcwa

UNIFORM, FFT, ANALYTICAL, HALF_INC, HALF_TRN = 0, 1, 2, 3, 4


class Expansion:
    """khepri/expansion.py:28-97 (g vectors only; no mutable k cache)."""

    def __init__(self, pw, lattice=None):
        self.pw = tuple(pw)
        if lattice is None:
            lattice = np.asarray([[1, 0], [0, 1]])
        self.lattice = np.asarray(lattice)
        self.g = rcwa.g_vectors(self.pw, self.lattice)
        self.lhs = self.rhs = None

    def rotate(self, angle_rad):
        """khepri/expansion.py:52-53 (radians, rotates G only)."""
        self.g = rcwa.rotation_matrix(angle_rad) @ self.g

    def __add__(self, rhs):
        """khepri/expansion.py:55-73."""
        e = Expansion((self.pw[0] ** 2, self.pw[1] ** 2))
        e.g = rcwa.minkowski_sum(self.g, rhs.g)
        e.lhs, e.rhs = self, rhs
        return e

    def k_vectors(self, kp, wl, eps=1.0):
        return rcwa.k_vectors(self.g, kp, wl, eps)


class Layer:
    """khepri/layer.py:63-193."""

    def __init__(self, expansion, kind, epsilon, depth):
        self.expansion = expansion
        self.kind = kind
        self.epsilon = epsilon
        self.depth = depth
        self.fields = False
        self.S = self.W = self.V = self.L = self.IC = None
        self._C = None

    @classmethod
    def pixmap(cls, e, eps, depth):
        return cls(e, FFT, eps, depth)

    @classmethod
    def uniform(cls, e, eps, depth):
        return cls(e, UNIFORM, eps, depth)

    @classmethod
    def analytical(cls, e, islands, eps_host, depth):
        """khepri/layer.py:112-119."""
        layer = cls(e, ANALYTICAL, islands, depth)
        layer.eps_host = eps_host
        return layer

    @classmethod
    def half_infinite(cls, e, side, eps):
        return cls(e, HALF_INC if side == "reflexion" else HALF_TRN, eps, 0)

    def solve(self, kp, wl, reuse_convmat=False):
        kx, ky, _ = self.expansion.k_vectors(kp, wl)
        W0, V0 = rcwa.free_space_eigenmodes(kx, ky)
        k0 = 2 * np.pi / wl
        if self.kind in (FFT, ANALYTICAL):
            # the reference recomputes C every solve (khepri/layer.py:157, 161-166)
            if self._C is None or not reuse_convmat:
                if self.kind == FFT:
                    self._C = rcwa.convolution_matrix(self.epsilon, self.expansion.pw)
                else:
                    self._C = rcwa.analytic_convolution_matrix(self.epsilon, self.eps_host, self.expansion.pw,
                                                               self.expansion.lattice)
            C = self._C
            self.IC = np.linalg.inv(C)
            self.W, self.V, self.L = rcwa.structured_layer_modes(kx, ky, C)
            self.S = rcwa.build_scatmat(self.W, self.V, W0, V0, self.L, self.depth, k0)
        elif self.kind == UNIFORM:
            self.W, self.V, self.L = rcwa.uniform_layer_modes(kx, ky, self.epsilon)
            self.S = rcwa.build_scatmat(self.W, self.V, W0, V0, self.L, self.depth, k0)
            self.IC = 1 / self.epsilon
        elif self.kind == HALF_INC:
            self.S, self.W, self.V, self.L = rcwa.halfspace_reflection(kx, ky, W0, V0, self.epsilon)
            self.IC = 1 / self.epsilon
        else:
            self.S, self.W, self.V, self.L = rcwa.halfspace_transmission(kx, ky, W0, V0, self.epsilon)
            self.IC = 1 / self.epsilon
        if not self.fields:
            self.W = self.V = self.L = self.IC = None


class ExtendedLayer:
    """khepri/extension.py:66-112."""

    def __init__(self, expansion, base):
        if expansion.lhs is base.expansion:
            self.mode, self.gs = 1, expansion.rhs.g
        elif expansion.rhs is base.expansion:
            self.mode, self.gs = 0, expansion.lhs.g
        else:
            raise NotImplementedError("base expansion must be part of the extended one")
        self.expansion = expansion
        self.base = base
        self.depth = base.depth
        self.fields = base.fields

    def solve(self, kp, wl, reuse_convmat=False):
        if isinstance(self.base, Layer):
            self.base.fields = self.fields
        Ss, Ws, Vs, Ls = [], [], [], []
        for g in self.gs.T:
            if isinstance(self.base, Layer):
                self.base.solve(np.asarray(kp) + g, wl, reuse_convmat=reuse_convmat)
            else:  # a Crystal as the base (khepri/extension.py:93-95)
                self.base.set_source(wl, np.nan, np.nan, kp=np.asarray(kp) + g)
                self.base.solve()
            Ss.append(self.base.S.copy())
            if self.fields:
                Ws.append(self.base.W.copy())
                Vs.append(self.base.V.copy())
                Ls.append(np.diag(self.base.L).copy())
        self.S = rcwa.joint_subspace_S(Ss, self.mode)
        if self.fields:
            self.W = rcwa.joint_subspace(Ws, self.mode)
            self.V = rcwa.joint_subspace(Vs, self.mode)
            W0s, V0s = [], []
            for g in self.gs.T:
                kx, ky, _ = self.base.expansion.k_vectors(np.asarray(kp) + g, wl)
                W0, V0 = rcwa.free_space_eigenmodes(kx, ky)
                W0s.append(W0)
                V0s.append(V0)
            self.W0 = rcwa.joint_subspace(W0s, self.mode)
            self.V0 = rcwa.joint_subspace(V0s, self.mode)
            self.L = np.diag(rcwa.joint_subspace(Ls, self.mode))
        self.IC = 1.0


class Crystal:
    """khepri/crystal.py:34-400."""

    def __init__(self, pw, lattice="square", lattice_pitch=1, void=False, epsi=1, epse=1):
        self.pw = tuple(pw)
        self.epsi, self.epse, self.void = epsi, epse, void
        self.lattice = rcwa.lattice_matrix(lattice, lattice_pitch)
        self.expansion = Expansion(self.pw, self.lattice)
        self.layers = {}
        self.S = None
        self.reuse_convmat = False

    @classmethod
    def from_expansion(cls, e, **kw):
        obj = cls(e.pw, **kw)
        obj.expansion = e
        return obj

    def add_layer_uniform(self, name, eps, depth):
        self.layers[name] = Layer.uniform(self.expansion, eps, depth)

    def add_layer_analytical(self, name, islands, eps_host, depth):
        """khepri/crystal.py:119-123."""
        self.layers[name] = Layer.analytical(self.expansion, islands, eps_host, depth)

    def add_layer_pixmap(self, name, eps, depth):
        self.layers[name] = Layer.pixmap(self.expansion, eps, depth)

    def add_layer(self, name, layer, extended=False):
        self.layers[name] = ExtendedLayer(self.expansion, layer) if extended else layer

    def set_device(self, stack, fields_mask=False):
        """khepri/crystal.py:131-164."""
        self.device_stack = copy(stack)
        gs = [] if self.void else ["Sref"]
        gs += list(stack)
        if not self.void:
            gs.append("Strans")
        self.global_stacking = gs
        if "Sref" in gs and "Sref" not in self.layers:
            self.layers["Sref"] = Layer.half_infinite(self.expansion, "reflexion", self.epsi)
            self.layers["Sref"].fields = True
        if "Strans" in gs and "Strans" not in self.layers:
            self.layers["Strans"] = Layer.half_infinite(self.expansion, "transmission", self.epse)
            self.layers["Strans"].fields = True
        if isinstance(fields_mask, bool):
            fields_mask = [fields_mask] * len(stack)
        self.stack_retain_mask = [True] + list(fields_mask) + [True]
        for name, on in zip(gs, self.stack_retain_mask):
            self.layers[name].fields |= on
            if hasattr(self.layers[name], "base"):
                self.layers[name].base.fields |= on

    @property
    def depth(self):
        return sum(self.layers[n].depth for n in self.device_stack)

    def set_source(self, wavelength, te=1.0, tm=1.0, theta=0.0, phi=0.0, kp=None):
        """khepri/crystal.py:345-360."""
        self.wavelength, self.te, self.tm = wavelength, te, tm
        if kp is not None:
            self.kp = tuple(kp)
        else:
            self.kp = tuple(rcwa.compute_kplanar(self.epsi, wavelength, theta, phi))
        self.k0 = 2 * np.pi / wavelength
        kxi, kyi = self.kp
        self.kzi = np.conj(cmath.sqrt(self.k0 ** 2 * self.epsi - kxi ** 2 - kyi ** 2))

    def solve(self, forward_only=False):
        """khepri/crystal.py:180-206."""
        for name in set(self.global_stacking):
            self.layers[name].solve(self.kp, self.wavelength, reuse_convmat=self.reuse_convmat)
        stacked = [self.layers[n] for n in self.global_stacking]
        pos = list(np.cumsum([l.depth for l in stacked]))
        pos[-1] = np.inf
        if not self.void:
            pos.insert(0, -np.inf)
        self.stack_positions = pos
        self.stacking_matrices, self.stacking_reverse_matrices, self.Stot = rcwa.stack_layers(
            [l.S for l in stacked], self.stack_retain_mask, forward_only=forward_only)
        self.S = self.Stot

    def poynting_flux_end(self, only_total=True):
        """khepri/crystal.py:363-396."""
        inc = rcwa.incident(self.pw, self.te, self.tm, (self.kp[0], self.kp[1], self.kzi))
        Wref = self.layers["Sref"].W
        c1p = np.linalg.inv(Wref) @ inc
        Wtr = self.layers["Strans"].W
        g = self.expansion.g
        T = rcwa.poynting_fluxes(g, Wtr @ self.Stot[1, 0] @ c1p, self.kp, self.wavelength,
                                 only_total, self.epsi, self.epse)
        R = rcwa.poynting_fluxes(g, Wref @ self.Stot[0, 0] @ c1p, self.kp, self.wavelength,
                                 only_total, self.epsi, self.epsi)
        if only_total:
            return R.real, T.real
        return R, T

    # ---------------------------------------------------------------- fields
    def locate_layer(self, z):
        """khepri/crystal.py:208-232."""
        idx = int(np.searchsorted(self.stack_positions, z)) - 1
        name = self.global_stacking[idx]
        zr = z if z <= 0 else z - self.stack_positions[idx]
        return self.layers[name], idx, zr

    def source_fields(self):
        """khepri/crystal.py:281-295."""
        e = rcwa.incident(self.pw, self.te, self.tm, (self.kp[0], self.kp[1], self.kzi))
        return e, np.zeros_like(e)

    def fourier_fields(self, z, inc):
        """khepri/crystal.py:234-279."""
        layer, li, zr = self.locate_layer(z)
        RI = rcwa.eigenbasis_matrix(layer.W, layer.V)
        kx, ky, _ = self.expansion.k_vectors(self.kp, self.wavelength)
        if isinstance(layer, Layer):
            W0, V0 = rcwa.free_space_eigenmodes(kx, ky)
        else:
            W0, V0 = layer.W0, layer.V0
        R0 = rcwa.eigenbasis_matrix(W0, V0)
        k0 = 2 * np.pi / self.wavelength
        Rref = rcwa.eigenbasis_matrix(self.layers["Sref"].W, self.layers["Sref"].V)
        c1p = np.split(np.linalg.solve(Rref, inc), 2)[0]
        c1m = self.Stot[0, 0] @ c1p
        c2p = self.Stot[1, 0] @ c1p
        cdp, cdm = rcwa.translate_mode_amplitudes(self.stacking_matrices[li],
                                                  self.stacking_reverse_matrices[li], c1p, c1m, c2p)
        sx, sy, ux, uy = rcwa.fourier_fields_from_modes(RI, layer.L, R0, (cdp, cdm),
                                                        k0 * (layer.depth - zr))
        sz, uz = rcwa.longitudinal_fields((sx, sy, ux, uy), kx, ky, layer.IC)
        return sx, sy, sz, ux, uy, uz

    def fields_coords_xy(self, x, y, z, incident_fields=None):
        """khepri/crystal.py:297-333."""
        x, y = np.asarray(x), np.asarray(y)
        if incident_fields is None:
            incident_fields = np.hstack(self.source_fields())
        elif isinstance(incident_fields, tuple):
            incident_fields = np.hstack(incident_fields)
        kx, ky, _ = self.expansion.k_vectors(self.kp, self.wavelength)
        ff = self.fourier_fields(z, incident_fields)
        k0 = 2 * np.pi / self.wavelength
        f = np.asarray([rcwa.idft(s, k0 * kx, k0 * ky, x, y) for s in ff])
        return f[:3], f[3:]

    def fields_volume(self, x, y, z, incident_fields=None):
        """khepri/crystal.py:335-343."""
        if incident_fields is None:
            incident_fields = self.source_fields()
        inc = np.hstack(incident_fields)
        out = np.array([self.fields_coords_xy(x, y, zi, inc) for zi in z])
        return out[:, 0], out[:, 1]


# --------------------------------------------------------------------------
# input rasteriser (khepri/draw.py:21-55, 90-95)
# --------------------------------------------------------------------------
class Drawing:
    def __init__(self, shape, eps_bg):
        # np.ones * eps: float canvas unless the background is complex
        self.canvas_ = np.ones(shape) * eps_bg
        xb = np.linspace(-0.5, 0.5, shape[0])
        yb = np.linspace(-0.5, 0.5, shape[1])
        self.X, self.Y = np.meshgrid(xb, yb, indexing="ij")

    def disc(self, xy, r, eps):
        """Strict inequality.  khepri/draw.py:41-44."""
        self.canvas_[np.sqrt((self.X - xy[0]) ** 2 + (self.Y - xy[1]) ** 2) < r] = eps

    def rectangle(self, xy, wh, eps):
        """Closed edges.  khepri/draw.py:46-55."""
        x0, y0 = xy[0] - wh[0] / 2, xy[1] - wh[1] / 2
        m = (self.X >= x0) & (self.X <= x0 + wh[0]) & (self.Y >= y0) & (self.Y <= y0 + wh[1])
        self.canvas_[m] = eps

    def canvas(self):
        return self.canvas_.copy()

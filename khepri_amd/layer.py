"""Layers of a stack and the star-product chain (mirrors khepri/layer.py).

A layer solves for a whole batch of (k_parallel, wavelength) pairs at once;
its S-matrix is a device tensor of shape (B, 2, 2, n, n).  The reference's
single-pair protocol ``solve(k_parallel, wavelength)`` is kept (B = 1).
"""
from __future__ import annotations

from enum import IntEnum

import os

import numpy as np
import torch

from . import ops
from .device import default_device


class Formulation(IntEnum):
    UNIFORM = 0
    FFT = 1
    ANALYTICAL = 2
    HALF_SPACE_INC = 3
    HALF_SPACE_TRN = 4


class Field(IntEnum):
    X = 0
    Y = 1
    Z = 2
    NORM = 3
    POYNTING = 4


def as_source_batch(k_parallel, wavelength, device):
    """(B, 2) complex kp and (B,) float wavelength tensors on `device`."""
    if isinstance(k_parallel, torch.Tensor):
        kp = k_parallel.to(device=device, dtype=torch.complex128)
    else:
        kp = torch.as_tensor(np.asarray(k_parallel, dtype=np.complex128), device=device)
    if isinstance(wavelength, torch.Tensor):
        wl = wavelength.to(device=device, dtype=torch.float64)
    else:
        wl = torch.as_tensor(np.asarray(wavelength, dtype=np.float64), device=device)
    kp = kp.reshape(-1, 2)
    wl = wl.reshape(-1)
    B = max(kp.shape[0], wl.shape[0])
    if kp.shape[0] == 1 and B > 1:
        kp = kp.expand(B, 2)
    if wl.shape[0] == 1 and B > 1:
        wl = wl.expand(B)
    if kp.shape[0] != B or wl.shape[0] != B:
        raise ValueError("k_parallel and wavelength batch sizes differ")
    return kp.contiguous(), wl.contiguous()


class Layer:
    """One layer of the stack (khepri/layer.py:63-193)."""

    def __init__(self) -> None:
        self.formulation = None
        self.expansion = None
        self.W = self.V = self.L = None
        self.S = None
        self.IC = None
        self.fields = False
        self.device = None
        self._cm_key = None
        self._C = self._IC = None
        self.info = None

    # ------------------------------------------------------------- builders
    # The reference's classmethods (khepri/layer.py:78-129): each records the
    # layer's formulation, its permittivity description and its depth.
    @classmethod
    def _build(cls, expansion, formulation, epsilon, depth, **extra):
        layer = cls()
        layer.expansion, layer.formulation = expansion, formulation
        layer.epsilon, layer.depth = epsilon, depth
        for key, value in extra.items():
            setattr(layer, key, value)
        return layer

    @classmethod
    def pixmap_or_uniform(cls, expansion, pixmap, depth):
        """A uniform layer when every pixel holds the same value, else a pixmap layer."""
        first = pixmap.flat[0]
        if (pixmap == first).all():
            return cls.uniform(expansion, first, depth)
        return cls.pixmap(expansion, pixmap, depth)

    @classmethod
    def pixmap(cls, expansion, pixmap, depth):
        """Convolution matrix from the 2-D DFT of the pixel map."""
        return cls._build(expansion, Formulation.FFT, pixmap, depth)

    @classmethod
    def uniform(cls, expansion, epsilon, depth):
        """Homogeneous layer: modes in closed form, no eigenproblem."""
        return cls._build(expansion, Formulation.UNIFORM, epsilon, depth)

    @classmethod
    def half_infinite(cls, expansion, type, epsilon):
        """Incidence ("reflexion") or emergence ("transmission") half space."""
        sides = {"reflexion": Formulation.HALF_SPACE_INC, "transmission": Formulation.HALF_SPACE_TRN}
        if type not in sides:
            raise ValueError(f"half space type must be 'reflexion' or 'transmission', got {type!r}")
        return cls._build(expansion, sides[type], epsilon, 0)

    @classmethod
    def analytical(cls, expansion, islands_description, eps_host, depth):
        """Island layer with closed-form Fourier coefficients
        (khepri/layer.py:112-119): islands_description is a list of
        {"type": "disc" | "rectangle", "params": [...], "epsilon": eps}."""
        return cls._build(expansion, Formulation.ANALYTICAL, islands_description, depth, eps_host=eps_host)

    @property
    def patterned(self):
        """True when the layer's modes come from the eigen-solver."""
        return self.formulation in (Formulation.FFT, Formulation.ANALYTICAL)

    # ------------------------------------------------------------- solve
    def convolution_matrices(self, device):
        """C and C^-1 of a pixmap or island layer, computed once per content."""
        if self.formulation == Formulation.ANALYTICAL:
            key = (str(device), repr(self.epsilon), complex(self.eps_host))
            if self._cm_key != key:
                e = self.expansion
                C = ops.convmat_islands(self.epsilon, self.eps_host, e.pw, e.reciprocal, e.sigma, device)
                IC, info = ops.inverse(C[None])
                ops.raise_on_info(info, "inverse of the convolution matrix", lu=True)
                self._C, self._IC, self._cm_key = C, IC[0], key
            return self._C, self._IC
        eps = np.asarray(self.epsilon)
        key = (str(device), eps.shape, eps.dtype.str, hash(eps.tobytes()))
        if self._cm_key != key:
            t = torch.as_tensor(np.ascontiguousarray(eps), device=device)
            if not t.is_complex():
                t = t.to(torch.float64)
            C = ops.convmat(t, self.expansion.pw)
            IC, info = ops.inverse(C)
            ops.raise_on_info(info, "inverse of the convolution matrix", lu=True)
            self._C, self._IC, self._cm_key = C[0], IC[0], key
        return self._C, self._IC

    def solve(self, k_parallel, wavelength):
        """Reference protocol (khepri/layer.py:145-193); batched when the
        arguments are arrays."""
        dev = self.device or default_device()
        kp, wl = as_source_batch(k_parallel, wavelength, dev)
        self.solve_batch(kp, wl)

    def chain_item(self):
        return (self.skind, self.X if self.skind == "prop" else self.S)

    @staticmethod
    def solve_batch_many(layers, kp, wl, rt_only=False):
        """solve_batch for several layers of one expansion.  When all their
        eigenproblems fit one round of the eigen-solver (khx_eig_slots), the
        patterned layers are solved together in one fused call so that they
        share the GPU (a single layer's small batch leaves most CUs idle); the
        bound also keeps the fused workspace within one round's worth of
        items.  Everything else as solve_batch."""
        pat = [ly for ly in layers if getattr(ly, "patterned", False)]
        B = kp.shape[0]
        fuse = len(pat) >= 2 and B * len(pat) <= ops.eig_slots(2 * pat[0].expansion.N) and \
            all(ly.expansion is pat[0].expansion for ly in pat) and os.environ.get("KHX_FUSE_LAYERS", "1") == "1"
        if not fuse:
            for ly in layers:
                ly.solve_batch(kp, wl, rt_only=rt_only)
            return
        dev = kp.device
        kx, ky = ops.kvectors(pat[0].expansion.g_device(dev), kp, wl)
        k0 = 2 * np.pi / wl
        mats = [ly.convolution_matrices(dev) for ly in pat]
        modes = any(bool(ly.fields) for ly in pat)
        res = ops.patterned_layers_fused([m[0] for m in mats], [m[1] for m in mats], kx, ky,
                                         [(k0 * ly.depth).contiguous() for ly in pat], modes=modes)
        for ly, (C, IC), (S, (W, V, L), info) in zip(pat, mats, res):
            ly.device = dev
            ly.skind, ly.X = "dense", None
            ly.S, ly.info, ly.C = S, info, C
            keep = bool(ly.fields)
            ly.IC = IC if keep else None
            ly.W, ly.V, ly.L = (W, V, L) if keep else (None, None, None)
            ly.kx, ly.ky = (kx, ky) if keep else (None, None)
        for ly in layers:
            if not getattr(ly, "patterned", False):
                ly.solve_batch(kp, wl, rt_only=rt_only)

    def solve_batch(self, kp, wl, ws=None, rt_only=False):
        """S-matrix for a batch.  With rt_only (R/T wanted, no fields), layers of
        the gap medium are kept as their exact structure instead of a dense S:
        a uniform eps = 1 layer as its propagation factors X ("prop"), an
        eps = 1 half space as the identity S-matrix ("identity")."""
        dev = kp.device
        self.device = dev
        kx, ky = ops.kvectors(self.expansion.g_device(dev), kp, wl)
        k0 = 2 * np.pi / wl
        modes = bool(self.fields)
        self.info = None
        self.skind, self.X = "dense", None
        gap_medium = np.isscalar(self.epsilon) and complex(self.epsilon) == 1
        # rt_only never reconstructs fields, so the retained eigenspaces of the
        # half spaces (fields flag) are not needed either
        if rt_only and gap_medium:
            if self.formulation == Formulation.UNIFORM:
                self.skind, self.S = "prop", None
                self.X = ops.propagation(kx, ky, (k0 * self.depth).contiguous())
                self.W = self.V = self.L = self.IC = None
                return
            if self.formulation in (Formulation.HALF_SPACE_INC, Formulation.HALF_SPACE_TRN):
                self.skind, self.S = "identity", None
                self.W = self.V = self.L = self.IC = None
                return
        if self.patterned:
            C, IC = self.convolution_matrices(dev)
            k0d = (k0 * self.depth).contiguous()
            self.S, (W, V, L), self.info = ops.patterned_layer(C, IC, kx, ky, k0d, modes=modes, ws=ws)
            self.C = C
            self.IC = IC if modes else None
        elif self.formulation == Formulation.UNIFORM:
            # 1 / eps as the reference's P = (1/er) [...] (khepri/alternative.py:138):
            # ZeroDivisionError for a Python zero, inf (non-finite R, T) for a numpy one
            inv = 1 / self.epsilon
            k0d = (k0 * self.depth).contiguous()
            self.S, (W, V, L) = ops.analytic_layer("uniform", kx, ky, self.epsilon, k0d, modes=modes)
            self.IC = inv if modes else None
        elif self.formulation in (Formulation.HALF_SPACE_INC, Formulation.HALF_SPACE_TRN):
            inv = 1 / self.epsilon
            side = "reflexion" if self.formulation == Formulation.HALF_SPACE_INC else "transmission"
            self.S, (W, V, L) = ops.analytic_layer(side, kx, ky, self.epsilon, modes=modes)
            self.IC = inv if modes else None
        else:
            raise NotImplementedError(f"formulation {self.formulation}")
        self.W, self.V, self.L = (W, V, L) if modes else (None, None, None)
        self.kx, self.ky = (kx, ky) if modes else (None, None)


def scattering_identity(B, n, device):
    """[[0, I], [I, 0]] per batch item (khepri/alternative.py:220-232)."""
    S = torch.zeros((B, 2, 2, n, n), dtype=ops.CDT, device=device)
    eye = torch.eye(n, dtype=ops.CDT, device=device)
    S[:, 0, 1] = eye
    S[:, 1, 0] = eye
    return S


def stack_layers(pw, layers, mask, reverse=True, ws=None, infos=None):
    """Forward chain S_I * S_0 * S_1 * ... (copies kept where mask) and the
    reverse chain (khepri/layer.py:35-60).

    S_I * S = S * S_I = S exactly, so the products with the identity are not
    formed.  The reverse chain is only needed for fields (``reverse=False``
    skips it and returns None for it).  ``infos``, when a list, receives the
    device status word of every star product (first zero pivot of D or F,
    where the reference's numpy solve raises LinAlgError).
    """
    Stot = layers[0].S
    Sls = [Stot if mask[0] else None]
    for keep, layer in zip(mask[1:], layers[1:]):
        Stot, info = ops.redheffer(Stot, layer.S, ws=ws)
        if infos is not None:
            infos.append(info)
        Sls.append(Stot if keep else None)
    if not reverse:
        return Sls, None, Stot
    rmask = list(reversed(mask[1:]))
    Srs = []
    Srev = None
    for keep, layer in zip(rmask, reversed(layers[1:])):
        if keep:
            Srs.append(Srev if Srev is not None else scattering_identity(layer.S.shape[0], layer.S.shape[-1],
                                                                         layer.S.device))
        else:
            Srs.append(None)
        if Srev is None:
            Srev = layer.S
        else:
            Srev, info = ops.redheffer(layer.S, Srev, ws=ws)
            if infos is not None:
                infos.append(info)
    Srs.append(Srev)
    Srs = list(reversed(Srs))
    return Sls, Srs, Stot

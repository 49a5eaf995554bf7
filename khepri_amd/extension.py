"""Extended (Moire / twisted-bilayer) layers (mirrors khepri/extension.py).

An ``ExtendedLayer`` solves its base layer at kp + g for every harmonic g of
the other lattice -- here as ONE batched launch of Nb base solves per
(kp, wavelength) pair -- and scatters the base matrices into the extended
2Nb^2 x 2Nb^2 space with the reference's stride interleave.
"""
from __future__ import annotations

import numpy as np
import torch

from . import ops
from .layer import Layer


class ExtendedLayer:
    """khepri/extension.py:66-112."""

    def __init__(self, expansion, base) -> None:
        if expansion.expansion_lhs is base.expansion:
            self.mode = 1
            self.gs = expansion.expansion_rhs.g_vectors
        elif expansion.expansion_rhs is base.expansion:
            self.mode = 0
            self.gs = expansion.expansion_lhs.g_vectors
        else:
            raise NotImplementedError("Base layer expansion should be in the extented expansion.")
        self.expansion = expansion
        self.base = base
        self.depth = self.base.depth
        self.fields = self.base.fields
        self.S = self.W = self.V = self.L = None
        self.IC = 1.0
        self.info = None

    def _base_sources(self, kp, wl):
        """(B*Nb, 2) kp + g and (B*Nb,) wavelengths; base solve m of item b is b*Nb + m."""
        g = torch.as_tensor(np.ascontiguousarray(self.gs.T), dtype=torch.complex128, device=kp.device)
        Nb = g.shape[0]
        kpb = (kp[:, None, :] + g[None, :, :]).reshape(-1, 2).contiguous()
        wlb = wl[:, None].expand(-1, Nb).reshape(-1).contiguous()
        return kpb, wlb, Nb

    def solve(self, k_parallel, wavelength):
        from .device import default_device
        from .layer import as_source_batch
        kp, wl = as_source_batch(k_parallel, wavelength, default_device())
        self.solve_batch(kp, wl)

    def chain_item(self):
        return ("dense", self.S)

    def solve_batch(self, kp, wl, ws=None, rt_only=False):
        if isinstance(self.base, Layer):
            self.base.fields = self.fields
        kpb, wlb, Nb = self._base_sources(kp, wl)
        self.base.solve_batch(kpb, wlb)
        self.info = getattr(self.base, "info", None)
        self.S = ops.joint_subspace(self.base.S, Nb, self.mode, 4)
        if self.fields:
            self.W = ops.joint_subspace(self.base.W, Nb, self.mode, 1)
            self.V = ops.joint_subspace(self.base.V, Nb, self.mode, 1)
            self.L = ops.joint_subspace_diag(self.base.L, Nb, self.mode)
            # free-space modes of the base expansion at every kp + g (khepri/extension.py:114-126)
            kx, ky = ops.kvectors(self.base.expansion.g_device(kp.device), kpb, wlb)
            _, (W0, V0, _) = ops.analytic_layer("uniform", kx, ky, 1.0,
                                                torch.zeros(kx.shape[0], dtype=torch.float64, device=kp.device),
                                                modes=True)
            self.W0 = ops.joint_subspace(W0, Nb, self.mode, 1)
            self.V0 = ops.joint_subspace(V0, Nb, self.mode, 1)
        else:
            self.W = self.V = self.L = None
        self.IC = 1.0
        self.base.S = None


def extended_freespace(e, gs, wl, mode, k_parallel):
    """Host helper kept for API parity (khepri/extension.py:114-126)."""
    raise NotImplementedError("extended free-space modes are built on the device by ExtendedLayer.solve_batch")

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
        # the base lives on one side of the product expansion; its solves are
        # shifted by the harmonics of the other side (mode: which side is the
        # base; the left side wins when both sides are the base's expansion)
        sides = {id(expansion.expansion_rhs): (0, expansion.expansion_lhs),
                 id(expansion.expansion_lhs): (1, expansion.expansion_rhs)}
        if id(base.expansion) not in sides:
            raise NotImplementedError("the base layer's expansion must be one of the two factors "
                                      "of the extended expansion")
        self.mode, other = sides[id(base.expansion)]
        self.gs = other.g_vectors
        self.expansion, self.base = expansion, base
        self.depth, self.fields = base.depth, base.fields
        self.S = self.W = self.V = self.L = None
        self.IC, self.info = 1.0, None

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
        if not isinstance(self.base, Layer):
            # a Crystal as the base (khepri/extension.py:88-95): its whole stack
            # solved at every kp + g, one batched solve, its total S-matrix used
            if self.fields:
                raise NotImplementedError("fields of an extended layer need a Layer base (a Crystal base keeps "
                                          "no layer eigenspace)")
            self.base.set_source(wlb.cpu().numpy(), np.nan, np.nan, kp=kpb.cpu().numpy())
            self.base.solve(rt_only=False)
            self.info = None
            self.S = ops.joint_subspace(self.base.Stot.contiguous(), Nb, self.mode, 4)
            self.W = self.V = self.L = None
            self.IC = 1.0
            return
        self.base.solve_batch(kpb, wlb)
        self.info = getattr(self.base, "info", None)
        self.S = ops.joint_subspace(self.base.S, Nb, self.mode, 4)
        if self.fields:
            self.W = ops.joint_subspace(self.base.W, Nb, self.mode, 1)
            self.V = ops.joint_subspace(self.base.V, Nb, self.mode, 1)
            self.L = ops.joint_subspace_diag(self.base.L, Nb, self.mode)
            self.W0, self.V0 = _extended_freespace_batch(self.base.expansion, kpb, wlb, Nb, self.mode)
        else:
            self.W = self.V = self.L = None
        self.IC = 1.0
        self.base.S = None


def _extended_freespace_batch(e, kpb, wlb, Nb, mode):
    """Free-space modes W0, V0 of the base expansion e at the base sources kpb
    (B Nb, 2) / wlb (B Nb), scattered into the joint space: (B, n, n) each."""
    kx, ky = ops.kvectors(e.g_device(kpb.device), kpb, wlb)
    _, (W0, V0, _) = ops.analytic_layer("uniform", kx, ky, 1.0,
                                        torch.zeros(kx.shape[0], dtype=torch.float64, device=kpb.device),
                                        modes=True)
    return ops.joint_subspace(W0, Nb, mode, 1), ops.joint_subspace(V0, Nb, mode, 1)


def extended_freespace(e, gs, wl, mode, k_parallel):
    """Free-space eigenmodes of the extended basis (khepri/extension.py:114-126):
    the base expansion e's W0, V0 at k_parallel + g for every g of gs
    ((Nb, 2), the other lattice's harmonics), scattered into the joint space
    with the kind-`mode` interleave.  Returns device tensors (n, n)."""
    from .device import default_device
    dev = default_device()
    g = torch.as_tensor(np.ascontiguousarray(np.asarray(gs, dtype=np.float64).reshape(-1, 2)),
                        dtype=torch.complex128, device=dev)
    kp = torch.as_tensor(np.asarray(k_parallel, dtype=np.complex128).reshape(1, 2), device=dev)
    kpb = (kp + g).contiguous()
    wlb = torch.full((g.shape[0],), float(wl), dtype=torch.float64, device=dev)
    W0, V0 = _extended_freespace_batch(e, kpb, wlb, g.shape[0], mode)
    return W0[0], V0[0]

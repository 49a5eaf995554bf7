"""Many independent devices solved together (uneven harmonic counts and
layer counts).

``CrystalBatch([cl_1, ..., cl_K]).solve()`` returns the R and T of every
Crystal (each with one scalar source), as if each had run
``solve(); poynting_flux_end()`` on its own:

* devices are grouped by expansion (same harmonics and lattice); within a
  group every layer instance of every device is solved in ONE batched
  launch per stage (patterned layers carry their own convolution matrix);
* instances are laid out device-major, so within a (group, layer count,
  incidence/emergence media) subgroup layer j of every device is a strided
  view, and the R/T-only chain (khepri_amd.chain) runs batched over it.

The reference solves one Crystal per call (khepri/crystal.py:180-206); this
object only reorganises the same computations into batches.
"""
from __future__ import annotations

import os
from collections import defaultdict

import numpy as np
import torch

from . import chain, ops
from .device import default_device
from .extension import ExtendedLayer
from .layer import Formulation, Layer


def _is_gap_medium(eps):
    return np.isscalar(eps) and complex(eps) == 1


class CrystalBatch:
    def __init__(self, crystals, device=None, chunk=None):
        self.crystals = list(crystals)
        self.device = torch.device(device) if device is not None else None
        env = os.environ.get("KHX_BATCH_CHUNK")
        # patterned layers per launch; default: the eigensolver's resident
        # slots at the group's order, so every launch is one full round
        self.chunk = chunk or (int(env) if env else None)

    def _dev(self):
        return self.device or default_device()

    def solve(self):
        dev = self._dev()
        K = len(self.crystals)
        R = np.full(K, np.nan)
        T = np.full(K, np.nan)
        groups = defaultdict(list)
        for i, cl in enumerate(self.crystals):
            if not cl.source_defined or cl.batch != 1:
                raise ValueError(f"crystal {i}: set one scalar source before batching")
            key = (cl.expansion.N, cl.expansion.g_vectors.tobytes())
            groups[key].append(i)
        for idx in groups.values():
            r, t = self._solve_group(idx, dev)
            R[idx] = r
            T[idx] = t
        self.R, self.T = R, T
        return R, T

    # ------------------------------------------------------------------
    def _instances(self, idx):
        """Dense layer instances (crystal, layer) of a group, device-major,
        devices sorted by (layer count, media) so subgroups are contiguous."""
        per = []
        for i in idx:
            cl = self.crystals[i]
            items = []
            for name in cl.global_stacking:
                layer = cl.layers[name]
                if isinstance(layer, ExtendedLayer):
                    raise NotImplementedError("extended layers are batched through Crystal.solve")
                half = layer.formulation in (Formulation.HALF_SPACE_INC, Formulation.HALF_SPACE_TRN)
                if half and _is_gap_medium(layer.epsilon):
                    continue                      # identity S-matrix
                items.append(layer)
            per.append((len(items), complex(cl.epsi), complex(cl.epse), i, items))
        per.sort(key=lambda p: (p[0], p[1].real, p[1].imag, p[2].real, p[2].imag, p[3]))
        return per

    def _solve_group(self, idx, dev):
        per = self._instances(idx)
        cl0 = self.crystals[idx[0]]
        e = cl0.expansion
        N = e.N
        n = 2 * N
        g = e.g_device(dev)
        inst = [(i, layer) for (_, _, _, i, items) in per for layer in items]
        M = len(inst)
        kp = torch.as_tensor(np.stack([self.crystals[i]._kp[0] for i, _ in inst]), device=dev)
        wl = torch.as_tensor(np.array([self.crystals[i]._wl[0] for i, _ in inst]), device=dev)
        S_all = torch.empty((M, 2, 2, n, n), dtype=ops.CDT, device=dev)
        # ---- patterned instances: per-item convolution matrices, chunked
        pat = [k for k, (_, L) in enumerate(inst) if L.formulation in (Formulation.FFT, Formulation.ANALYTICAL)]
        chunk = self.chunk or ops.eig_slots(n)
        for c0 in range(0, len(pat), chunk):
            ks = pat[c0:c0 + chunk]
            C, IC = self._convmats([inst[k][1] for k in ks], e.pw, dev)
            kt = torch.as_tensor(ks, device=dev)
            kx, ky = ops.kvectors(g, kp[kt].contiguous(), wl[kt].contiguous())
            depth = torch.as_tensor([inst[k][1].depth for k in ks], dtype=torch.float64, device=dev)
            k0d = (2 * np.pi / wl[kt]) * depth
            S, _, info = ops.patterned_layer(C, IC, kx, ky, k0d.contiguous())
            ops.raise_on_info(info, "patterned layers")
            S_all[kt] = S
        # ---- analytic instances (uniform layers, half spaces of eps != 1), grouped by kind and eps
        ana = defaultdict(list)
        for k, (_, L) in enumerate(inst):
            if L.formulation == Formulation.UNIFORM:
                ana[("uniform", complex(L.epsilon))].append(k)
            elif L.formulation == Formulation.HALF_SPACE_INC:
                ana[("reflexion", complex(L.epsilon))].append(k)
            elif L.formulation == Formulation.HALF_SPACE_TRN:
                ana[("transmission", complex(L.epsilon))].append(k)
        for (kind, eps), ks in ana.items():
            kt = torch.as_tensor(ks, device=dev)
            kx, ky = ops.kvectors(g, kp[kt].contiguous(), wl[kt].contiguous())
            depth = torch.as_tensor([inst[k][1].depth for k in ks], dtype=torch.float64, device=dev)
            k0d = ((2 * np.pi / wl[kt]) * depth).contiguous()
            S, _ = ops.analytic_layer(kind, kx, ky, eps, k0d)
            S_all[kt] = S
        # ---- R/T-only chain per (layer count, media) subgroup
        out_r, out_t = {}, {}
        off = 0
        j = 0
        while j < len(per):
            m, ei, ee = per[j][0], per[j][1], per[j][2]
            j1 = j
            while j1 < len(per) and per[j1][:3] == (m, ei, ee):
                j1 += 1
            crys = [per[q][3] for q in range(j, j1)]
            G = len(crys)
            kpc = torch.as_tensor(np.stack([self.crystals[i]._kp[0] for i in crys]), device=dev)
            wlc = torch.as_tensor(np.array([self.crystals[i]._wl[0] for i in crys]), device=dev)
            tec = torch.as_tensor(np.array([self.crystals[i]._te[0] for i in crys]), device=dev)
            tmc = torch.as_tensor(np.array([self.crystals[i]._tm[0] for i in crys]), device=dev)
            c = ops.incident(N, kpc, wlc, tec, tmc, ei)
            if m == 0:
                sR, sT = torch.zeros_like(c), c
            else:
                V = S_all[off:off + G * m].view(G, m, 2, 2, n, n)
                sR, sT, infos = chain.rt_amplitudes([("dense", V[:, q]) for q in range(m)], c)
                for info in infos:
                    ops.raise_on_info(info, "star-product junction", lu=True)
            r, t, _, _ = ops.poynting_vec(sR, sT, g, kpc, wlc, ei, ee)
            for q, i in enumerate(crys):
                out_r[i] = r[q]
                out_t[i] = t[q]
            off += G * m
            j = j1
        Rg = torch.stack([out_r[i] for i in idx]).cpu().numpy()
        Tg = torch.stack([out_t[i] for i in idx]).cpu().numpy()
        return Rg, Tg

    @staticmethod
    def _convmats(layers, pw, dev):
        """(count, N, N) C and C^-1 of the pixmap layers (one convmat launch per pixmap shape) and
        island layers."""
        N = pw[0] * pw[1]
        C = torch.empty((len(layers), N, N), dtype=ops.CDT, device=dev)
        by_shape = defaultdict(list)
        for k, L in enumerate(layers):
            if L.formulation == Formulation.ANALYTICAL:   # closed-form island coefficients
                C[k] = L.convolution_matrices(dev)[0]
                continue
            eps = np.asarray(L.epsilon)
            by_shape[(eps.shape, np.iscomplexobj(eps))].append(k)
        for (shape, cplx), ks in by_shape.items():
            arr = np.stack([np.asarray(layers[k].epsilon, dtype=np.complex128 if cplx else np.float64) for k in ks])
            C[torch.as_tensor(ks, device=dev)] = ops.convmat(torch.as_tensor(arr, device=dev), pw)
        IC, info = ops.inverse(C)
        ops.raise_on_info(info, "inverse of the convolution matrices", lu=True)
        return C, IC

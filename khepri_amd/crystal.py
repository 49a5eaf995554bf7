"""User-facing device object (mirrors khepri/crystal.py ``Crystal``).

Same methods, arguments and return conventions as the reference, plus
batching: ``set_source`` accepts arrays of wavelengths / angles / in-plane
wave vectors (one solve per entry) and ``poynting_flux_end`` then returns
arrays.  All numerics run on the HIP device through libkhx; the Python code
only orchestrates.
"""
from __future__ import annotations

import cmath
import contextlib
import os
from types import SimpleNamespace

import numpy as np
import torch

from . import chain, ops
from .device import default_device
from .expansion import Expansion, compute_kplanar
from .extension import ExtendedLayer
from .layer import Layer, stack_layers


def _bytes_per_item(n, unique, length):
    # layer S-matrices, patterned-layer workspace, star workspace, chain
    return 16 * n * n * (4 * unique + 22 + 8 + 4 * (length + 2) + 4)


# Named lattices: rows are the lattice vectors in units of the pitch
# (khepri/crystal.py:60-72 accepts the same two names).
_NAMED_LATTICES = {
    "square": ((1.0, 0.0), (0.0, 1.0)),
    "hexagonal": ((np.sqrt(3) / 2, 0.5), (np.sqrt(3) / 2, -0.5)),
}

# The half spaces framing every stack: layer name, side, medium attribute.
_HALF_SPACES = (("Sref", "reflexion", "epsi"), ("Strans", "transmission", "epse"))


def _lattice_vectors(lattice, pitch):
    """Explicit 2 x 2 lattice vectors, or a named lattice scaled by pitch.
    An unknown name raises NotImplementedError, as the reference does."""
    if not isinstance(lattice, str):
        return lattice
    rows = _NAMED_LATTICES.get(lattice)
    if rows is None:
        raise NotImplementedError(f"lattice {lattice!r} is not available; use one of "
                                  f"{sorted(_NAMED_LATTICES)} or explicit lattice vectors")
    return pitch * np.asarray(rows)


class Crystal:
    def __init__(self, pw, lattice="square", lattice_pitch=1, void=False, epsi=1, epse=1, device=None) -> None:
        self.pw, self.a, self.void = pw, lattice_pitch, void
        self.epsi, self.epse = epsi, epse
        self.device = None if device is None else torch.device(device)
        self.lattice = _lattice_vectors(lattice, lattice_pitch)
        self.expansion = Expansion(pw, self.lattice)
        self.layers = {}
        self.source = None
        # solve results: full S-matrix chain (S, stacking matrices) or R/T only
        self.S, self._RT = None, None
        self.stacking_matrices, self.stacking_reverse_matrices = [], None
        self.stack_positions = []

    @classmethod
    def from_expansion(cls, expansion, **kwargs):
        """A crystal over a given (e.g. twisted) expansion; kwargs go to __init__."""
        crystal = cls(expansion.pw, **kwargs)
        crystal.expansion = expansion
        return crystal

    # ------------------------------------------------------------------ layers
    def add_layer_uniform(self, name, epsilon, depth):
        self.layers[name] = Layer.uniform(self.expansion, epsilon, depth)

    def add_layer_pixmap(self, name, epsilon, depth):
        self.layers[name] = Layer.pixmap(self.expansion, epsilon, depth)

    def add_pixmap_or_uniform(self, name, epsilon, depth):
        """A constant map becomes a uniform layer of that value."""
        first = epsilon.flat[0]
        if (epsilon == first).all():
            self.add_layer_uniform(name, first, depth)
        else:
            self.add_layer_pixmap(name, epsilon, depth)

    def add_layer_analytical(self, name, epsilon, epsilon_host, depth):
        self.layers[name] = Layer.analytical(self.expansion, epsilon, epsilon_host, depth)

    def add_layer(self, name, layer, extended=False):
        self.layers[name] = ExtendedLayer(self.expansion, layer) if extended else layer

    def set_device(self, layers_stack, fields_mask=False):
        """The user's stack framed by the incidence and emergence half spaces
        (unless void), which are created on first use; fields_mask (one flag,
        or one per user layer) marks the layers whose fields are kept
        (khepri/crystal.py:131-164)."""
        self.device_stack = list(layers_stack)
        framed = not self.void
        self.global_stacking = (["Sref"] if framed else []) + self.device_stack + (["Strans"] if framed else [])
        for name, side, medium in _HALF_SPACES:
            if name in self.global_stacking and name not in self.layers:
                half = Layer.half_infinite(self.expansion, side, getattr(self, medium))
                half.fields = True
                self.layers[name] = half
        if isinstance(fields_mask, bool):
            fields_mask = [fields_mask] * len(self.device_stack)
        self.stack_retain_mask = [True, *fields_mask, True]
        self._fields_requested = any(fields_mask)
        for name, keep in zip(self.global_stacking, self.stack_retain_mask):
            layer = self.layers[name]
            layer.fields |= keep
            base = getattr(layer, "base", None)
            if base is not None:
                base.fields |= keep

    @property
    def depth(self):
        return sum(self.layers[name].depth for name in self.device_stack)

    @property
    def source_defined(self):
        return self.source is not None

    @property
    def solved(self):
        return self.S is not None or self._RT is not None

    @property
    def zmax(self):
        return self.stack_positions[-2]

    # ------------------------------------------------------------------ source
    def set_source(self, wavelength, te=1.0, tm=1.0, theta=0.0, phi=0.0, kp=None):
        """khepri/crystal.py:345-360; every argument may be an array (batch).

        kp is one in-plane wave vector (kx, ky), shape (2,), or a batch of
        them, shape (B, 2) with one (kx, ky) per row.  Any other shape raises
        ValueError: a (2, B) array such as a flattened Brillouin-zone grid
        must be transposed first (a (2, 2) array is read as two rows)."""
        wl = np.atleast_1d(np.asarray(wavelength, dtype=np.float64))
        self._scalar = np.ndim(wavelength) == 0 and (kp is None or np.ndim(kp) == 1) \
            and np.ndim(theta) == 0 and np.ndim(phi) == 0
        if kp is not None:
            kpa = np.asarray(kp, dtype=np.complex128)
            if kpa.ndim not in (1, 2) or kpa.shape[-1] != 2:
                raise ValueError(f"kp must have shape (2,) or (B, 2), got {kpa.shape}")
            kpa = kpa.reshape(-1, 2)
        else:
            th = np.atleast_1d(np.asarray(theta, dtype=np.float64))
            ph = np.atleast_1d(np.asarray(phi, dtype=np.float64))
            B0 = max(len(wl), len(th), len(ph))
            wlb, thb, phb = (np.broadcast_to(v, (B0,)) for v in (wl, th, ph))
            kpa = np.stack([compute_kplanar(self.epsi, w, t, p) for w, t, p in zip(wlb, thb, phb)])
        B = max(len(wl), kpa.shape[0])
        wl = np.broadcast_to(wl, (B,)).copy()
        kpa = np.broadcast_to(kpa, (B, 2)).copy()
        te_a = np.broadcast_to(np.asarray(te, dtype=np.complex128), (B,)).copy()
        tm_a = np.broadcast_to(np.asarray(tm, dtype=np.complex128), (B,)).copy()
        self.source = SimpleNamespace()
        self.source.te, self.source.tm, self.source.theta, self.source.phi = te, tm, theta, phi
        self.source.wavelength = wavelength
        self._wl, self._kp, self._te, self._tm = wl, kpa, te_a, tm_a
        self.batch = B
        if self._scalar:
            self.kp = (kpa[0, 0], kpa[0, 1]) if kp is None else tuple(kp)
            self.k0 = 2 * np.pi / wavelength
            kxi, kyi = self.kp
            self.kzi = np.conj(cmath.sqrt(self.k0 ** 2 * self.epsi - kxi ** 2 - kyi ** 2))
        else:
            self.kp = kpa
            self.k0 = 2 * np.pi / wl
            self.kzi = np.conj(np.sqrt((self.k0 ** 2 * self.epsi - kpa[:, 0] ** 2 - kpa[:, 1] ** 2).astype(complex)))
        self.S = None
        self._RT = None

    # ------------------------------------------------------------------ solve
    def _dev(self):
        return self.device or default_device()

    def _chunk(self, B):
        env = os.environ.get("KHX_CHUNK")
        if env:
            return max(1, int(env))
        budget = float(os.environ.get("KHX_MEM_BUDGET_GB", "48")) * 2 ** 30
        n = 2 * self.expansion.N
        per = _bytes_per_item(n, len(set(self.global_stacking)), len(self.global_stacking))
        chunk = int(max(1, min(B, budget // per)))
        if chunk < B and self._has_patterned():
            # every chunk's eigensolver launch is a whole number of rounds of
            # resident workgroups: a partly filled last round costs a full one
            slots = ops.eig_slots(n)
            if chunk >= slots:
                chunk = chunk // slots * slots
        return chunk

    def _has_patterned(self):
        return any(getattr(self.layers[name], "patterned", False)
                   for name in set(self.global_stacking) if name in self.layers)

    def solve(self, rt_only=None):
        """khepri/crystal.py:180-206.

        rt_only: compute only what poynting_flux_end needs (Stot11 c1p and
        Stot21 c1p, see khepri_amd.chain) instead of every S-matrix of the
        chain.  Default: on when no fields are requested and the batch does
        not fit one memory chunk -- then the batch is streamed chunk by chunk
        and only R/T (total and per order) are kept."""
        assert self.source_defined, "Call set_source before solving."
        dev = self._dev()
        names = list(dict.fromkeys(self.global_stacking))   # unique, in stack order
        stacked = [self.layers[name] for name in self.global_stacking]
        self.stack_positions = list(np.cumsum([layer.depth for layer in stacked]))
        self.stack_positions[-1] = np.inf
        if not self.void:
            self.stack_positions.insert(0, -np.inf)

        B = self.batch
        chunk = self._chunk(B)
        static = getattr(self, "_static_sources", None)
        if static is not None:   # device buffers of a captured graph (khepri_amd.graph)
            kp_all, wl_all, te_all, tm_all = static
        else:
            kp_all = torch.as_tensor(self._kp, device=dev)
            wl_all = torch.as_tensor(self._wl, device=dev)
            te_all = torch.as_tensor(self._te, device=dev)
            tm_all = torch.as_tensor(self._tm, device=dev)
        self._sources_dev = (kp_all, wl_all, te_all, tm_all)
        want_reverse = getattr(self, "_fields_requested", False)
        if rt_only is None:
            rt_only = (not want_reverse) and chunk < B
        if rt_only and want_reverse:
            raise ValueError("fields need the full S-matrix chain (rt_only=False)")
        self.S = None
        self._RT = None
        if not rt_only and chunk >= B:
            Layer.solve_batch_many([self.layers[name] for name in names], kp_all, wl_all)
            self._check_layers(names)
            infos = []
            self.stacking_matrices, self.stacking_reverse_matrices, self.Stot = stack_layers(
                self.expansion.pw, stacked, self.stack_retain_mask, reverse=want_reverse, infos=infos)
            for info in infos:
                self._check_info(info, "star product", lu=True)
            self.S = self.Stot
            return
        if want_reverse:
            raise ValueError("fields need the whole batch resident; lower the batch size (KHX_CHUNK)")
        g = self.expansion.g_device(dev)
        N = self.expansion.N
        R = torch.empty(B, dtype=torch.float64, device=dev)
        T = torch.empty_like(R)
        Ro = torch.empty((B, N), dtype=torch.float64, device=dev)
        To = torch.empty_like(Ro)
        out = (R, T, Ro, To)
        # the status words of every chunk are checked once, after the last chunk
        # is enqueued (a check is a host synchronisation, which would leave the
        # GPU idle while the host enqueues the next chunk); inside an outer
        # deferral (graph capture) they go to its list instead
        own = not ops._DEFERRED
        sink = []
        with ops.deferred_status(sink) if own else contextlib.nullcontext():
            for c0 in range(0, B, chunk):
                self._solve_chunk(slice(c0, min(B, c0 + chunk)), names, stacked, rt_only, g, N, out)
        if own:
            ops.check_status(sink)
        self._RT = (R, T, Ro, To)

    def _solve_chunk(self, sl, names, stacked, rt_only, g, N, out):
        """Sources sl of the batch: layer solves, chain, R/T into out[...][sl]."""
        kp_all, wl_all, te_all, tm_all = self._sources_dev
        kp, wl = kp_all[sl], wl_all[sl]
        Layer.solve_batch_many([self.layers[name] for name in names], kp, wl, rt_only=rt_only)
        self._check_layers(names)
        if rt_only:
            c = ops.incident(N, kp, wl, te_all[sl], tm_all[sl], self.epsi)
            sR, sT, infos = chain.rt_amplitudes([layer.chain_item() for layer in stacked], c)
            for info in infos:
                self._check_info(info, "star-product junction", lu=True)
            r, t, ro, to = ops.poynting_vec(sR, sT, g, kp, wl, self.epsi, self.epse, orders=True)
        else:
            infos = []
            _, _, Stot = stack_layers(self.expansion.pw, stacked, self.stack_retain_mask, reverse=False,
                                      infos=infos)
            for info in infos:
                self._check_info(info, "star product", lu=True)
            r, t, ro, to = ops.poynting(Stot, g, kp, wl, te_all[sl], tm_all[sl], self.epsi, self.epse,
                                        orders=True)
        R, T, Ro, To = out
        R[sl], T[sl], Ro[sl], To[sl] = r, t, ro, to
        for name in names:
            self.layers[name].S = None

    def _check_info(self, info, what, lu=False):
        if info is not None and os.environ.get("KHX_CHECK_INFO", "1") == "1":
            ops.raise_on_info(info, what, lu=lu)

    def _check_layers(self, names):
        for name in names:
            self._check_info(getattr(self.layers[name], "info", None), f"layer {name}")

    # ------------------------------------------------------------------ outputs
    def poynting_flux_end(self, only_total=True):
        """(R, T) -- khepri/crystal.py:363-396."""
        assert self.solved, "Call solve first"
        if self._RT is not None:
            R, T, Ro, To = (x.cpu().numpy() for x in self._RT)
        else:
            kp, wl, te, tm = self._sources_dev
            R, T, Ro, To = ops.poynting(self.Stot, self.expansion.g_device(self._dev()), kp, wl, te, tm,
                                        self.epsi, self.epse, orders=not only_total)
            R, T = R.cpu().numpy(), T.cpu().numpy()
            if not only_total:
                Ro, To = Ro.cpu().numpy(), To.cpu().numpy()
        if not only_total:
            if self._scalar:
                return (R[0], Ro[0]), (T[0], To[0])
            return (R, Ro), (T, To)
        if self._scalar:
            return R[0], T[0]
        return R, T

    # fields are provided by khepri_amd.fields (bound below)


from . import fields as _fields  # noqa: E402

Crystal.locate_layer = _fields.locate_layer
Crystal.get_source_as_field_vectors = _fields.get_source_as_field_vectors
Crystal.fields_coords_xy = _fields.fields_coords_xy
Crystal.fields_volume = _fields.fields_volume
Crystal._fourier_fields = _fields.fourier_fields


class Multilayer(Crystal):
    def __init__(self, epsi=1, epse=1):
        super().__init__((1, 1), lattice="square", lattice_pitch=1, void=False, epsi=epsi, epse=epse)

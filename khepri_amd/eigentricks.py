"""Band-structure tool (khepri/eigentricks.py) on the GPU.

Bloch modes of one period from its S-matrix.  S is one (4ng x 4ng) matrix in
the (+, -) block layout of the reference's band-structure code, or a batch
(B, 4ng, 4ng) of them (numpy or a device tensor): rows / columns 0:2ng are
the forward amplitudes (p), 2ng:4ng the backward ones (m); S[p,p] is the
forward transmission, S[m,p] the reflection from the left, S[p,m] the
reflection from the right and S[m,m] the backward transmission.
``scattering_matrix(crystal)`` assembles that layout from a solved
``Crystal`` (its stack between the two half spaces is then the period).

The pencil (Sl, Sr) is solved by libkhx (``khx_scattering_eig``) through its
Cayley transform instead of the reference's QZ call: same eigenvalues and
eigenvectors (unit 2-norm columns; the phase of each column differs from
scipy's, as between any two eigensolvers), with the propagating Bloch
factors |w| = 1 at full accuracy.  A batch of S-matrices -- a band diagram's
whole (frequency, k) sweep -- is solved in one call.
"""
from __future__ import annotations

import numpy as np
import torch

from . import _lib
from ._lib import check
from .device import default_device

KHX_INFO_NAN_INPUT = 0x40000000
# shifts on the unit circle for the Cayley transform; the second is used for
# the items where the first happens to be a Bloch factor
_SHIFTS = (np.exp(1j * 1.0), np.exp(1j * 2.5))
_CDT = torch.complex128


def _to_dev(S):
    """(device tensor (B, 4ng, 4ng), batched?, was_numpy?)."""
    numpy_in = not isinstance(S, torch.Tensor)
    t = torch.as_tensor(np.asarray(S), dtype=_CDT) if numpy_in else S.to(_CDT)
    if t.device.type != "cuda":
        t = t.to(default_device())
    batched = t.dim() == 3
    if not batched:
        t = t[None]
    if t.dim() != 3 or t.shape[1] != t.shape[2] or t.shape[1] % 4 != 0:
        raise ValueError(f"S must be (4ng, 4ng) or (B, 4ng, 4ng), got {tuple(np.shape(S))}")
    return t.contiguous(), batched, numpy_in


def _out(t, batched, numpy_in):
    t = t if batched else t[0]
    return t.cpu().numpy() if numpy_in else t


def _stream():
    return torch.cuda.current_stream().cuda_stream


def scattering_splitlr(S):
    """(Sl, Sr) = ([[Spp, 0], [Smp, -I]], [[I, -Spm], [0, -Smm]]).
    khepri/eigentricks.py:5-22."""
    lib = _lib.load()
    t, batched, numpy_in = _to_dev(S)
    B, n2 = t.shape[0], t.shape[1]
    Sl = torch.empty_like(t)
    Sr = torch.empty_like(t)
    check(lib.khx_scattering_splitlr(n2 // 4, t.data_ptr(), B, Sl.data_ptr(), Sr.data_ptr(), _stream()),
          "khx_scattering_splitlr")
    return _out(Sl, batched, numpy_in), _out(Sr, batched, numpy_in)


def scattering_det(S):
    """det(Spp - Spm Smm^-1 Smp) det(Smm) per S-matrix; raises
    numpy.linalg.LinAlgError when Smm is singular, like the reference's inv.
    khepri/eigentricks.py:24-28."""
    lib = _lib.load()
    t, batched, numpy_in = _to_dev(S)
    B, n2 = t.shape[0], t.shape[1]
    ng = n2 // 4
    det = torch.empty(B, dtype=_CDT, device=t.device)
    info = torch.empty(B, dtype=torch.int32, device=t.device)
    ws = torch.empty(max(1, lib.khx_scattering_det_ws_bytes(ng, B)), dtype=torch.uint8, device=t.device)
    check(lib.khx_scattering_det(ng, t.data_ptr(), B, det.data_ptr(), info.data_ptr(), ws.data_ptr(), _stream()),
          "khx_scattering_det")
    bad = (info != 0).nonzero()
    if len(bad):
        raise np.linalg.LinAlgError(f"Singular matrix (S[m,m] of item {int(bad[0, 0])})")
    if not batched:
        return complex(det[0]) if numpy_in else det[0]
    return det.cpu().numpy() if numpy_in else det


def _eig_dev(t):
    """(w, v, info) of a device batch, with a second shift for the items the
    first one hits."""
    w, v, info = _eig_dev_shift(t, _SHIFTS[0])
    retry = (info < 0).nonzero()[:, 0]
    if len(retry):
        w2, v2, i2 = _eig_dev_shift(t[retry].contiguous(), _SHIFTS[1])
        w[retry], v[retry], info[retry] = w2, v2, i2
    return w, v, info


def _eig_dev_shift(t, s):
    lib = _lib.load()
    B, n2 = t.shape[0], t.shape[1]
    ng = n2 // 4
    w = torch.empty((B, n2), dtype=_CDT, device=t.device)
    v = torch.empty((B, n2, n2), dtype=_CDT, device=t.device)
    info = torch.empty(B, dtype=torch.int32, device=t.device)
    ws = torch.empty(max(1, lib.khx_scattering_eig_ws_bytes(ng, B)), dtype=torch.uint8, device=t.device)
    check(lib.khx_scattering_eig(ng, t.data_ptr(), B, s.real, s.imag, w.data_ptr(), v.data_ptr(), info.data_ptr(),
                                 ws.data_ptr(), _stream()), "khx_scattering_eig")
    return w, v, info


def scattering_eigenvalues(S, dos=False):
    """Bloch factors w and unit-norm modes v (column j for w[j]) of the pencil
    Sl v = w Sr v; with dos also scattering_det(S).  One S: None when it holds
    a NaN, as in the reference.  A batch (B, 4ng, 4ng): arrays (B, 4ng) and
    (B, 4ng, 4ng), NaN for the items that hold a NaN.  Raises
    numpy.linalg.LinAlgError where the eigensolver does not converge.
    khepri/eigentricks.py:29-40."""
    t, batched, numpy_in = _to_dev(S)
    w, v, info = _eig_dev(t)
    nan_items = info == KHX_INFO_NAN_INPUT
    failed = ((info != 0) & ~nan_items).nonzero()
    if len(failed):
        raise np.linalg.LinAlgError(f"eigen-solver did not converge (item {int(failed[0, 0])})")
    if not batched and bool(nan_items[0]):
        return None
    out = (_out(w, batched, numpy_in), _out(v, batched, numpy_in))
    if dos:
        if batched:
            ok = (~nan_items).nonzero()[:, 0]
            det = torch.full((t.shape[0],), complex("nan+nanj"), dtype=_CDT, device=t.device)
            if len(ok):
                det[ok] = scattering_det(t[ok].contiguous())
            out = out + ((det.cpu().numpy() if numpy_in else det),)
        else:
            out = out + (scattering_det(S),)
    return out


def on_shell(eigenvalues, tol=1e-10):
    """|w| == 1 within tol.  khepri/eigentricks.py:42-43."""
    return np.isclose(np.abs(eigenvalues), 1.0, rtol=0.0, atol=tol)


def band_structure(S):
    """Real parts in (0, 1) of the Bloch factors, kept where on shell
    (khepri/eigentricks.py:47-55); None for an S holding a NaN.  A batch gives
    the list of the per-item results."""
    t, batched, _ = _to_dev(S)
    w, _, info = _eig_dev(t)
    failed = ((info != 0) & (info != KHX_INFO_NAN_INPUT)).nonzero()
    if len(failed):
        raise np.linalg.LinAlgError(f"eigen-solver did not converge (item {int(failed[0, 0])})")
    wr = w.real.cpu().numpy()
    nan_items = (info == KHX_INFO_NAN_INPUT).cpu().numpy()
    res = []
    for b in range(wr.shape[0]):
        if nan_items[b]:
            res.append(None)
            continue
        x = wr[b]
        x = x[x > 0]
        x = x[x < 1]
        res.append(x[on_shell(x)])
    return res if batched else res[0]


def scattering_matrix(crystal):
    """(B, 4ng, 4ng) device tensor in the layout above from a Crystal solved
    with its full S-matrix chain (``solve(rt_only=False)``): with the stack's
    Redheffer blocks [[S11, S12], [S21, S22]] (S11 reflection, S21
    transmission of light coming from the left), S[p,p] = S21, S[p,m] = S22,
    S[m,p] = S11, S[m,m] = S12.  No reference counterpart: the reference's
    band diagrams assemble S with its legacy transfer-matrix code."""
    St = getattr(crystal, "Stot", None)
    if St is None or crystal.S is None:
        raise ValueError("solve the crystal with its full S-matrix chain first (solve(rt_only=False))")
    top = torch.cat([St[:, 1, 0], St[:, 1, 1]], dim=-1)
    bot = torch.cat([St[:, 0, 0], St[:, 0, 1]], dim=-1)
    return torch.cat([top, bot], dim=-2).contiguous()

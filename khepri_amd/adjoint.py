"""Adjoint gradient of the Poynting fluxes with respect to pixel permittivities.

The reference computes R and T only (khepri/crystal.py:363-396); it has no
derivative.  This module adds reverse-mode differentiation of the whole solve
path, written by hand and run in libkhx:

    eps (pixmap) -> C = Toeplitz(fft2(eps)) (khepri/tools.py:33-56), IC = C^-1
    -> P, Q -> (mu, W) = eig(P Q), lam = sqrt(mu), V = Q W / lam
    -> layer S-matrix (khepri/alternative.py:158-195)
    -> Redheffer chain (khepri/alternative.py:19-30, khepri/layer.py:35-60)
    -> R, T (khepri/alternative.py:235-245)

Backward pass (``flux_gradient``):
  1. the forward quantities are recomputed, not stored: each unique layer's
     S-matrix and modes (one batched launch per layer), then the left-to-right
     chain products P_k = P_{k-1} * S_k;
  2. the Poynting adjoint gives sRbar, sTbar, hence Stot11bar = sRbar c^H and
     Stot21bar = sTbar c^H (c: incident vector);
  3. ``khx_redheffer_bwd`` walks the chain back to every layer's Sbar;
     instances of one named layer add up (the layer adjoint is linear in Sbar);
  4. ``khx_patterned_layer_bwd`` maps Sbar to Cbar (scattering-matrix
     intermediates recomputed, eigen adjoint with Lorentzian broadening
     E / (|E|^2 + eta^2) of 1/conj(mu_j - mu_i), C^-1 differentiated);
  5. ``khx_convmat_adjoint`` maps Cbar to the pixels (Toeplitz gather and
     partial DFT transposed).

``FluxFunction`` wraps it as a torch.autograd.Function over pixmap tensors.
"""
from __future__ import annotations

import os

import numpy as np
import torch

from . import ops
from .extension import ExtendedLayer
from .layer import Formulation

_ONE = complex(1.0)


def _weights(w, B, dev):
    if w is None:
        return None
    if isinstance(w, torch.Tensor):
        t = w.detach().to(device=dev, dtype=torch.float64).reshape(-1)
    else:
        t = torch.as_tensor(np.asarray(w, dtype=np.float64).reshape(-1), device=dev)
    if t.numel() == 1 and B > 1:
        t = t.expand(B)
    if t.numel() != B:
        raise ValueError(f"weights: expected 1 or {B} values, got {t.numel()}")
    return t.contiguous()


def _bytes_per_source(n, unique, length):
    # layer S + modes, chain products and their adjoints, layer / star backward workspaces
    return 16 * n * n * (7 * unique + 8 * length + 46 + 16)


def _gap(layer):
    return np.isscalar(layer.epsilon) and complex(layer.epsilon) == 1


def flux_gradient(crystal, wT=1.0, wR=0.0, wrt=None, reduce=False, eta_rel=1e-12, chunk=None):
    """Gradient of  sum_b wR[b] R[b] + wT[b] T[b]  with respect to the pixmap
    permittivity of each layer named in ``wrt`` (default: every pixmap layer).

    Returns (R, T, grads): R, T numpy (B,) as poynting_flux_end would give, and
    grads[name] a float64 tensor on the device of shape (B, nx, ny) -- one map
    per source, d(wR_b R_b + wT_b T_b)/d eps -- or (nx, ny) summed over the
    sources when ``reduce``.  Real pixmaps give the derivative with respect to
    the real permittivity; complex pixmaps give dL/dRe + i dL/dIm.
    """
    R, T, grads = _flux_gradient_device(crystal, wT, wR, wrt, reduce, eta_rel, chunk)
    return R.cpu().numpy(), T.cpu().numpy(), grads


def _flux_gradient_device(crystal, wT=1.0, wR=0.0, wrt=None, reduce=False, eta_rel=1e-12, chunk=None):
    """flux_gradient with R and T left on the device (no host sync: the
    sequence can be captured in a graph; khepri_amd.graph.GradientGraph)."""
    cl = crystal
    assert cl.source_defined, "Call set_source before differentiating."
    dev = cl._dev()
    names = list(dict.fromkeys(cl.global_stacking))
    for name in names:
        if isinstance(cl.layers[name], ExtendedLayer):
            raise NotImplementedError("gradients through extended (twisted) layers are not provided")
    pixmaps = [nm for nm in names if cl.layers[nm].formulation == Formulation.FFT]
    wrt = list(pixmaps if wrt is None else wrt)
    for nm in wrt:
        if nm not in pixmaps:
            raise ValueError(f"layer {nm!r} is not a pixmap layer of the device")
    B = cl.batch
    N = cl.expansion.N
    n = 2 * N
    g = cl.expansion.g_device(dev)
    static = getattr(cl, "_static_sources", None)
    if static is not None:   # device buffers of a captured graph (khepri_amd.graph)
        kp_all, wl_all, te_all, tm_all = static
    else:
        kp_all = torch.as_tensor(cl._kp, device=dev)
        wl_all = torch.as_tensor(cl._wl, device=dev)
        te_all = torch.as_tensor(cl._te, device=dev)
        tm_all = torch.as_tensor(cl._tm, device=dev)
    wT_all = _weights(wT, B, dev)
    wR_all = _weights(wR, B, dev)
    if chunk is None:
        env = os.environ.get("KHX_GRAD_CHUNK")
        budget = float(os.environ.get("KHX_MEM_BUDGET_GB", "48")) * 2 ** 30
        chunk = int(env) if env else max(1, int(budget // _bytes_per_source(n, len(names), len(cl.global_stacking))))
    shapes = {nm: np.asarray(cl.layers[nm].epsilon).shape for nm in wrt}
    cplx = {nm: np.iscomplexobj(cl.layers[nm].epsilon) for nm in wrt}
    if reduce:
        grads = {nm: torch.zeros(shapes[nm], dtype=torch.complex128 if cplx[nm] else torch.float64, device=dev)
                 for nm in wrt}
    else:
        grads = {nm: torch.empty((B,) + shapes[nm], dtype=torch.complex128 if cplx[nm] else torch.float64,
                                 device=dev) for nm in wrt}
    R = torch.empty(B, dtype=torch.float64, device=dev)
    T = torch.empty(B, dtype=torch.float64, device=dev)
    for c0 in range(0, B, chunk):
        sl = slice(c0, min(B, c0 + chunk))
        r, t, gch = _gradient_chunk(cl, names, wrt, g, N, kp_all[sl], wl_all[sl], te_all[sl], tm_all[sl],
                                    None if wR_all is None else wR_all[sl], None if wT_all is None else wT_all[sl],
                                    eta_rel, reduce)
        R[sl], T[sl] = r, t
        for nm in wrt:
            if reduce:
                grads[nm] += gch[nm]
            else:
                grads[nm][sl] = gch[nm]
    return R, T, grads


def _layers_forward(cl, names, kx, ky, k0):
    """_layer_forward for every named layer; the patterned ones in one fused
    call when their eigenproblems would not fill the GPU on their own."""
    pat = [nm for nm in names if cl.layers[nm].patterned]
    B, N = kx.shape
    out = {}
    if len(pat) >= 2 and B * len(pat) <= ops.eig_slots(2 * N) and os.environ.get("KHX_FUSE_LAYERS", "1") == "1":
        mats = [cl.layers[nm].convolution_matrices(kx.device) for nm in pat]
        k0ds = [(k0 * cl.layers[nm].depth).contiguous() for nm in pat]
        res = ops.patterned_layers_fused([m[0] for m in mats], [m[1] for m in mats], kx, ky, k0ds, modes=True,
                                         cache=True)
        for nm, (C, IC), k0d, (S, (W, V, lam), info, cache) in zip(pat, mats, k0ds, res):
            ops.raise_on_info(info, "patterned layer")
            out[nm] = ("dense", S, dict(C=C, IC=IC, W=W, V=V, lam=lam, k0d=k0d, cache=cache))
    for nm in names:
        if nm not in out:
            out[nm] = _layer_forward(cl.layers[nm], kx, ky, k0)
    return out


def _layer_forward(layer, kx, ky, k0):
    """("identity", None) | ("dense", S, extras) for one layer over a batch."""
    k0d = (k0 * layer.depth).contiguous()
    f = layer.formulation
    if f in (Formulation.FFT, Formulation.ANALYTICAL):
        C, IC = layer.convolution_matrices(kx.device)
        B, N = kx.shape
        cache = torch.empty((B, 9, 2 * N, 2 * N), dtype=ops.CDT, device=kx.device)
        S, (W, V, lam), info = ops.patterned_layer(C, IC, kx, ky, k0d, modes=True, cache=cache)
        ops.raise_on_info(info, "patterned layer")
        return ("dense", S, dict(C=C, IC=IC, W=W, V=V, lam=lam, k0d=k0d, cache=cache))
    if f in (Formulation.HALF_SPACE_INC, Formulation.HALF_SPACE_TRN):
        if _gap(layer):
            return ("identity", None, None)
        kind = "reflexion" if f == Formulation.HALF_SPACE_INC else "transmission"
        S, _ = ops.analytic_layer(kind, kx, ky, layer.epsilon, None)
        return ("dense", S, None)
    if f == Formulation.UNIFORM:
        S, _ = ops.analytic_layer("uniform", kx, ky, layer.epsilon, k0d)
        return ("dense", S, None)
    raise NotImplementedError(f"layer formulation {f} is not differentiable here")


def _gradient_chunk(cl, names, wrt, g, N, kp, wl, te, tm, wR, wT, eta_rel, reduce):
    dev = kp.device
    Bc = kp.shape[0]
    n = 2 * N
    kx, ky = ops.kvectors(g, kp, wl)
    k0 = 2 * np.pi / wl
    # 1. layer S-matrices and modes (recomputed)
    fwd = _layers_forward(cl, names, kx, ky, k0)
    seq = [nm for nm in cl.global_stacking if fwd[nm][0] == "dense"]
    c = ops.incident(N, kp, wl, te, tm, cl.epsi)
    # 2. chain products P_k = P_{k-1} * S_k
    if not seq:
        raise ValueError("the device has no scattering layer")
    Ps = [fwd[seq[0]][1]]
    for nm in seq[1:]:
        P, info = ops.redheffer(Ps[-1], fwd[nm][1])
        ops.raise_on_info(info, "star product", lu=True)
        Ps.append(P)
    Stot = Ps[-1]
    sR = ops.matmul(Stot[:, 0, 0], c[:, :, None])[..., 0]
    sT = ops.matmul(Stot[:, 1, 0], c[:, :, None])[..., 0]
    r, t, _, _ = ops.poynting_vec(sR, sT, g, kp, wl, cl.epsi, cl.epse)
    # 3. flux adjoint -> Stot adjoint (rank one in the 11 and 21 blocks)
    sRbar, sTbar = ops.poynting_vec_adjoint(sR, sT, g, kp, wl, cl.epsi, cl.epse, wR, wT)
    G = torch.zeros((Bc, 2, 2, n, n), dtype=ops.CDT, device=dev)
    ops.matmul(sRbar[:, :, None], c[:, :, None], C=G[:, 0, 0], opb="C")
    ops.matmul(sTbar[:, :, None], c[:, :, None], C=G[:, 1, 0], opb="C")
    # 4. chain backward -> per-instance Sbar, summed per layer name
    Sbar = {}
    for k in range(len(seq) - 1, 0, -1):
        Pbar, SkBar, info = ops.redheffer_bwd(Ps[k - 1], fwd[seq[k]][1], G)
        ops.raise_on_info(info, "star product adjoint", lu=True)
        _accumulate(Sbar, seq[k], SkBar)
        G = Pbar
    _accumulate(Sbar, seq[0], G)
    # 5. layer adjoints -> Cbar -> pixels (the layers together in one call
    #    when one layer's batch would leave the GPU mostly idle)
    out = {}
    live = [nm for nm in wrt if nm in Sbar]
    Cbars = {}
    if len(live) >= 2 and Bc * len(live) <= ops.eig_slots(n) and os.environ.get("KHX_FUSE_LAYERS", "1") == "1":
        L = len(live)
        exs = [fwd[nm][2] for nm in live]
        Cf = torch.stack([ex["C"].expand(Bc, N, N) for ex in exs]).reshape(L * Bc, N, N)
        ICf = torch.stack([ex["IC"].expand(Bc, N, N) for ex in exs]).reshape(L * Bc, N, N)
        Cbar_all, info = ops.patterned_layer_bwd(
            Cf, ICf, kx.repeat(L, 1), ky.repeat(L, 1), torch.cat([ex["k0d"].reshape(Bc) for ex in exs]),
            torch.cat([ex["W"] for ex in exs]), torch.cat([ex["V"] for ex in exs]),
            torch.cat([ex["lam"] for ex in exs]), torch.cat([Sbar[nm] for nm in live]), eta_rel,
            cache=torch.cat([ex["cache"] for ex in exs]))
        ops.raise_on_info(info, "layer adjoint")
        for l, nm in enumerate(live):
            Cbars[nm] = Cbar_all[l * Bc:(l + 1) * Bc]
    for nm in wrt:
        ex = fwd[nm][2]
        if nm not in Sbar:
            shape = np.asarray(cl.layers[nm].epsilon).shape
            out[nm] = torch.zeros(((Bc,) if not reduce else ()) + shape, device=dev,
                                  dtype=torch.complex128 if np.iscomplexobj(cl.layers[nm].epsilon) else torch.float64)
            continue
        if nm in Cbars:
            Cbar = Cbars[nm]
        else:
            Cbar, info = ops.patterned_layer_bwd(ex["C"], ex["IC"], kx, ky, ex["k0d"], ex["W"], ex["V"], ex["lam"],
                                                 Sbar[nm], eta_rel, cache=ex["cache"])
            ops.raise_on_info(info, "layer adjoint")
        eps = np.asarray(cl.layers[nm].epsilon)
        if reduce:
            ones = torch.ones((1, Bc), dtype=ops.CDT, device=dev)
            Cbar = ops.matmul(ones, Cbar.reshape(1, Bc, N * N)).reshape(1, N, N)
        m = ops.convmat_adjoint(Cbar, cl.expansion.pw, eps.shape, np.iscomplexobj(eps))
        out[nm] = m[0] if reduce else m
    return r, t, out


def _accumulate(acc, name, S):
    if name in acc:
        ops.axpby(1.0, acc[name].reshape(-1, *S.shape[-2:]), 1.0, S.reshape(-1, *S.shape[-2:]),
                  out=acc[name].reshape(-1, *S.shape[-2:]))
    else:
        acc[name] = S


class FluxFunction(torch.autograd.Function):
    """(R, T) of a Crystal as a differentiable function of pixmap tensors.

        R, T = FluxFunction.apply(crystal, ("L1", "L2"), eps1, eps2)
        (T.sum()).backward()      # eps1.grad, eps2.grad

    Forward sets the layers' permittivities and runs the batched R/T solve;
    backward recomputes the S-matrices and runs ``flux_gradient`` with the
    incoming output gradients as the per-source weights.
    """

    @staticmethod
    def forward(ctx, crystal, names, *eps):
        for nm, e in zip(names, eps):
            crystal.layers[nm].epsilon = e.detach().cpu().numpy()
        crystal.solve(rt_only=True)
        R, T = crystal.poynting_flux_end()
        ctx.crystal, ctx.names = crystal, tuple(names)
        ctx.meta = [(e.dtype, e.device) for e in eps]
        dev = eps[0].device if eps else torch.device("cpu")
        return (torch.as_tensor(np.atleast_1d(R), dtype=torch.float64, device=dev),
                torch.as_tensor(np.atleast_1d(T), dtype=torch.float64, device=dev))

    @staticmethod
    def backward(ctx, gR, gT):
        cl = ctx.crystal
        _, _, grads = flux_gradient(cl, wT=gT, wR=gR, wrt=list(ctx.names), reduce=True)
        out = [grads[nm].to(device=dv, dtype=dt) for nm, (dt, dv) in zip(ctx.names, ctx.meta)]
        return (None, None, *out)

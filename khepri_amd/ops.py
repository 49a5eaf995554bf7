"""Device ops of the RCWA path: thin wrappers that allocate outputs and
workspaces as torch tensors and launch the libkhx kernels on the current
stream.  Tensors are complex128 / float64 on a HIP device.

Numerical status words (eigen-solver convergence, exact zero pivots) are
returned as device int tensors; ``raise_on_info`` turns them into the
exceptions the reference's numpy calls raise.
"""
from __future__ import annotations

import numpy as np
import torch

from . import _lib
from ._lib import check

CDT = torch.complex128
RDT = torch.float64


def _p(t):
    return None if t is None else t.data_ptr()


def _stream():
    return torch.cuda.current_stream().cuda_stream


def _ws(nbytes, device):
    return torch.empty(max(int(nbytes), 16), dtype=torch.uint8, device=device)


def _require(t, dtype, name):
    if t.dtype != dtype or not t.is_contiguous() or t.device.type != "cuda":
        raise TypeError(f"{name}: expected a contiguous {dtype} tensor on a HIP device")


_SLOTS = {}


def eig_slots(n):
    """Matrices the eigensolver runs at once on the current device (cached per
    order and device)."""
    key = (n, torch.cuda.current_device())
    if key not in _SLOTS:
        rc = _lib.load().khx_eig_slots(n)
        if rc < 0:
            check(rc, "khx_eig_slots")
        _SLOTS[key] = rc
    return _SLOTS[key]


_DEFERRED = []   # stack of lists: status words recorded (not checked) while a graph is captured


class deferred_status:
    """Context in which raise_on_info records the status words in `sink`
    instead of reading them (a read is a host sync, which graph capture does
    not allow); the caller checks them later with check_status(sink)."""

    def __init__(self, sink):
        self.sink = sink

    def __enter__(self):
        _DEFERRED.append(self.sink)
        return self.sink

    def __exit__(self, *exc):
        _DEFERRED.pop()
        return False


def check_status(sink):
    for info, what, lu in sink:
        _raise_on_words(info, what, lu)


def raise_on_info(info, what, lu=False):
    """Map per-item status words to numpy-style exceptions (one host sync).

    Eigen-solver words (lu=False): > 0 no convergence, < 0 an LU of the
    S-matrix algebra met an exact zero pivot.  LU words (lu=True): nonzero =
    first zero pivot + 1.  Both raise numpy's LinAlgError, as numpy's eig,
    solve and inv do in the reference.  Inside deferred_status the words are
    recorded for a later check_status instead."""
    if info is None:
        return
    if _DEFERRED:
        _DEFERRED[-1].append((info, what, lu))
        return
    _raise_on_words(info, what, lu)


def _raise_on_words(info, what, lu):
    v = info.cpu().numpy()
    if lu and (v != 0).any():
        raise np.linalg.LinAlgError(f"{what}: singular matrix (items {np.nonzero(v)[0][:8]})")
    if (v > 0).any():
        raise np.linalg.LinAlgError(f"{what}: eigenvalues did not converge (items {np.nonzero(v > 0)[0][:8]})")
    if (v < 0).any():
        raise np.linalg.LinAlgError(f"{what}: singular matrix (items {np.nonzero(v < 0)[0][:8]})")


# --------------------------------------------------------------------------
def kvectors(g, kp, wl):
    """g (2, N) float64; kp (B, 2) complex; wl (B,) float64 -> kx, ky (B, N)."""
    lib = _lib.load()
    B, N = kp.shape[0], g.shape[1]
    kx = torch.empty((B, N), dtype=CDT, device=kp.device)
    ky = torch.empty_like(kx)
    check(lib.khx_kvectors(_p(g), N, _p(kp), _p(wl), B, _p(kx), _p(ky), _stream()), "khx_kvectors")
    return kx, ky


def convmat(eps, pw):
    """eps (count, nx, ny) float64 or complex128 -> C (count, N, N)."""
    lib = _lib.load()
    P, Q = pw
    if eps.dim() == 2:
        eps = eps.unsqueeze(0)
    count, nx, ny = eps.shape
    if nx // 2 + P - 1 >= nx or ny // 2 + Q - 1 >= ny:
        raise IndexError("pixmap too small for the requested number of plane waves")
    cplx = 1 if eps.is_complex() else 0
    eps = eps.contiguous()
    C = torch.empty((count, P * Q, P * Q), dtype=CDT, device=eps.device)
    ws = _ws(lib.khx_convmat_ws_bytes(P, Q, count), eps.device)
    check(lib.khx_convmat(_p(eps), cplx, nx, ny, P, Q, count, _p(C), _p(ws), _stream()), "khx_convmat")
    return C


def _owned(t, overwrite):
    """t itself when the caller hands it over (overwrite) and it is contiguous,
    else a contiguous copy: the kernels work in place."""
    if overwrite:
        return t.contiguous()
    return t.clone(memory_format=torch.contiguous_format)


def solve(A, B, overwrite_a=False, overwrite_b=False):
    """X = A^-1 B for batches (A (b, n, n), B (b, n, m)).  The inputs are left
    untouched unless overwrite_a / overwrite_b hand them over (A then holds
    the LU factors and B is returned as X)."""
    lib = _lib.load()
    Af = _owned(A, overwrite_a)
    X = _owned(B, overwrite_b)
    b, n, m = X.shape
    info = torch.empty(b, dtype=torch.int32, device=A.device)
    check(lib.khx_zgesv(n, m, _p(Af), _p(X), b, _p(info), _stream()), "khx_zgesv")
    return X, info


def inverse(A):
    b, n, _ = A.shape
    I = torch.eye(n, dtype=CDT, device=A.device).expand(b, n, n).contiguous()
    return solve(A, I, overwrite_b=True)


def _mat_view(t, name):
    """(ld, batch stride) in complex elements of a (b, rows, cols) or (rows, cols) view."""
    if t.stride(-1) != 1:
        raise ValueError(f"{name}: last dimension must be contiguous")
    if t.dim() == 2:
        return t.stride(0), 0, 1
    if t.dim() != 3:
        raise ValueError(f"{name}: expected a (batch, rows, cols) view")
    sb = t.stride(0) if t.shape[0] > 1 else 0
    return t.stride(1), sb, t.shape[0]


_OPS = {"N": 0, "T": 1, "C": 2}


def matmul(A, B, alpha=1.0, beta=0.0, C=None, opa="N", opb="N"):
    """alpha op(A) op(B) + beta C for strided batches (views of S-matrix blocks
    welcome); op is "N" (as is), "T" (transpose) or "C" (conjugate transpose).
    A 2-D operand is shared by the batch.  C, if given, is written in place."""
    lib = _lib.load()
    M, K = A.shape[-2:] if opa == "N" else A.shape[-2:][::-1]
    N = B.shape[-1] if opb == "N" else B.shape[-2]
    lda, sa, ba = _mat_view(A, "A")
    ldb, sb, bb = _mat_view(B, "B")
    batch = max(ba, bb)
    if C is None:
        C = torch.empty((batch, M, N), dtype=CDT, device=A.device)
        if complex(beta) != 0:
            raise ValueError("beta != 0 needs C")
    ldc, sc, bc = _mat_view(C, "C")
    if bc != batch or (sc == 0 and batch > 1):
        raise ValueError("C must hold one matrix per batch item")
    al = np.array([complex(alpha).real, complex(alpha).imag])
    be = np.array([complex(beta).real, complex(beta).imag])
    if opa == "N" and opb == "N":
        check(lib.khx_zgemm(M, N, K, al.ctypes.data, _p(A), lda, sa, _p(B), ldb, sb, be.ctypes.data, _p(C), ldc,
                            sc, batch, _stream()), "khx_zgemm")
    else:
        check(lib.khx_zgemm_op(_OPS[opa], _OPS[opb], M, N, K, al.ctypes.data, _p(A), lda, sa, _p(B), ldb, sb,
                               be.ctypes.data, _p(C), ldc, sc, batch, _stream()), "khx_zgemm_op")
    return C


def conj_transpose(A, out=None):
    """A^H of a (b, rows, cols) strided batch -> (b, cols, rows)."""
    lib = _lib.load()
    rows, cols = A.shape[-2:]
    lda, sa, ba = _mat_view(A, "A")
    if out is None:
        out = torch.empty((ba, cols, rows), dtype=CDT, device=A.device)
    ldb, sb, _ = _mat_view(out, "out")
    check(lib.khx_conj_transpose(_p(A), lda, sa, rows, cols, _p(out), ldb, sb, ba, _stream()), "khx_conj_transpose")
    return out


def reduce_dot(A, B, axis, out=None, accumulate=False):
    """sum over the other axis of A * conj(B): (b, rows) for axis 1, (b, cols) for axis 0."""
    lib = _lib.load()
    rows, cols = A.shape[-2:]
    lda, sa, ba = _mat_view(A, "A")
    ldb, sb, bb = _mat_view(B, "B")
    batch = max(ba, bb)
    nout = rows if axis == 1 else cols
    if out is None:
        if accumulate:
            raise ValueError("accumulate needs out")
        out = torch.empty((batch, nout), dtype=CDT, device=A.device)
    if out.stride(-1) != 1 or out.shape[-1] != nout:
        raise ValueError("out: (batch, outputs) with contiguous rows")
    check(lib.khx_reduce_dot(_p(A), lda, sa, _p(B), ldb, sb, rows, cols, axis, int(accumulate), _p(out),
                             out.stride(0), batch, _stream()), "khx_reduce_dot")
    return out


def eig_adjoint_mix(G, mu, mubar, eta_rel):
    lib = _lib.load()
    _require(G, CDT, "G")
    b, n, _ = G.shape
    mu, mubar = mu.contiguous(), mubar.contiguous()   # kept alive until the launch is enqueued
    check(lib.khx_eig_adjoint_mix(_p(G), _p(mu), _p(mubar), n, b, float(eta_rel),
                                  _stream()), "khx_eig_adjoint_mix")
    return G


def lambda_adjoint(lam, k0d, xbar=None, ilbar=None):
    lib = _lib.load()
    b, n = lam.shape
    mubar = torch.empty_like(lam)
    check(lib.khx_lambda_adjoint(_p(lam), _p(k0d), _p(xbar), _p(ilbar), n, b, _p(mubar), _stream()),
          "khx_lambda_adjoint")
    return mubar


def build_pq(C, IC, kx, ky):
    lib = _lib.load()
    B, N = kx.shape
    n = 2 * N
    per_item = int(C.dim() == 3)
    P = torch.empty((B, n, n), dtype=CDT, device=kx.device)
    Q = torch.empty_like(P)
    check(lib.khx_build_pq(_p(C), _p(IC), per_item, _p(kx), _p(ky), N, B, _p(P), _p(Q), _stream()), "khx_build_pq")
    return P, Q


def pq_adjoint(Pbar, Qbar, kx, ky):
    lib = _lib.load()
    B, N = kx.shape
    ICbar = torch.empty((B, N, N), dtype=CDT, device=kx.device)
    Cbar = torch.empty_like(ICbar)
    check(lib.khx_pq_adjoint(_p(Pbar), _p(Qbar), _p(kx), _p(ky), N, B, _p(ICbar), _p(Cbar), _stream()),
          "khx_pq_adjoint")
    return ICbar, Cbar


def convmat_adjoint(Cbar, pw, shape, is_complex=False):
    """Pixel adjoint (b, nx, ny) (float64, or complex128 when is_complex) of Cbar (b, N, N)."""
    lib = _lib.load()
    _require(Cbar, CDT, "Cbar")
    b = Cbar.shape[0]
    nx, ny = shape
    out = torch.empty((b, nx, ny), dtype=CDT if is_complex else RDT, device=Cbar.device)
    ws = _ws(lib.khx_convmat_adjoint_ws_bytes(pw[0], pw[1], b), Cbar.device)
    check(lib.khx_convmat_adjoint(_p(Cbar), pw[0], pw[1], nx, ny, b, int(is_complex), _p(out), _p(ws), _stream()),
          "khx_convmat_adjoint")
    return out


def poynting_vec_adjoint(sR, sT, g, kp, wl, epsi, epse, wR=None, wT=None):
    lib = _lib.load()
    B, n = sT.shape
    ei = np.array([complex(epsi).real, complex(epsi).imag])
    ee = np.array([complex(epse).real, complex(epse).imag])
    sRbar = torch.empty_like(sR)
    sTbar = torch.empty_like(sT)
    check(lib.khx_poynting_vec_adjoint(_p(sR), _p(sT), n // 2, _p(g), _p(kp), _p(wl), ei.ctypes.data, ee.ctypes.data,
                                       _p(wR), _p(wT), B, _p(sRbar), _p(sTbar), _stream()),
          "khx_poynting_vec_adjoint")
    return sRbar, sTbar


def eig(A, overwrite_a=False):
    """Eigen-decomposition of a batch of general complex matrices (A is left
    untouched unless overwrite_a hands it over as workspace)."""
    lib = _lib.load()
    Af = _owned(A, overwrite_a)
    b, n, _ = Af.shape
    lam = torch.empty((b, n), dtype=CDT, device=A.device)
    W = torch.empty((b, n, n), dtype=CDT, device=A.device)
    info = torch.empty(b, dtype=torch.int32, device=A.device)
    ws = _ws(lib.khx_eig_ws_bytes(n, b), A.device)
    check(lib.khx_eig(n, _p(Af), b, _p(lam), _p(W), _p(info), _p(ws), _stream()), "khx_eig")
    return lam, W, info


def patterned_layer(C, IC, kx, ky, k0d, modes=False, ws=None, cache=None):
    """S (B, 2, 2, n, n) of a patterned layer; (W, V, lam) too when modes.
    C, IC: (N, N) shared by the batch or (B, N, N) one per item.  cache: a
    (B, 9, n, n) complex tensor that receives the solves' results for
    patterned_layer_bwd (khx_patterned_layer_cached)."""
    lib = _lib.load()
    B, N = kx.shape
    n = 2 * N
    dev = kx.device
    S = torch.empty((B, 2, 2, n, n), dtype=CDT, device=dev)
    W = V = lam = None
    if modes:
        W = torch.empty((B, n, n), dtype=CDT, device=dev)
        V = torch.empty_like(W)
        lam = torch.empty((B, n), dtype=CDT, device=dev)
    info = torch.empty(B, dtype=torch.int32, device=dev)
    need = lib.khx_patterned_layer_ws_bytes(n, B)
    if ws is None or ws.numel() < need:
        ws = _ws(need, dev)
    per_item = 1 if C.dim() == 3 else 0
    C, IC = C.contiguous(), IC.contiguous()
    if per_item and (C.shape[0] != B or IC.shape[0] != B):
        raise ValueError("per-item convolution matrices must match the batch")
    if cache is not None:
        _require(cache, CDT, "cache")
        if tuple(cache.shape) != (B, 9, n, n) or not cache.is_contiguous():
            raise ValueError(f"cache must be a contiguous ({B}, 9, {n}, {n}) complex tensor")
        check(lib.khx_patterned_layer_cached(_p(C), _p(IC), per_item, _p(kx), _p(ky), N, B, _p(k0d), _p(S), _p(W),
                                             _p(V), _p(lam), _p(info), _p(cache), _p(ws), _stream()),
              "khx_patterned_layer_cached")
        return S, (W, V, lam), info
    check(lib.khx_patterned_layer(_p(C), _p(IC), per_item, _p(kx), _p(ky), N, B, _p(k0d), _p(S), _p(W), _p(V),
                                  _p(lam), _p(info), _p(ws), _stream()), "khx_patterned_layer")
    return S, (W, V, lam), info


_ISLAND_KINDS = {"disc": 0, "rectangle": 1, "uniform": 2}


def convmat_islands(islands, eps_host, pw, reciprocal, sigma, device):
    """(N, N) convolution matrix of an island layer: islands = [{"type":
    "disc" | "rectangle" | "uniform", "params": [...], "epsilon": eps}, ...]
    as in khepri/layer.py:112-119, 161-166."""
    lib = _lib.load()
    P, Q = pw
    nisl = len(islands)
    kinds = np.zeros(max(nisl, 1), dtype=np.int32)
    params = np.zeros((max(nisl, 1), 4), dtype=np.float64)
    eps = np.zeros(max(nisl, 1), dtype=np.complex128)
    for i, isl in enumerate(islands):
        if isl["type"] not in _ISLAND_KINDS:
            raise NotImplementedError(f"island type {isl['type']!r}")
        kinds[i] = _ISLAND_KINDS[isl["type"]]
        pr = list(isl.get("params", []))
        params[i, :len(pr)] = pr
        eps[i] = complex(isl["epsilon"])
    kd = torch.as_tensor(kinds, device=device)
    pd = torch.as_tensor(params, device=device)
    ed = torch.as_tensor(eps, device=device)
    C = torch.empty((P * Q, P * Q), dtype=CDT, device=device)
    ws = _ws(lib.khx_convmat_ws_bytes(P, Q, 1), device)
    eh = np.array([complex(eps_host).real, complex(eps_host).imag])
    rec = np.ascontiguousarray(np.asarray(reciprocal, dtype=np.float64).reshape(4))
    check(lib.khx_convmat_islands(_p(kd), _p(pd), _p(ed), nisl, eh.ctypes.data, rec.ctypes.data, float(sigma), P, Q,
                                  _p(C), _p(ws), _stream()), "khx_convmat_islands")
    return C


def patterned_layers_fused(Cs, ICs, kx, ky, k0ds, modes=False, cache=False):
    """patterned_layer for several layers over the same sources in one call:
    item l*B + b is layer l at source b, so the eigen-solver and every other
    stage runs once over L*B items (a small batch of one layer leaves most CUs
    idle).  Returns one (S, (W, V, lam), info) per layer, as views; with
    cache, one (S, (W, V, lam), info, cache) per layer (cache: the
    patterned_layer_bwd intermediates, (B, 9, n, n))."""
    L = len(Cs)
    B, N = kx.shape
    n = 2 * N
    C = torch.stack([c.expand(B, N, N) for c in Cs]).reshape(L * B, N, N)
    IC = torch.stack([c.expand(B, N, N) for c in ICs]).reshape(L * B, N, N)
    kxl = kx.repeat(L, 1)
    kyl = ky.repeat(L, 1)
    k0dl = torch.cat([k.reshape(B) for k in k0ds])
    cb = torch.empty((L * B, 9, n, n), dtype=CDT, device=kx.device) if cache else None
    S, (W, V, lam), info = patterned_layer(C, IC, kxl, kyl, k0dl, modes=modes, cache=cb)
    out = []
    for l in range(L):
        sl = slice(l * B, (l + 1) * B)
        item = (S[sl], (W[sl] if modes else None, V[sl] if modes else None, lam[sl] if modes else None), info[sl])
        out.append(item + (cb[sl],) if cache else item)
    return out


def analytic_layer(kind, kx, ky, eps, k0d=None, modes=False):
    """kind 'uniform' | 'reflexion' | 'transmission'."""
    lib = _lib.load()
    B, N = kx.shape
    n = 2 * N
    dev = kx.device
    S = torch.empty((B, 2, 2, n, n), dtype=CDT, device=dev)
    W = V = lam = None
    if modes:
        W = torch.empty((B, n, n), dtype=CDT, device=dev)
        V = torch.empty_like(W)
        lam = torch.empty((B, n), dtype=CDT, device=dev)
    e = np.array([complex(eps).real, complex(eps).imag])
    if kind == "uniform":
        check(lib.khx_uniform_layer(_p(kx), _p(ky), N, B, e.ctypes.data, _p(k0d), _p(S), _p(W), _p(V), _p(lam),
                                    _stream()), "khx_uniform_layer")
    else:
        side = 1 if kind == "reflexion" else 2
        check(lib.khx_halfspace(side, _p(kx), _p(ky), N, B, e.ctypes.data, _p(S), _p(W), _p(V), _p(lam),
                                _stream()), "khx_halfspace")
    return S, (W, V, lam)


def redheffer(SA, SB, out=None, ws=None):
    lib = _lib.load()
    ab = SA.dim() == 4
    bb = SB.dim() == 4
    n = SA.shape[-1]
    B = max(1 if ab else SA.shape[0], 1 if bb else SB.shape[0])
    dev = SA.device
    S = out if out is not None else torch.empty((B, 2, 2, n, n), dtype=CDT, device=dev)
    info = torch.empty(B, dtype=torch.int32, device=dev)
    need = lib.khx_redheffer_ws_bytes(n, B)
    if ws is None or ws.numel() < need:
        ws = _ws(need, dev)
    SA, SB = SA.contiguous(), SB.contiguous()   # kept alive until the launch is enqueued
    check(lib.khx_redheffer(_p(SA), int(ab), _p(SB), int(bb), n, B, _p(S), _p(info),
                            _p(ws), _stream()), "khx_redheffer")
    return S, info


def patterned_layer_bwd(C, IC, kx, ky, k0d, W, V, lam, Sbar, eta_rel=1e-12, cache=None):
    """Cbar (B, N, N): adjoint of the convolution matrix of a patterned layer
    given the adjoint Sbar (B, 2, 2, n, n) of its S-matrix and its modes;
    cache: the forward pass's intermediates (patterned_layer(..., cache=)),
    which spare the backward pass four solves."""
    lib = _lib.load()
    B, N = kx.shape
    n = 2 * N
    dev = kx.device
    per_item = 1 if C.dim() == 3 else 0
    C, IC = C.contiguous(), IC.contiguous()
    W, V, lam, Sbar = W.contiguous(), V.contiguous(), lam.contiguous(), Sbar.contiguous()
    Cbar = torch.empty((B, N, N), dtype=CDT, device=dev)
    info = torch.empty(B, dtype=torch.int32, device=dev)
    ws = _ws(lib.khx_patterned_layer_bwd_ws_bytes(n, B), dev)
    if cache is not None:
        cache = cache.contiguous()
        if tuple(cache.shape) != (B, 9, n, n):
            raise ValueError(f"cache must be ({B}, 9, {n}, {n})")
        check(lib.khx_patterned_layer_bwd_cached(_p(C), _p(IC), per_item, _p(kx), _p(ky), N, B, _p(k0d), _p(W),
                                                 _p(V), _p(lam), _p(cache), _p(Sbar), float(eta_rel), _p(Cbar),
                                                 _p(info), _p(ws), _stream()),
              "khx_patterned_layer_bwd_cached")
        return Cbar, info
    check(lib.khx_patterned_layer_bwd(_p(C), _p(IC), per_item, _p(kx), _p(ky), N, B, _p(k0d), _p(W),
                                      _p(V), _p(lam), _p(Sbar),
                                      float(eta_rel), _p(Cbar), _p(info), _p(ws), _stream()),
          "khx_patterned_layer_bwd")
    return Cbar, info


def redheffer_bwd(SA, SB, Sbar):
    """(SAbar, SBbar, info) for S = SA * SB (both (B, 2, 2, n, n))."""
    lib = _lib.load()
    B, n = SA.shape[0], SA.shape[-1]
    dev = SA.device
    SAbar = torch.empty((B, 2, 2, n, n), dtype=CDT, device=dev)
    SBbar = torch.empty_like(SAbar)
    info = torch.empty(B, dtype=torch.int32, device=dev)
    ws = _ws(lib.khx_redheffer_bwd_ws_bytes(n, B), dev)
    SA, SB, Sbar = SA.contiguous(), SB.contiguous(), Sbar.contiguous()
    check(lib.khx_redheffer_bwd(_p(SA), _p(SB), n, B, _p(Sbar), _p(SAbar),
                                _p(SBbar), _p(info), _p(ws), _stream()), "khx_redheffer_bwd")
    return SAbar, SBbar, info


def poynting(Stot, g, kp, wl, te, tm, epsi, epse, orders=False):
    lib = _lib.load()
    B = Stot.shape[0]
    N = g.shape[1]
    dev = Stot.device
    ws = _ws(lib.khx_poynting_ws_bytes(N, B), dev)
    R = torch.empty(B, dtype=RDT, device=dev)
    T = torch.empty(B, dtype=RDT, device=dev)
    Ro = To = None
    if orders:
        Ro = torch.empty((B, N), dtype=RDT, device=dev)
        To = torch.empty((B, N), dtype=RDT, device=dev)
    ei = np.array([complex(epsi).real, complex(epsi).imag])
    ee = np.array([complex(epse).real, complex(epse).imag])
    check(lib.khx_poynting(_p(Stot), N, _p(g), _p(kp), _p(wl), _p(te), _p(tm), ei.ctypes.data, ee.ctypes.data, B,
                           _p(R), _p(T), _p(Ro), _p(To), _p(ws), _stream()), "khx_poynting")
    return R, T, Ro, To


def incident(N, kp, wl, te, tm, epsi):
    """Incident vectors c1p (B, n) on the device."""
    lib = _lib.load()
    B = kp.shape[0]
    c = torch.empty((B, 2 * N), dtype=CDT, device=kp.device)
    ei = np.array([complex(epsi).real, complex(epsi).imag])
    check(lib.khx_incident(N, _p(kp), _p(wl), _p(te), _p(tm), ei.ctypes.data, B, _p(c), _stream()), "khx_incident")
    return c


def poynting_vec(sR, sT, g, kp, wl, epsi, epse, orders=False):
    lib = _lib.load()
    B, n = sR.shape
    N = n // 2
    dev = sR.device
    R = torch.empty(B, dtype=RDT, device=dev)
    T = torch.empty(B, dtype=RDT, device=dev)
    Ro = To = None
    if orders:
        Ro = torch.empty((B, N), dtype=RDT, device=dev)
        To = torch.empty((B, N), dtype=RDT, device=dev)
    ei = np.array([complex(epsi).real, complex(epsi).imag])
    ee = np.array([complex(epse).real, complex(epse).imag])
    sR, sT = sR.contiguous(), sT.contiguous()
    check(lib.khx_poynting_vec(_p(sR), _p(sT), N, _p(g), _p(kp), _p(wl), ei.ctypes.data, ee.ctypes.data, B, _p(R),
                               _p(T), _p(Ro), _p(To), _stream()), "khx_poynting_vec")
    return R, T, Ro, To


def lu_factor(A, overwrite_a=False):
    """(LU, ipiv, info) of a batch (A untouched unless overwrite_a)."""
    lib = _lib.load()
    LU = _owned(A, overwrite_a)
    b, n, _ = LU.shape
    piv = torch.empty((b, n), dtype=torch.int32, device=A.device)
    info = torch.empty(b, dtype=torch.int32, device=A.device)
    check(lib.khx_lu_factor(n, _p(LU), b, _p(piv), _p(info), _stream()), "khx_lu_factor")
    return LU, piv, info


def lu_factor_solve(A, B, overwrite_a=False, overwrite_b=False):
    """(LU, ipiv, info, X = A^-1 B): the factors kept for later solves (inputs
    untouched unless handed over with overwrite_a / overwrite_b)."""
    lib = _lib.load()
    LU = _owned(A, overwrite_a)
    X = _owned(B, overwrite_b)
    b, n, m = X.shape
    piv = torch.empty((b, n), dtype=torch.int32, device=A.device)
    info = torch.empty(b, dtype=torch.int32, device=A.device)
    check(lib.khx_lu_factor_solve(n, m, _p(LU), _p(X), b, _p(piv), _p(info), _stream()), "khx_lu_factor_solve")
    return LU, piv, info, X


def lu_solve(LU, piv, B, overwrite_b=False):
    lib = _lib.load()
    X = _owned(B, overwrite_b)
    if X.dim() == 2:
        X = X.unsqueeze(-1)
    b, n, m = X.shape
    check(lib.khx_lu_solve(n, m, _p(LU), _p(piv), _p(X), b, _stream()), "khx_lu_solve")
    return X


def diag_scale(A, dl=None, dr=None, out=None):
    """diag(dl) A diag(dr) for a strided batch (A (b, rows, cols), dl (b, rows), dr (b, cols))."""
    lib = _lib.load()
    lda, sa, b = _mat_view(A, "A")
    rows, cols = A.shape[-2:]
    out = torch.empty((b, rows, cols), dtype=CDT, device=A.device) if out is None else out
    ldo, so, _ = _mat_view(out, "out")
    dl = None if dl is None else dl.contiguous()
    dr = None if dr is None else dr.contiguous()
    check(lib.khx_diag_scale(_p(A), lda, sa, _p(dl), _p(dr), rows, cols, b, _p(out), ldo, so, _stream()),
          "khx_diag_scale")
    return out


def axpby(a, A, b=0.0, B=None, out=None):
    """a A + b B for strided batches of matrices (out may alias A or B)."""
    lib = _lib.load()
    rows, cols = A.shape[-2:]
    lda, sa, ba = _mat_view(A, "A")
    if B is not None:
        ldb, sb, bb = _mat_view(B, "B")
    else:
        ldb, sb, bb = cols, 0, ba
    batch = max(ba, bb)
    if out is None:
        out = torch.empty((batch, rows, cols), dtype=CDT, device=A.device)
    ldd, sd, _ = _mat_view(out, "out")
    al = np.array([complex(a).real, complex(a).imag])
    be = np.array([complex(b).real, complex(b).imag])
    check(lib.khx_zaxpby(_p(out), ldd, sd, al.ctypes.data, _p(A), lda, sa, be.ctypes.data, _p(B), ldb, sb, rows, cols,
                         batch, _stream()), "khx_zaxpby")
    return out


def identity_minus(A, out=None):
    lib = _lib.load()
    A = A.contiguous()
    b, n, _ = A.shape
    out = torch.empty_like(A) if out is None else out
    check(lib.khx_identity_minus(_p(A), n, b, _p(out), _stream()), "khx_identity_minus")
    return out


def propagation(kx, ky, k0d):
    lib = _lib.load()
    B, N = kx.shape
    X = torch.empty((B, 2 * N), dtype=CDT, device=kx.device)
    check(lib.khx_propagation(_p(kx), _p(ky), N, B, _p(k0d), _p(X), _stream()), "khx_propagation")
    return X


def joint_subspace(x, Nb, kind, blocks):
    """x (items*Nb, blocks..., 2Nb, 2Nb) -> (items, blocks..., 2Nb^2, 2Nb^2)."""
    lib = _lib.load()
    x = x.contiguous()
    items = x.shape[0] // Nb
    ne = 2 * Nb * Nb
    shape = (items, 2, 2, ne, ne) if blocks == 4 else (items, ne, ne)
    out = torch.empty(shape, dtype=CDT, device=x.device)
    check(lib.khx_joint_subspace(_p(x), Nb, blocks, kind, items, _p(out), _stream()), "khx_joint_subspace")
    return out


def joint_subspace_diag(v, Nb, kind):
    """v (items*Nb, 2Nb) -> (items, 2Nb^2)."""
    lib = _lib.load()
    v = v.contiguous()
    items = v.shape[0] // Nb
    out = torch.empty((items, 2 * Nb * Nb), dtype=CDT, device=v.device)
    check(lib.khx_joint_subspace_diag(_p(v), Nb, kind, items, _p(out), _stream()), "khx_joint_subspace_diag")
    return out


def eigenbasis(W, V):
    """[[W, W], [-V, V]] for (b, n, n) batches."""
    lib = _lib.load()
    W, V = W.contiguous(), V.contiguous()
    b, n, _ = W.shape
    R = torch.empty((b, 2 * n, 2 * n), dtype=CDT, device=W.device)
    check(lib.khx_eigenbasis(_p(W), _p(V), n, b, _p(R), _stream()), "khx_eigenbasis")
    return R


def mode_propagate(lam, a, zbar):
    lib = _lib.load()
    n = lam.shape[-1]
    lam, a, zbar = lam.contiguous(), a.contiguous(), zbar.contiguous()
    M = torch.empty((2 * n, zbar.shape[0]), dtype=CDT, device=lam.device)
    check(lib.khx_mode_propagate(_p(lam), _p(a), _p(zbar), n, zbar.shape[0], _p(M), _stream()), "khx_mode_propagate")
    return M


def longitudinal(kx, ky, F):
    lib = _lib.load()
    N = kx.shape[-1]
    nz = F.shape[-1]
    F, kx, ky = F.contiguous(), kx.contiguous(), ky.contiguous()
    uz = torch.empty((N, nz), dtype=CDT, device=F.device)
    w = torch.empty_like(uz)
    check(lib.khx_longitudinal(_p(kx), _p(ky), _p(F), N, nz, _p(uz), _p(w), _stream()),
          "khx_longitudinal")
    return uz, w


def mode_propagate_batch(lam, a, zbar):
    """lam (B, n), a (B, 2n), zbar (nz) -> M (B, 2n, nz)."""
    lib = _lib.load()
    B, n = lam.shape
    lam, a, zbar = lam.contiguous(), a.contiguous(), zbar.contiguous()
    M = torch.empty((B, 2 * n, zbar.shape[0]), dtype=CDT, device=lam.device)
    check(lib.khx_mode_propagate_batch(_p(lam), _p(a), _p(zbar), n, zbar.shape[0], B, _p(M), _stream()),
          "khx_mode_propagate_batch")
    return M


def longitudinal_batch(kx, ky, F):
    """kx, ky (B, N), F (B, 4N, nz) -> uz, w (B, N, nz)."""
    lib = _lib.load()
    B, N = kx.shape
    nz = F.shape[-1]
    F, kx, ky = F.contiguous(), kx.contiguous(), ky.contiguous()
    uz = torch.empty((B, N, nz), dtype=CDT, device=F.device)
    w = torch.empty_like(uz)
    check(lib.khx_longitudinal_batch(_p(kx), _p(ky), _p(F), N, nz, B, _p(uz), _p(w), _stream()),
          "khx_longitudinal_batch")
    return uz, w


def idft_phase_batch(kx, ky, wl, x, y):
    """exp(i k0_b (Kx x + Ky y)) for normalised K (B, N), k0_b = 2 pi / wl[b]:
    (B, npts, N)."""
    lib = _lib.load()
    B, N = kx.shape
    x, y, kx, ky, wl = x.contiguous(), y.contiguous(), kx.contiguous(), ky.contiguous(), wl.contiguous()
    E = torch.empty((B, x.shape[0], N), dtype=CDT, device=x.device)
    check(lib.khx_idft_phase_batch(_p(kx), _p(ky), _p(wl), _p(x), _p(y), N, x.shape[0], B, _p(E), _stream()),
          "khx_idft_phase_batch")
    return E


def idft_phase(kx, ky, k0, x, y):
    """exp(i k0 (Kx x + Ky y)) for normalised K, (npts, N)."""
    lib = _lib.load()
    N = kx.shape[-1]
    x, y, kx, ky = x.contiguous(), y.contiguous(), kx.contiguous(), ky.contiguous()
    E = torch.empty((x.shape[0], N), dtype=CDT, device=x.device)
    check(lib.khx_idft_phase(_p(kx), _p(ky), float(k0), _p(x), _p(y), N, x.shape[0], _p(E),
                             _stream()), "khx_idft_phase")
    return E

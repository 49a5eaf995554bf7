"""R/T-only evaluation of the star-product chain.

``poynting_flux_end`` (khepri/crystal.py:363-396) reads only Stot11 c1p and
Stot21 c1p, where c1p is the incident vector.  With
Stot = S_1 * R_2, R_j = S_j * R_{j+1} (Redheffer products, khepri/alternative.py:19-30):

    D_j    = I - R11_{j+1} S22_j
    R11_j  = S11_j + S12_j D_j^-1 R11_{j+1} S21_j          (needed for j >= 2)
    Stot11 c = S11_1 c + S12_1 D_1^-1 R11_2 S21_1 c
    Stot21 c = R21_2 F_1^-1 S21_1 c,  F^-1 y = y + S22 D^-1 R11 y   (push-through)

so each junction costs one GEMM and one LU (plus one solve and two GEMMs to
build R11_j when more junctions follow) and the rest is matrix-vector work --
the same R and T as the full chain, without forming the other three blocks.

Layers whose medium is the gap medium (uniform eps = 1: S = [[0, X], [X, 0]])
fold into a neighbour as diagonal scalings, and eps = 1 half spaces are the
identity S-matrix; both are exact.
"""
from __future__ import annotations

import torch

from . import ops


def _fold_left(X, S):
    """P * S with P = [[0, X], [X, 0]]: [[X S11 X, X S12], [S21 X, S22]]."""
    out = torch.empty_like(S)
    ops.diag_scale(S[:, 0, 0], X, X, out=out[:, 0, 0])
    ops.diag_scale(S[:, 0, 1], X, None, out=out[:, 0, 1])
    ops.diag_scale(S[:, 1, 0], None, X, out=out[:, 1, 0])
    out[:, 1, 1].copy_(S[:, 1, 1])
    return out


def _fold_right(S, X):
    """S * P: [[S11, S12 X], [X S21, X S22 X]]."""
    out = torch.empty_like(S)
    out[:, 0, 0].copy_(S[:, 0, 0])
    ops.diag_scale(S[:, 0, 1], None, X, out=out[:, 0, 1])
    ops.diag_scale(S[:, 1, 0], X, None, out=out[:, 1, 0])
    ops.diag_scale(S[:, 1, 1], X, X, out=out[:, 1, 1])
    return out


def dense_sequence(items):
    """items: [("dense", S) | ("prop", X) | ("identity", None)] in stack order.
    Returns the dense S-matrices with propagation layers folded in, or
    ("prop", X_total) when the stack has no dense layer."""
    items = [it for it in items if it[0] != "identity"]
    out = []
    pending = None
    for kind, val in items:
        if kind == "prop":
            pending = val if pending is None else ops.diag_scale(pending[:, None, :], None, val)[:, 0, :]
        else:
            out.append(_fold_left(pending, val) if pending is not None else val)
            pending = None
    if pending is not None:
        if out:
            out[-1] = _fold_right(out[-1], pending)
        else:
            return ("prop", pending)
    return out


def _mv(M, v, add=None):
    """M v (+ add) for batches of matrices and vectors."""
    C = None if add is None else add.unsqueeze(-1).clone()
    return ops.matmul(M, v.unsqueeze(-1), beta=0.0 if add is None else 1.0, C=C)[..., 0]


def rt_amplitudes(items, c):
    """(Stot11 c, Stot21 c, LU status words) for incident vectors c (B, n)."""
    seq = dense_sequence(items)
    if isinstance(seq, tuple):   # no scattering layer: pure propagation
        return torch.zeros_like(c), ops.diag_scale(c[:, None, :], None, seq[1])[:, 0, :], []
    m = len(seq)
    if m == 1:
        S = seq[0]
        return _mv(S[:, 0, 0], c), _mv(S[:, 1, 0], c), []
    # right-to-left: R11 of every right sub-stack and the LU of each junction
    R11 = [None] * m
    LU = [None] * m
    R11[m - 1] = seq[m - 1][:, 0, 0]
    infos = []
    for j in range(m - 2, -1, -1):
        Sj = seq[j]
        P = ops.matmul(R11[j + 1], Sj[:, 1, 1])
        D = ops.identity_minus(P, out=P)            # temporaries: factored in place
        if j >= 1:  # factor D_j and solve for R11_{j+1} S21_j in the same pass
            lu, piv, info, Y = ops.lu_factor_solve(D, ops.matmul(R11[j + 1], Sj[:, 1, 0]),
                                                   overwrite_a=True, overwrite_b=True)
            R11[j] = ops.matmul(Sj[:, 0, 1], Y, beta=1.0, C=Sj[:, 0, 0].clone())
        else:
            lu, piv, info = ops.lu_factor(D, overwrite_a=True)
        infos.append(info)
        LU[j] = (lu, piv)
    # left-to-right vector pass
    S1 = seq[0]
    v = _mv(S1[:, 1, 0], c)
    u = ops.lu_solve(LU[0][0], LU[0][1], _mv(R11[1], v), overwrite_b=True)[..., 0]
    sR = _mv(S1[:, 0, 1], u, add=_mv(S1[:, 0, 0], c))
    y = _mv(S1[:, 1, 1], u, add=v)
    for j in range(1, m - 1):
        Sj = seq[j]
        v = _mv(Sj[:, 1, 0], y)
        u = ops.lu_solve(LU[j][0], LU[j][1], _mv(R11[j + 1], v), overwrite_b=True)[..., 0]
        y = _mv(Sj[:, 1, 1], u, add=v)
    sT = _mv(seq[m - 1][:, 1, 0], y)
    return sR, sT, infos

"""numpy / scipy restatement of the reference's band-structure tool
(khepri/eigentricks.py) -- TEST ORACLE ONLY.

The S-matrix here is one (4 ng x 4 ng) array in the (+, -) block layout of
the reference's band-structure code: rows / columns 0:2ng are the forward
(p) amplitudes, 2ng:4ng the backward (m) ones, with S[p,p] the forward
transmission, S[m,p] the reflection from the left, S[p,m] the reflection
from the right and S[m,m] the backward transmission.  The Bloch modes of a
period with that S are the generalized eigenpairs of (Sl, Sr).

Pinned by closed forms (tests/test_oracle.py): a homogeneous slab has the
Bloch factors exp(+- i kz d); a two-layer Bragg period at normal incidence
satisfies w + 1/w = 2 cos(k1 d1) cos(k2 d2) - (n1/n2 + n2/n1) sin(k1 d1)
sin(k2 d2); the Schur-complement determinant equals det(S); the pencil's
eigenpairs are those of Sr^-1 Sl.
"""
from __future__ import annotations

import numpy as np
from scipy.linalg import eig


def scattering_splitlr(S):
    """(Sl, Sr) = ([[Spp, 0], [Smp, -I]], [[I, -Spm], [0, -Smm]]).  khepri/eigentricks.py:5-22."""
    ng = S.shape[0] // 4
    p, m = slice(0, 2 * ng), slice(2 * ng, 4 * ng)
    eye = np.eye(2 * ng, dtype=S.dtype)
    Sl = np.zeros_like(S)
    Sl[p, p] = S[p, p]
    Sl[m, p] = S[m, p]
    Sl[m, m] = -eye
    Sr = np.zeros_like(S)
    Sr[p, p] = eye
    Sr[p, m] = -S[p, m]
    Sr[m, m] = -S[m, m]
    return Sl, Sr


def scattering_det(S):
    """det(Spp - Spm Smm^-1 Smp) det(Smm).  khepri/eigentricks.py:24-28."""
    ng = S.shape[0] // 4
    p, m = slice(0, 2 * ng), slice(2 * ng, 4 * ng)
    return np.linalg.det(S[p, p] - S[p, m] @ np.linalg.inv(S[m, m]) @ S[m, p]) * np.linalg.det(S[m, m])


def scattering_eigenvalues(S, dos=False):
    """Generalized eigenpairs of (Sl, Sr) (unit-norm eigenvectors), None when
    Sl or Sr holds a NaN; with dos also the determinant.  khepri/eigentricks.py:29-40."""
    Sl, Sr = scattering_splitlr(S)
    if np.any(np.isnan(Sr)) or np.any(np.isnan(Sl)):
        return None
    w, v = eig(Sl, Sr)
    return (w, v, scattering_det(S)) if dos else (w, v)


def on_shell(eigenvalues, tol=1e-10):
    """|w| == 1 within tol.  khepri/eigentricks.py:42-43."""
    return np.isclose(np.abs(eigenvalues), 1.0, rtol=0.0, atol=tol)


def band_structure(S):
    """Real parts in (0, 1) of the eigenvalues, kept where on shell.  khepri/eigentricks.py:47-55."""
    r = scattering_eigenvalues(S)
    if r is None:
        return None
    w = r[0].real
    w = w[w > 0]
    w = w[w < 1]
    return w[on_shell(w)]


def from_blocks(S2x2):
    """(4 ng x 4 ng) band-structure S from a Redheffer S-matrix [[S11, S12],
    [S21, S22]] (S11 reflection and S21 transmission of light from the left):
    S[p,p] = S21, S[p,m] = S22, S[m,p] = S11, S[m,m] = S12."""
    S2x2 = np.asarray(S2x2)
    return np.block([[S2x2[1, 0], S2x2[1, 1]], [S2x2[0, 0], S2x2[0, 1]]])

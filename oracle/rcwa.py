"""numpy complex128 restatement of the reference RCWA math (test oracle only).

Each function names the reference function and file:line it restates.  The
operation order follows the reference (dense solves, dense GEMMs, one eig
per patterned layer) so the same code doubles as the CPU baseline.
"""
from __future__ import annotations

import cmath
from math import floor, prod

import numpy as np

TWOPI = 2.0 * np.pi


# --------------------------------------------------------------------------
# basis & geometry
# --------------------------------------------------------------------------
def expansion_indices(pw):
    """Harmonic index pairs, m (x) fastest.  khepri/expansion.py:4-16."""
    P, Q = pw
    if P % 2 != 1 or Q % 2 != 1:
        raise AssertionError("plane-wave counts must be odd")
    m = np.arange(-(P // 2), P // 2 + 1)
    n = np.arange(-(Q // 2), Q // 2 + 1)
    mm = np.tile(m, Q)            # m varies fastest
    nn = np.repeat(n, P)
    return np.vstack([mm, nn])


def reciproc(a1, a2):
    """Reciprocal basis.  khepri/tools.py:64-69."""
    det = a1[0] * a2[1] - a1[1] * a2[0]
    c = TWOPI / det
    return (a2[1] * c, -a2[0] * c), (-a1[1] * c, a1[0] * c)


def lattice_matrix(lattice="square", pitch=1.0):
    """Rows are lattice vectors.  khepri/crystal.py:60-74."""
    if isinstance(lattice, str):
        if lattice == "square":
            return pitch * np.asarray([[1, 0], [0, 1]])
        if lattice == "hexagonal":
            return pitch * np.asarray([[np.sqrt(3) / 2, 0.5], [np.sqrt(3) / 2, -0.5]])
        raise NotImplementedError(lattice)
    return np.asarray(lattice)


def g_vectors(pw, lattice):
    """G = m*b1 + n*b2, shape (2, N).  khepri/expansion.py:30-41."""
    lattice = np.asarray(lattice)
    b = np.asarray(reciproc(lattice[0], lattice[1]))   # (2 basis, 2 xy)
    idx = expansion_indices(pw)                         # (2, N)
    return b[0][:, None] * idx[0][None, :] + b[1][:, None] * idx[1][None, :]


def rotation_matrix(theta):
    """khepri/tools.py:23-24."""
    c, s = np.cos(theta), np.sin(theta)
    return np.array([[c, -s], [s, c]])


def minkowski_sum(g_lhs, g_rhs):
    """Extended basis: index i_lhs*N_rhs + i_rhs holds g_rhs + g_lhs.
    khepri/expansion.py:55-73."""
    out = (g_rhs[:, None, :] + g_lhs[:, :, None]).reshape(2, -1)
    return out


def kz_from_kplanar(kx, ky, k0, eps):
    """kz = conj(sqrt(k0^2 conj(eps) - kx^2 - ky^2)).  khepri/expansion.py:19-26."""
    arg = (k0 ** 2 * np.conj(eps) - kx ** 2 - ky ** 2).astype(complex)
    return np.conj(np.sqrt(arg))


def k_vectors(g, kp, wavelength, eps=1.0):
    """(Kx, Ky, Kz)/k0, complex (3, N).  khepri/expansion.py:43-50, 95-97."""
    k0 = TWOPI / wavelength
    kv = np.zeros((3, g.shape[1]), dtype=np.complex128)
    kv[0:2] = np.asarray(kp)[:, None] + g
    kv[2] = kz_from_kplanar(kv[0], kv[1], k0, eps)
    return kv / k0


def compute_kplanar(eps_inc, wavelength, theta_deg=0.0, phi_deg=0.0):
    """k_par from incidence angles in degrees.  khepri/tools.py:17-21."""
    phi = np.deg2rad(phi_deg)
    theta = np.deg2rad(theta_deg)
    d = np.array([np.cos(phi), np.sin(phi)], dtype=complex)
    return d * cmath.sqrt(eps_inc) * TWOPI / wavelength * np.sin(theta)


def gen_bzi_grid(shape, a=1, reciproc_basis=None):
    """Cell-centred Bloch grid over the first BZ.  khepri/beams.py:193-207."""
    if reciproc_basis is None:
        b1, b2 = [TWOPI / a, 0], [0, TWOPI / a]
    else:
        b1, b2 = reciproc_basis
    si, sj = 1 / shape[0], 1 / shape[1]
    i, j = np.meshgrid(np.arange(-0.5 + si / 2, 0.5, si),
                       np.arange(-0.5 + sj / 2, 0.5, sj), indexing="ij")
    return np.stack([b1[0] * i + b2[0] * j, b1[1] * i + b2[1] * j])


# --------------------------------------------------------------------------
# convolution matrix
# --------------------------------------------------------------------------
def fourier_coefficients(eps2d):
    """fftshift(fft2(eps))/(Nx*Ny).  khepri/tools.py:34."""
    return np.fft.fftshift(np.fft.fft2(eps2d)) / prod(eps2d.shape)


def toeplitz_from_fourier(coefs, pw):
    """C[q*P+p, q'*P+p'] = coefs[g0x + p - p', g0y + q - q'] (python negative
    index wrap kept).  khepri/tools.py:38-56."""
    P, Q = pw
    Nx, Ny = coefs.shape
    g0x, g0y = floor(Nx / 2), floor(Ny / 2)
    N = P * Q
    p = np.arange(N) % P
    q = np.arange(N) // P
    ix = g0x + p[:, None] - p[None, :]
    iy = g0y + q[:, None] - q[None, :]
    if ix.max() >= Nx or iy.max() >= Ny:
        raise IndexError("pixmap too small for the requested harmonics")
    ix = np.where(ix < 0, ix + Nx, ix)
    iy = np.where(iy < 0, iy + Ny, iy)
    return coefs[ix, iy].astype(np.complex128)


def convolution_matrix(eps2d, pw):
    """khepri/tools.py:33-35."""
    return toeplitz_from_fourier(fourier_coefficients(eps2d), pw)


# --------------------------------------------------------------------------
# analytical islands (closed-form Fourier coefficients)
# --------------------------------------------------------------------------
def unit_cell_area(lattice):
    """|a1 x a2|.  khepri/tools.py:58-60."""
    a1, a2 = np.asarray(lattice)
    return abs(a1[0] * a2[1] - a1[1] * a2[0])


def fexpz(z):
    """(e^z - 1)/z, with the reference's truncated series for |z| <= 1e-2.
    khepri/fourier.py:7-20."""
    z = np.asarray(z, dtype=complex)
    out = np.empty_like(z)
    big = np.abs(z) > 1e-2
    out[big] = (np.exp(z[big]) - 1.0) / z[big]
    w = z[~big]
    out[~big] = 1 + w / 2.0 * (1 + w / 3.0 * (1 + w / 4.0 * (1 + w / 5.0 * (1 + w / 6.0 * (1 + w / 7)))))
    return out


def transform_rectangle(ll0, ll1, ur0, ur1, Gx, Gy, sigma):
    """Fourier coefficient of the indicator of [ll0, ur0] x [ll1, ur1].
    khepri/fourier.py:28-39."""
    a, b = ur0 - ll0, ur1 - ll1
    return a * b / sigma * fexpz(-1j * Gx * a) * fexpz(-1j * Gy * b) * np.exp(-1j * (ll0 * Gx + ll1 * Gy))


def transform_disc(c0, c1, radius, Gx, Gy, sigma):
    """Fourier coefficient of the indicator of a disc: pi r^2 / sigma *
    2 J1(|G| r) / (|G| r) * e^{-i G.c}, pi r^2 / sigma at G = 0.
    khepri/fourier.py:42-58 (scipy.special.jv)."""
    from scipy import special
    Gx, Gy = np.asarray(Gx, dtype=float), np.asarray(Gy, dtype=float)
    norm = np.sqrt(Gx * Gx + Gy * Gy) * radius
    zero = np.isclose(Gx, 0.0) & np.isclose(Gy, 0.0)
    out = np.zeros_like(Gx, dtype=complex)
    nz = ~zero
    out[nz] = (np.pi * radius ** 2 / sigma * 2 * np.exp(-1j * (c0 * Gx[nz] + c1 * Gy[nz]))
               * special.jv(1.0, norm[nz]) / norm[nz])
    out[zero] = np.pi * radius ** 2 / sigma
    return out


def island_transform(island, Gx, Gy, sigma):
    """khepri/fourier.py:23-25 (dispatch on the island type)."""
    kind, p = island["type"], island["params"]
    if kind == "disc":
        return transform_disc(*p, Gx, Gy, sigma)
    if kind == "rectangle":
        return transform_rectangle(*p, Gx, Gy, sigma)
    if kind == "uniform":
        return np.zeros_like(np.asarray(Gx), dtype=complex)
    raise NotImplementedError(kind)


def combine_fourier_masks(islands_data, eps_host, inverse=False):
    """eps_g = eps_host delta_0 + sum_i (eps_i - eps_host) T_i.
    khepri/fourier.py:145-159."""
    length = islands_data[0][0].shape[0]
    center = (length - 1) // 2
    trsf = (lambda x: x, lambda x: 1 / x)[inverse]
    eps_g = np.zeros((length,), dtype=complex)
    eps_g[center] = trsf(eps_host)
    for bf, island_eps in islands_data:
        eps_g += (trsf(island_eps) - trsf(eps_host)) * bf
    return eps_g


def expanded_pw(pw, mul=3):
    """khepri/expansion.py:86-92 (the plane-wave counts of the expanded grid)."""
    return [e * mul if e > 1 else 1 for e in pw]


def analytic_convolution_matrix(islands, eps_host, pw, lattice):
    """C of an island layer: the island coefficients on the 3x-expanded
    harmonic grid (m fastest), the flat array reshaped to the expanded
    (P', Q') shape and gathered like a pixmap's spectrum.  The reshape pairs
    the flat index i*Q' + j with (row i, column j), which for P = Q swaps the
    roles of the two harmonic indices; this is reproduced as it is.
    khepri/layer.py:161-166, khepri/expansion.py:86-92, khepri/tools.py:38-56."""
    epw = expanded_pw(pw)
    Gx, Gy = g_vectors(epw, lattice)
    sigma = unit_cell_area(lattice)
    data = [(island_transform(isl, Gx, Gy, sigma), isl["epsilon"]) for isl in islands]
    if not data:
        data = [(np.zeros(Gx.shape, dtype=complex), eps_host)]
    fourier = combine_fourier_masks(data, eps_host).T
    return toeplitz_from_fourier(fourier.reshape(epw), pw)


# --------------------------------------------------------------------------
# modes of the layers
# --------------------------------------------------------------------------
def _kz_masked(arg):
    """Branch rule of free-space / uniform layers: Re(kz^2) < 0 -> -i sqrt(-kz^2).
    khepri/alternative.py:92-96, 145-149."""
    kz = np.asarray(arg, dtype=complex).copy()
    neg = kz.real < 0
    kz[neg] = -1j * np.sqrt(-kz[neg])
    kz[~neg] = np.sqrt(kz[~neg])
    return kz


def free_space_eigenmodes(kx, ky):
    """W0 = I, V0 = Q0 / (i kz) column-scaled.  khepri/alternative.py:84-99."""
    N = len(kx)
    KX, KY = np.diag(kx), np.diag(ky)
    I = np.eye(N)
    Q = np.block([[KX * KY, I - KX ** 2], [KY ** 2 - I, -KY * KX]])
    kz = _kz_masked(np.diag(I - KX ** 2 - KY ** 2))
    lam = np.hstack((1j * kz, 1j * kz))
    return np.eye(2 * N), Q / lam


def uniform_layer_modes(kx, ky, eps):
    """khepri/alternative.py:130-156."""
    N = len(kx)
    KX, KY = np.diag(kx), np.diag(ky)
    I = np.eye(N, dtype=np.complex128)
    P = (1 / eps) * np.block([[KX * KY, eps * I - KX ** 2], [KY ** 2 - eps * I, -KY * KX]])
    Q = eps * P
    kz = _kz_masked(np.diag(eps * I - KX ** 2 - KY ** 2))
    lam = np.hstack((1j * kz, 1j * kz))
    return np.eye(2 * N), Q / lam, lam


def structured_pq(kx, ky, C):
    """P and Q of a patterned layer.  khepri/alternative.py:158-170."""
    KX, KY = np.diag(kx), np.diag(ky)
    I = np.eye(len(kx))
    CiKy = np.linalg.solve(C, KY)
    CiKx = np.linalg.solve(C, KX)
    P = np.block([[KX @ CiKy, I - KX @ CiKx], [KY @ CiKy - I, -KY @ CiKx]]).astype(np.complex128)
    Q = np.block([[KX @ KY, C - KX @ KX], [KY @ KY - C, -KY @ KX]]).astype(np.complex128)
    return P, Q


def structured_layer_modes(kx, ky, C):
    """eig(P Q), lam = principal sqrt, V = Q W / lam.  khepri/alternative.py:158-178."""
    P, Q = structured_pq(kx, ky, C)
    lam2, W = np.linalg.eig(P @ Q)
    lam = np.sqrt(lam2 + 0j)
    V = Q @ W / lam
    return W, V, lam


def build_scatmat(W, V, W0, V0, lam, depth, k0):
    """Symmetric layer S-matrix.  khepri/alternative.py:181-195."""
    t1 = np.linalg.solve(W, W0)
    t2 = np.linalg.solve(V, V0)
    A, B = t1 + t2, t1 - t2
    X = np.diag(np.exp(-lam * depth * k0))
    T = A - X @ B @ np.linalg.solve(A, X) @ B
    S11 = np.linalg.solve(T, X @ B @ np.linalg.solve(A, X) @ A - B)
    S12 = np.linalg.solve(T, X @ (A - B @ np.linalg.solve(A, B)))
    return np.array([[S11, S12], [S12, S11]])


def _halfspace_AB(kx, ky, W0, V0, eps):
    N = len(kx)
    KX, KY = np.diag(kx), np.diag(ky)
    I = np.eye(N)
    Qh = np.block([[KX @ KY, eps * I - KX @ KX], [KY @ KY - eps * I, -KY @ KX]])
    Kz = np.conj(np.sqrt(np.diag(eps * I - KX ** 2 - KY ** 2).astype(complex)))
    lam = np.hstack((1j * Kz, 1j * Kz))
    Wh = np.eye(2 * N)
    Vh = Qh / lam
    A = np.linalg.solve(W0, Wh) + np.linalg.solve(V0, Vh)
    B = np.linalg.solve(W0, Wh) - np.linalg.solve(V0, Vh)
    return A, B, Wh, Vh, lam


def halfspace_reflection(kx, ky, W0, V0, eps):
    """Incidence half-space.  khepri/alternative.py:32-56."""
    A, B, Wh, Vh, lam = _halfspace_AB(kx, ky, W0, V0, eps)
    S11 = -np.linalg.solve(A, B)
    S12 = 2 * np.linalg.inv(A)
    S21 = 0.5 * (A - B @ np.linalg.solve(A, B))
    S22 = B @ np.linalg.inv(A)
    return np.array([[S11, S12], [S21, S22]]), Wh, Vh, lam


def halfspace_transmission(kx, ky, W0, V0, eps):
    """Emergence half-space.  khepri/alternative.py:59-82."""
    A, B, Wh, Vh, lam = _halfspace_AB(kx, ky, W0, V0, eps)
    S11 = B @ np.linalg.inv(A)
    S12 = 0.5 * (A - B @ np.linalg.solve(A, B))
    S21 = 2 * np.linalg.inv(A)
    S22 = -np.linalg.solve(A, B)
    return np.array([[S11, S12], [S21, S22]]), Wh, Vh, lam


def scattering_identity(n):
    """khepri/alternative.py:220-232 (block form)."""
    I = np.eye(n, dtype=complex)
    Z = np.zeros_like(I)
    return np.array([[Z, I], [I, Z]])


def redheffer(SA, SB):
    """Star product SA * SB.  khepri/alternative.py:19-30."""
    I = np.eye(SA[0, 0].shape[0], dtype=np.complex128)
    D = I - SB[0, 0] @ SA[1, 1]
    F = I - SA[1, 1] @ SB[0, 0]
    S11 = SA[0, 0] + SA[0, 1] @ np.linalg.solve(D, SB[0, 0]) @ SA[1, 0]
    S12 = SA[0, 1] @ np.linalg.solve(D, SB[0, 1])
    S21 = SB[1, 0] @ np.linalg.solve(F, SA[1, 0])
    S22 = SB[1, 1] + SB[1, 0] @ np.linalg.solve(F, SA[1, 1]) @ SB[0, 1]
    return np.array([[S11, S12], [S21, S22]])


def stack_layers(Ss, mask, forward_only=False):
    """Forward chain (with retained copies) and the reverse chain.
    khepri/layer.py:35-60.  Returns (Sls, Srs, Stot)."""
    n = Ss[0].shape[-1]
    Stot = scattering_identity(n)
    Sls = []
    for S, keep in zip(Ss, mask):
        Stot = redheffer(Stot, S)
        Sls.append(Stot.copy() if keep else None)
    if forward_only:
        return Sls, None, Stot
    Srev = scattering_identity(n)
    rmask = list(reversed(mask[1:]))
    Srs = []
    for keep, S in zip(rmask, reversed(Ss[1:])):
        Srs.append(Srev.copy() if keep else None)
        Srev = redheffer(S, Srev)
    Srs.append(Srev.copy())
    return Sls, list(reversed(Srs)), Stot


# --------------------------------------------------------------------------
# source & fluxes
# --------------------------------------------------------------------------
def incident(pw, te, tm, k_vector, normalize=True):
    """Incident plane-wave vector.  khepri/alternative.py:101-128."""
    pn = np.linalg.norm((abs(te), abs(tm)))
    if normalize:
        te = te / pn
        tm = tm / pn
    kvec = np.array(k_vector)
    kpn = np.linalg.norm(kvec[0:2])
    kbar = kvec / np.linalg.norm(kvec)
    if abs(kpn) < 1e-8:
        aTE = np.array([1, 0, 0])
        aTM = np.array([0, 1, 0])
    else:
        aTE = -np.cross(np.array([0, 0, -1], dtype=np.complex128), kbar)
        aTE = aTE / np.linalg.norm(aTE)
        aTM = np.cross(aTE, kbar)
        aTM = aTM / np.linalg.norm(aTM)
    N = prod(pw)
    delta = np.zeros(N, dtype=np.complex128)
    delta[(N - 1) // 2] = 1
    pxy = te * aTE + tm * aTM
    return np.hstack([delta * pxy[0], delta * pxy[1]])


def poynting_fluxes(g, c_output, kp, wavelength, only_total=True, epsi=1, epse=1):
    """Diffraction-order power.  khepri/alternative.py:235-245."""
    from numpy.lib.scimath import sqrt as csqrt
    k0 = TWOPI / wavelength
    kzi = np.conj(csqrt(k0 ** 2 * epsi - kp[0] ** 2 - kp[1] ** 2)) / k0
    sx, sy = np.split(c_output, 2)
    kx, ky, kz = k_vectors(g, kp, wavelength, epse)
    with np.errstate(divide="ignore", invalid="ignore"):
        sz = -(kx * sx + ky * sy) / kz
        t = kz.real / np.real(kzi) * (np.abs(sx) ** 2 + np.abs(sy) ** 2 + np.abs(sz) ** 2)
    if only_total:
        return np.sum(t)
    return np.sum(t), t


# --------------------------------------------------------------------------
# fields
# --------------------------------------------------------------------------
def translate_mode_amplitudes(Sl, Sr, c1p, c1m, c2p):
    """khepri/fields.py:18-27."""
    F = np.eye(len(c1p)) - Sl[1, 1] @ Sr[0, 0]
    cdp = np.linalg.solve(F, Sl[1, 0] @ c1p)
    cdm = Sr[0, 0] @ cdp
    return cdp, cdm


def eigenbasis_matrix(W, V):
    """[[W, W], [-V, V]].  khepri/fields.py:46-51."""
    return np.block([[W, W], [-V, V]])


def fourier_fields_from_modes(RI, LI, R0, amps, zbar):
    """khepri/fields.py:53-62 (|L| clamp at 1e14)."""
    arg = LI * zbar
    L = np.hstack((np.exp(arg), np.exp(-arg)))
    big = np.abs(L) > 1e14
    L[big] /= np.abs(L[big]) / 1e14
    return np.split(RI @ (L * np.linalg.solve(RI, R0 @ np.hstack(amps))), 4)


def longitudinal_fields(tf, kx, ky, IC=1.0):
    """khepri/fields.py:68-76."""
    sx, sy, ux, uy = tf
    uz = -1j * (kx * sy - ky * sx)
    if np.isscalar(IC):
        sz = -1j * IC * (kx * uy - ky * ux)
    else:
        sz = -1j * IC @ (kx * uy - ky * ux)
    return sz, uz


def idft(ffield, kx, ky, x, y):
    """sum_G f_G exp(i(kx x + ky y)).  khepri/fourier.py:136-142."""
    shp = x.shape
    xf, yf = x.ravel(), y.ravel()
    ph = np.exp(1j * (kx[:, None] * xf[None, :] + ky[:, None] * yf[None, :]))
    return (ffield[:, None] * ph).sum(axis=0).reshape(shp)


# --------------------------------------------------------------------------
# extended (twisted) bases
# --------------------------------------------------------------------------
def joint_subspace(mats, kind=0):
    """Scatter K (2N x 2N) matrices into one (2N^2 x 2N^2) matrix.
    Restates the as_strided views of khepri/extension.py:10-51 as index maps:
      kind 0: (bi*N^2 + m*N + a, bj*N^2 + m*N + b) <- mats[m][bi*N + a, bj*N + b]
      kind 1: (bi*N^2 + a*N + m, bj*N^2 + b*N + m) <- mats[m][bi*N + a, bj*N + b]
    """
    N = mats[0].shape[0] // 2
    out = np.zeros((2 * N * N, 2 * N * N), dtype=mats[0].dtype)
    a = np.arange(N)
    for m, M in enumerate(mats):
        for bi in range(2):
            for bj in range(2):
                blk = M[bi * N:(bi + 1) * N, bj * N:(bj + 1) * N]
                if kind == 0:
                    r = bi * N * N + m * N + a
                    c = bj * N * N + m * N + a
                elif kind == 1:
                    r = bi * N * N + a * N + m
                    c = bj * N * N + a * N + m
                else:
                    raise NotImplementedError("two lattices at most")
                out[np.ix_(r, c)] = blk
    return out


def joint_subspace_S(Ss, kind=0):
    """khepri/extension.py:54-63."""
    return np.array([[joint_subspace([S[i, j] for S in Ss], kind) for j in range(2)]
                     for i in range(2)])

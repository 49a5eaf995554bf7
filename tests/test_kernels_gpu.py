"""Kernel-level parity on the GPU: each libkhx entry point against numpy /
the oracle on the same seeded inputs."""
import numpy as np
import pytest
import torch

from oracle import rcwa
from khepri_amd import workloads as W

pytestmark = pytest.mark.gpu


def crand(rng, *shape):
    return rng.normal(size=shape) + 1j * rng.normal(size=shape)


def dev(a, gpu):
    return torch.as_tensor(np.ascontiguousarray(a), device=gpu)


@pytest.mark.parametrize("M,N,K,batch", [(50, 50, 50, 3), (162, 162, 162, 2), (13, 27, 9, 4), (162, 81, 81, 2)])
def test_zgemm(gpu, M, N, K, batch):
    from khepri_amd import ops
    rng = np.random.default_rng(M + N + K)
    A, B, C0 = crand(rng, batch, M, K), crand(rng, batch, K, N), crand(rng, batch, M, N)
    out = ops.matmul(dev(A, gpu), dev(B, gpu), alpha=1 - 0.5j, beta=0.25 + 1j, C=dev(C0, gpu).clone())
    ref = (1 - 0.5j) * A @ B + (0.25 + 1j) * C0
    np.testing.assert_allclose(out.cpu().numpy(), ref, atol=1e-12 * np.abs(ref).max())
    # broadcast B
    out = ops.matmul(dev(A, gpu), dev(np.ascontiguousarray(B[0]), gpu))
    np.testing.assert_allclose(out.cpu().numpy(), A @ B[0], atol=1e-12 * np.abs(A @ B[0]).max())


@pytest.mark.parametrize("n,m,batch", [(1, 1, 3), (50, 50, 4), (162, 324, 2), (97, 3, 3), (130, 64, 3),
                                       (242, 100, 2), (339, 70, 1), (1250, 300, 1)])
def test_zgesv(gpu, n, m, batch):
    from khepri_amd import ops
    rng = np.random.default_rng(n * 7 + m)
    A, B = crand(rng, batch, n, n), crand(rng, batch, n, m)
    X, info = ops.solve(dev(A, gpu), dev(B, gpu))
    assert (info.cpu().numpy() == 0).all()
    ref = np.linalg.solve(A, B)
    np.testing.assert_allclose(X.cpu().numpy(), ref, atol=1e-10 * np.abs(ref).max())


def test_zgesv_singular_reports_info(gpu):
    from khepri_amd import ops
    A = np.zeros((2, 4, 4), dtype=complex)
    A[0] = np.eye(4)
    A[1, :, :2] = 1.0
    _, info = ops.solve(dev(A, gpu), dev(np.ones((2, 4, 1), dtype=complex), gpu))
    info = info.cpu().numpy()
    assert info[0] == 0 and info[1] > 0
    # the same through the multi-launch LU with right-hand sides solved after the factorization
    A = np.zeros((2, 100, 100), dtype=complex)
    A[0] = np.eye(100)
    A[1] = np.eye(100)
    A[1, 40, 40] = 0.0
    _, info = ops.solve(dev(A, gpu), dev(np.ones((2, 100, 64), dtype=complex), gpu))
    info = info.cpu().numpy()
    assert info[0] == 0 and info[1] == 41


@pytest.mark.parametrize("n,batch", [(1, 2), (2, 3), (5, 4), (50, 6), (162, 2)])
def test_eig_random(gpu, n, batch):
    from khepri_amd import ops
    rng = np.random.default_rng(n)
    A = crand(rng, batch, n, n)
    lam, Wm, info = ops.eig(dev(A, gpu))
    assert (info.cpu().numpy() == 0).all()
    lam, Wm = lam.cpu().numpy(), Wm.cpu().numpy()
    for b in range(batch):
        res = np.abs(A[b] @ Wm[b] - Wm[b] * lam[b]).max() / np.abs(A[b]).max()
        assert res < 1e-12 * max(1, n / 10)
        np.testing.assert_allclose(np.linalg.norm(Wm[b], axis=0), 1.0, atol=1e-13)
        ref = np.linalg.eigvals(A[b])
        d = np.abs(lam[b][:, None] - ref[None, :])
        assert d.min(axis=1).max() < 1e-10 * np.abs(ref).max()


def test_eig_structured_degenerate(gpu):
    """P*Q of the disc layer at normal incidence has degenerate eigenvalue pairs."""
    from khepri_amd import ops
    pix = W.disc_pixmap()
    C = rcwa.convolution_matrix(pix, (5, 5))
    g = rcwa.g_vectors((5, 5), np.eye(2))
    kx, ky, _ = rcwa.k_vectors(g, (0, 0), 2.0)
    P, Q = rcwa.structured_pq(kx, ky, C)
    M = P @ Q
    lam, Wm, info = ops.eig(dev(M[None], gpu))
    lam, Wm = lam.cpu().numpy()[0], Wm.cpu().numpy()[0]
    assert np.abs(M @ Wm - Wm * lam).max() / np.abs(M).max() < 1e-12
    assert np.linalg.cond(Wm) < 1e6


@pytest.mark.parametrize("pw,cplx", [((5, 5), False), ((9, 9), False), ((3, 7), True)])
def test_convmat(gpu, pw, cplx):
    from khepri_amd import ops
    rng = np.random.default_rng(5)
    pix = W.disc_pixmap()
    if cplx:
        pix = pix + 1j * rng.uniform(0, 0.1, pix.shape)
    C = ops.convmat(dev(pix, gpu), pw)[0].cpu().numpy()
    ref = rcwa.convolution_matrix(pix, pw)
    np.testing.assert_allclose(C, ref, atol=1e-14 * np.abs(ref).max())
    # gather is exact: entries with equal (dp, dq) are bit-identical
    P = pw[0]
    assert C[0, 0] == C[P + 1, P + 1]


def _layer_inputs(pw, kp, wl):
    g = rcwa.g_vectors(pw, np.eye(2))
    kx, ky, _ = rcwa.k_vectors(g, kp, wl)
    return g, kx, ky


@pytest.mark.parametrize("env", [{}, {"KHX_LAYER_AINV": "0"}, {"KHX_LAYER_WINV": "0"}, {"KHX_LAYER_VIDENT": "1"}],
                         ids=["default", "AINV=0", "WINV=0", "VIDENT=1"])
@pytest.mark.parametrize("pw,kp,wl", [((5, 5), (0.0, 0.0), 2.0), ((5, 5), (0.4, -0.3), 1.7),
                                      ((9, 9), (0.2, 0.1), 1.9)])
def test_patterned_layer(gpu, monkeypatch, pw, kp, wl, env):
    """The layer S-matrix against the oracle's build_scatmat: by default W^-1
    from the eigensolver's Schur vectors and triangular eigenvectors and
    [G1 | G2] through A^-1; also with the reference's solves for either
    (KHX_LAYER_WINV=0, KHX_LAYER_AINV=0) and with V^-1 V0 through the
    identity V^-1 = diag(1/lam) W^-1 P (KHX_LAYER_VIDENT=1)."""
    from khepri_amd import ops
    for k, v in env.items():
        monkeypatch.setenv(k, v)
    pix = W.disc_pixmap()
    C = rcwa.convolution_matrix(pix, pw)
    g, kx, ky = _layer_inputs(pw, kp, wl)
    k0 = 2 * np.pi / wl
    Wr, Vr, Lr = rcwa.structured_layer_modes(kx, ky, C)
    W0, V0 = rcwa.free_space_eigenmodes(kx, ky)
    Sref = rcwa.build_scatmat(Wr, Vr, W0, V0, Lr, 0.55, k0)
    Cd = dev(C, gpu)
    IC, _ = ops.inverse(Cd[None])
    k0d = dev(np.array([k0 * 0.55]), gpu)
    S, (Wm, Vm, L), info = ops.patterned_layer(Cd, IC[0], dev(kx[None], gpu), dev(ky[None], gpu), k0d, modes=True)
    assert (info.cpu().numpy() == 0).all()
    np.testing.assert_allclose(S.cpu().numpy()[0], Sref, atol=1e-10)
    # modes are a valid eigen-decomposition with the principal branch
    P, Q = rcwa.structured_pq(kx, ky, C)
    Wm, L = Wm.cpu().numpy()[0], L.cpu().numpy()[0]
    assert np.abs(P @ Q @ Wm - Wm * L ** 2).max() / np.abs(P @ Q).max() < 1e-11
    assert (L.real >= 0).all()


@pytest.mark.parametrize("kind,eps", [("uniform", 4.0 - 0.2j), ("uniform", 1.0), ("reflexion", 2.25),
                                      ("transmission", 1.44)])
def test_analytic_layers(gpu, kind, eps):
    from khepri_amd import ops
    pw, kp, wl, d = (3, 3), (0.5, 0.3), 1.3, 0.7
    g, kx, ky = _layer_inputs(pw, kp, wl)
    W0, V0 = rcwa.free_space_eigenmodes(kx, ky)
    k0 = 2 * np.pi / wl
    if kind == "uniform":
        Wr, Vr, Lr = rcwa.uniform_layer_modes(kx, ky, eps)
        Sref = rcwa.build_scatmat(Wr, Vr, W0, V0, Lr, d, k0)
    elif kind == "reflexion":
        Sref, Wr, Vr, Lr = rcwa.halfspace_reflection(kx, ky, W0, V0, eps)
    else:
        Sref, Wr, Vr, Lr = rcwa.halfspace_transmission(kx, ky, W0, V0, eps)
    S, (Wm, Vm, L) = ops.analytic_layer(kind, dev(kx[None], gpu), dev(ky[None], gpu), eps,
                                        dev(np.array([k0 * d]), gpu), modes=True)
    np.testing.assert_allclose(S.cpu().numpy()[0], Sref, atol=1e-12)
    np.testing.assert_allclose(Vm.cpu().numpy()[0], Vr, atol=1e-12)
    np.testing.assert_allclose(L.cpu().numpy()[0], Lr, atol=1e-12)


def test_redheffer(gpu):
    from khepri_amd import ops
    rng = np.random.default_rng(11)
    n, B = 50, 3
    SA = 0.2 * crand(rng, B, 2, 2, n, n)
    SB = 0.2 * crand(rng, B, 2, 2, n, n)
    S, info = ops.redheffer(dev(SA, gpu), dev(SB, gpu))
    ref = np.array([rcwa.redheffer(SA[b], SB[b]) for b in range(B)])
    np.testing.assert_allclose(S.cpu().numpy(), ref, atol=1e-12)


def test_poynting_matches_oracle(gpu):
    from khepri_amd import ops
    from oracle import crystal as OC
    pix = W.disc_pixmap(64, 6.0, 0.3, 1.0)
    cl = OC.Crystal((3, 3), epsi=1.0, epse=2.25)
    cl.add_layer_pixmap("P", pix, 0.3)
    cl.set_device(["P"])
    cl.set_source(1 / 0.6, 1.0, 0.4, 25.0, 70.0)
    cl.solve()
    (R, Ro), (T, To) = cl.poynting_flux_end(only_total=False)
    kp = np.array([cl.kp], dtype=complex)
    out = ops.poynting(dev(cl.Stot[None], gpu), dev(cl.expansion.g, gpu), dev(kp, gpu),
                       dev(np.array([cl.wavelength]), gpu), dev(np.array([1.0 + 0j]), gpu),
                       dev(np.array([0.4 + 0j]), gpu), 1.0, 2.25, orders=True)
    Rg, Tg, Rog, Tog = (x.cpu().numpy()[0] for x in out)
    assert abs(Rg - R.real) < 1e-13 and abs(Tg - T.real) < 1e-13
    np.testing.assert_allclose(Rog, Ro.real, atol=1e-14)
    np.testing.assert_allclose(Tog, To.real, atol=1e-14)


@pytest.mark.parametrize("n,m,batch", [(1, 1, 2), (50, 1, 3), (162, 7, 2), (162, 1, 3), (242, 3, 2), (97, 4, 2),
                                       (1250, 1, 1), (162, 162, 2), (131, 80, 2)])
def test_lu_factor_solve(gpu, n, m, batch):
    from khepri_amd import ops
    rng = np.random.default_rng(n + 100 * m)
    A, B = crand(rng, batch, n, n), crand(rng, batch, n, m)
    lu, piv, info = ops.lu_factor(dev(A, gpu))
    assert (info.cpu().numpy() == 0).all()
    X = ops.lu_solve(lu, piv, dev(B, gpu)).cpu().numpy()
    ref = np.linalg.solve(A, B)
    np.testing.assert_allclose(X, ref, atol=1e-10 * np.abs(ref).max())
    # factor and solve in one pass: same solution, factors reusable
    lu2, piv2, info2, X2 = ops.lu_factor_solve(dev(A, gpu), dev(B, gpu))
    assert (info2.cpu().numpy() == 0).all()
    np.testing.assert_allclose(X2.cpu().numpy(), ref, atol=1e-10 * np.abs(ref).max())
    X3 = ops.lu_solve(lu2, piv2, dev(B, gpu)).cpu().numpy()
    np.testing.assert_allclose(X3, ref, atol=1e-10 * np.abs(ref).max())
    # the factors reproduce P A = L U
    lu, piv = lu.cpu().numpy(), piv.cpu().numpy()
    for b in range(batch):
        L = np.tril(lu[b], -1) + np.eye(n)
        U = np.triu(lu[b])
        PA = A[b].copy()
        for k in range(n):
            PA[[k, piv[b, k]]] = PA[[piv[b, k], k]]
        np.testing.assert_allclose(L @ U, PA, atol=1e-12 * np.abs(A[b]).max())


def test_strided_views_and_elementwise(gpu):
    from khepri_amd import ops
    rng = np.random.default_rng(21)
    S = crand(rng, 3, 2, 2, 9, 9)
    St = dev(S, gpu)
    out = ops.matmul(St[:, 0, 1], St[:, 1, 0])
    np.testing.assert_allclose(out.cpu().numpy(), S[:, 0, 1] @ S[:, 1, 0], atol=1e-12)
    dl, dr = crand(rng, 3, 9), crand(rng, 3, 9)
    d = ops.diag_scale(St[:, 1, 1], dev(dl, gpu), dev(dr, gpu)).cpu().numpy()
    np.testing.assert_allclose(d, dl[:, :, None] * S[:, 1, 1] * dr[:, None, :], atol=1e-13)
    im = ops.identity_minus(dev(S[:, 0, 0].copy(), gpu)).cpu().numpy()
    np.testing.assert_allclose(im, np.eye(9) - S[:, 0, 0], atol=0)


def test_propagation_layer_is_exact_free_space(gpu):
    """A uniform eps = 1 layer: S = [[0, X], [X, 0]] (the dense analytic S agrees)."""
    from khepri_amd import ops
    g, kx, ky = _layer_inputs((3, 3), (0.7, 0.2), 1.1)
    k0d = dev(np.array([2 * np.pi / 1.1 * 0.9]), gpu)
    X = ops.propagation(dev(kx[None], gpu), dev(ky[None], gpu), k0d).cpu().numpy()[0]
    S, _ = ops.analytic_layer("uniform", dev(kx[None], gpu), dev(ky[None], gpu), 1.0, k0d)
    S = S.cpu().numpy()[0]
    np.testing.assert_allclose(S[0, 1], np.diag(X), atol=1e-14)
    np.testing.assert_allclose(S[0, 0], 0, atol=1e-14)


@pytest.mark.parametrize("n,m,batch", [(24, 5, 3), (50, 100, 2), (67, 1, 2), (162, 324, 2), (242, 3, 1), (338, 20, 1)])
def test_blocked_lu_paths(gpu, n, m, batch):
    """Panel-blocked LU (n >= 24): solve and factor/solve agree with numpy, incl.
    ragged last panels and right-hand-side tiles."""
    from khepri_amd import ops
    rng = np.random.default_rng(n + m)
    A, B = crand(rng, batch, n, n), crand(rng, batch, n, m)
    ref = np.linalg.solve(A, B)
    X, info = ops.solve(dev(A, gpu), dev(B, gpu))
    assert (info.cpu().numpy() == 0).all()
    np.testing.assert_allclose(X.cpu().numpy(), ref, atol=1e-9 * np.abs(ref).max())
    lu, piv, info = ops.lu_factor(dev(A, gpu))
    X2 = ops.lu_solve(lu, piv, dev(B, gpu)).cpu().numpy()
    np.testing.assert_allclose(X2, ref, atol=1e-9 * np.abs(ref).max())


def test_eig_slots_fill_whole_cus(gpu):
    from khepri_amd import ops
    cus = torch.cuda.get_device_properties(0).multi_processor_count
    for n in (50, 162, 242):
        slots = ops.eig_slots(n)
        assert slots >= cus and slots % cus == 0, (n, slots, cus)


@pytest.mark.parametrize("reg_chase", ["0", "1"])
def test_eig_chase_variants(gpu, monkeypatch, reg_chase):
    """Both QR chases (LDS block chase, register chase) on layer matrices with
    windows that span several chase blocks and on random matrices."""
    from khepri_amd import ops
    monkeypatch.setenv("KHX_EIG_REG_CHASE", reg_chase)
    pix = W.disc_pixmap()
    C = rcwa.convolution_matrix(pix, (9, 9))
    g = rcwa.g_vectors((9, 9), np.eye(2))
    mats = []
    for wl, kp in [(2.0, (0.1, 0.3)), (1.7, (0.0, 0.0)), (2.2, (0.4, -0.2))]:
        kx, ky, _ = rcwa.k_vectors(g, kp, wl)
        P, Q = rcwa.structured_pq(kx, ky, C)
        mats.append(P @ Q)
    rng = np.random.default_rng(9)
    mats += [crand(rng, 162, 162), crand(rng, 100, 100)[:, :]]
    for M in mats:
        lam, Wm, info = ops.eig(dev(M[None], gpu))
        assert int(info.cpu()[0]) == 0
        lam, Wm = lam.cpu().numpy()[0], Wm.cpu().numpy()[0]
        assert np.abs(M @ Wm - Wm * lam).max() / np.abs(M).max() < 1e-12
        ref = np.linalg.eigvals(M)
        d = np.abs(lam[:, None] - ref[None, :])
        assert d.min(axis=1).max() < 1e-10 * np.abs(ref).max()


@pytest.mark.parametrize("n", [97, 130, 162, 250])
@pytest.mark.parametrize("recfactor", ["0", "1"])
def test_lu_recursive_paths(gpu, monkeypatch, n, recfactor):
    """Multi-launch LU with the factorization recursive or panel by panel and
    the triangular solves after it (right-hand sides >= 64)."""
    from khepri_amd import ops
    monkeypatch.setenv("KHX_LU_RECFACTOR", recfactor)
    rng = np.random.default_rng(n + 7)
    A, B = crand(rng, 2, n, n), crand(rng, 2, n, 64)
    X, info = ops.solve(dev(A, gpu), dev(B, gpu))
    assert (info.cpu().numpy() == 0).all()
    ref = np.linalg.solve(A, B)
    np.testing.assert_allclose(X.cpu().numpy(), ref, atol=1e-10 * np.abs(ref).max())
    lu, piv, info = ops.lu_factor(dev(A, gpu))
    lu, piv = lu.cpu().numpy(), piv.cpu().numpy()
    for b in range(2):
        L = np.tril(lu[b], -1) + np.eye(n)
        U = np.triu(lu[b])
        PA = A[b].copy()
        for k in range(n):
            PA[[k, piv[b, k]]] = PA[[piv[b, k], k]]
        np.testing.assert_allclose(L @ U, PA, atol=1e-12 * np.abs(A[b]).max())


@pytest.mark.parametrize("direct", ["0", "1"])
@pytest.mark.parametrize("shape,pw", [((128, 128), (13, 13)), ((37, 52), (5, 7)), ((64, 64), (11, 11))])
def test_convmat_dft_variants(gpu, monkeypatch, direct, shape, pw):
    """Separable and direct DFT of the pixmap give the oracle's coefficients."""
    from khepri_amd import ops
    if direct == "1":
        monkeypatch.setenv("KHX_DIRECT_DFT", "1")
    rng = np.random.default_rng(sum(shape))
    pix = rng.uniform(1.0, 12.0, size=shape) + 1j * rng.uniform(0.0, 0.2, size=shape)
    C = ops.convmat(dev(pix, gpu), pw)[0].cpu().numpy()
    ref = rcwa.convolution_matrix(pix, pw)
    np.testing.assert_allclose(C, ref, atol=2e-15 * np.abs(ref).max() * 8)


EIG_VARIANTS = [
    {"KHX_EIG_MULTISHIFT": "1"}, {"KHX_EIG_MULTISHIFT": "1", "KHX_EIG_QR_WG": "3"},
    {"KHX_EIG_MULTISHIFT": "1", "KHX_EIG_QR_WG": "4"}, {"KHX_EIG_MULTISHIFT": "1", "KHX_EIG_THREADS": "512"},
    {"KHX_EIG_REG_CHASE": "0"}, {"KHX_EIG_QR_THREADS": "128"},
    {"KHX_EIG_CHASE_WAVE": "0", "KHX_EIG_REG_CHASE": "0"},
    {"KHX_EIG_THREADS": "64"}, {"KHX_EIG_THREADS": "128"}, {"KHX_EIG_THREADS": "512"},
    {"KHX_EIG_THREADS": "1024"}, {"KHX_EIG_LDS_PAD": "8192"},
]


@pytest.mark.parametrize("env", EIG_VARIANTS, ids=lambda e: "-".join(f"{k[4:]}={v}" for k, v in e.items()))
def test_eig_tuning_variants(gpu, monkeypatch, env):
    """Every launch-shape / chase variant selectable through the environment
    solves the layer eigenproblem (n = 50 in LDS, n = 162 in HBM)."""
    from khepri_amd import ops
    for k, v in env.items():
        monkeypatch.setenv(k, v)
    pix = W.disc_pixmap()
    for pw in ((5, 5), (9, 9)):
        C = rcwa.convolution_matrix(pix, pw)
        g = rcwa.g_vectors(pw, np.eye(2))
        mats = []
        for wl, kp in [(2.0, (0.1, 0.3)), (1.7, (0.0, 0.0))]:
            kx, ky, _ = rcwa.k_vectors(g, kp, wl)
            P, Q = rcwa.structured_pq(kx, ky, C)
            mats.append(P @ Q)
        M = np.array(mats)
        lam, Wm, info = ops.eig(dev(M, gpu))
        assert (info.cpu().numpy() == 0).all()
        lam, Wm = lam.cpu().numpy(), Wm.cpu().numpy()
        for b in range(len(M)):
            assert np.abs(M[b] @ Wm[b] - Wm[b] * lam[b]).max() / np.abs(M[b]).max() < 1e-12
            ref = np.linalg.eigvals(M[b])
            assert np.abs(lam[b][:, None] - ref[None, :]).min(axis=1).max() < 1e-10 * np.abs(ref).max()


LU_VARIANTS = [{"KHX_LU_NB": "32"}, {"KHX_LU_PATH": "0"}, {"KHX_LU_PATH": "1"}, {"KHX_LU_PATH": "2"},
               {"KHX_LU_UNBLOCKED": "1"}, {"KHX_LU_RECSOLVE": "0"}, {"KHX_LU_NB": "32", "KHX_LU_RECFACTOR": "1"},
               {"KHX_LU_PANEL_REG": "0"}]


@pytest.mark.parametrize("pw", [(3, 3), (5, 5), (1, 31)])
def test_eig_small_register_chase_matches_lds_chase(gpu, monkeypatch, pw):
    """n <= 64 (H in LDS): the register chase over the whole window against
    the barrier-per-rotation LDS chase (KHX_EIG_REG_CHASE=0), on layer
    matrices (n = 18, 50, 62) and random ones: the same rotations, so the
    eigenvalues agree to rounding; residuals at rounding level both ways."""
    from khepri_amd import ops
    pix = W.disc_pixmap()
    C = rcwa.convolution_matrix(pix, pw)
    g = rcwa.g_vectors(pw, np.eye(2))
    mats = []
    for wl, kp in [(2.0, (0.1, 0.3)), (1.7, (0.0, 0.0)), (0.9, (0.4, -0.2))]:
        kx, ky, _ = rcwa.k_vectors(g, kp, wl)
        P, Q = rcwa.structured_pq(kx, ky, C)
        mats.append(P @ Q)
    n = mats[0].shape[0]
    rng = np.random.default_rng(n)
    mats += [crand(rng, n, n), crand(rng, n, n)]
    M = np.stack(mats)
    out = {}
    for rc in ("0", "1"):
        monkeypatch.setenv("KHX_EIG_REG_CHASE", rc)
        lam, Wm, info = ops.eig(dev(M, gpu))
        assert (info.cpu().numpy() == 0).all()
        out[rc] = (lam.cpu().numpy(), Wm.cpu().numpy())
        for b in range(len(M)):
            assert np.abs(M[b] @ out[rc][1][b] - out[rc][1][b] * out[rc][0][b]).max() / np.abs(M[b]).max() < 1e-12
    for b in range(len(M)):
        ref = np.linalg.eigvals(M[b])
        for rc in ("0", "1"):
            d = np.abs(out[rc][0][b][:, None] - ref[None, :]).min(axis=1).max()
            assert d < 1e-10 * np.abs(ref).max()
    np.testing.assert_allclose(out["1"][0], out["0"][0], atol=1e-11 * np.abs(out["0"][0]).max())


@pytest.mark.parametrize("pw", [(7, 7), (9, 9), (11, 11), (13, 13)])
def test_eig_vectors_per_column_match_row_sweep(gpu, monkeypatch, pw):
    """The eigenvector stage with one thread per eigenvector (row blocks in
    registers) against the row-by-row sweep (KHX_EIG_VEC_ROWS=1): same
    operations in the same order, so eigenvalues and eigenvectors agree bit
    for bit; and the residual |M W - W diag(lam)| stays at rounding level."""
    from khepri_amd import ops
    pix = W.disc_pixmap()
    C = rcwa.convolution_matrix(pix, pw)
    g = rcwa.g_vectors(pw, np.eye(2))
    mats = []
    for wl, kp in [(2.0, (0.1, 0.3)), (1.7, (0.0, 0.0))]:
        kx, ky, _ = rcwa.k_vectors(g, kp, wl)
        P, Q = rcwa.structured_pq(kx, ky, C)
        mats.append(P @ Q)
    n = mats[0].shape[0]
    mats.append(crand(np.random.default_rng(n), n, n))
    M = np.stack(mats)
    out = {}
    for rows in ("1", "0"):
        monkeypatch.setenv("KHX_EIG_VEC_ROWS", rows)
        lam, Wm, info = ops.eig(dev(M, gpu))
        assert (info.cpu().numpy() == 0).all()
        out[rows] = (lam.cpu().numpy(), Wm.cpu().numpy())
    np.testing.assert_array_equal(out["0"][0], out["1"][0])
    np.testing.assert_array_equal(out["0"][1], out["1"][1])
    lam, Wm = out["0"]
    for b in range(len(M)):
        assert np.abs(M[b] @ Wm[b] - Wm[b] * lam[b]).max() / np.abs(M[b]).max() < 1e-11


@pytest.mark.parametrize("mode", ["1", "2"])
@pytest.mark.parametrize("pw", [(7, 7), (9, 9)])
def test_eig_paired_schur_matches_single(gpu, monkeypatch, pw, mode):
    """The Schur stage with two matrices per workgroup (one chase wavefront
    for both, blocks of 31; KHX_EIG_PAIR=1: 256 threads, 2: 512) against one
    matrix per workgroup (KHX_EIG_PAIR=0, blocks of 40) on an odd batch larger than the CU count
    (the paired launch's shape; the last workgroup holds one matrix): layer
    matrices and random ones in a shuffled order, so partners deflate at
    different sweeps.  Same rotations in the same order per matrix, so
    eigenvalues, eigenvectors and status words agree bit for bit; so do the
    status words of a run stopped by a low iteration cap."""
    from khepri_amd import ops
    cus = torch.cuda.get_device_properties(0).multi_processor_count
    pix = W.disc_pixmap()
    C = rcwa.convolution_matrix(pix, pw)
    g = rcwa.g_vectors(pw, np.eye(2))
    rng = np.random.default_rng(sum(pw))
    mats = []
    for i in range(12):
        kx, ky, _ = rcwa.k_vectors(g, tuple(rng.uniform(-0.4, 0.4, 2)), rng.uniform(0.9, 2.2))
        P, Q = rcwa.structured_pq(kx, ky, C)
        mats.append(P @ Q)
    n = mats[0].shape[0]
    mats += [crand(rng, n, n) for _ in range(4)]
    batch = cus + 37
    M = np.stack(mats)[rng.integers(0, len(mats), batch)]
    out = {}
    for pair in ("0", mode):
        monkeypatch.setenv("KHX_EIG_PAIR", pair)
        lam, Wm, info = ops.eig(dev(M, gpu))
        out[pair] = (lam.cpu().numpy(), Wm.cpu().numpy(), info.cpu().numpy())
    assert (out[mode][2] == 0).all()
    for a, b in zip(out["0"], out[mode]):
        np.testing.assert_array_equal(a, b)
    lam, Wm, _ = out[mode]
    for b in list(range(6)) + [batch - 1]:
        assert np.abs(M[b] @ Wm[b] - Wm[b] * lam[b]).max() / np.abs(M[b]).max() < 1e-11
    monkeypatch.setenv("KHX_EIG_ITMAX", str(n // 2))
    infos = {}
    for pair in ("0", mode):
        monkeypatch.setenv("KHX_EIG_PAIR", pair)
        infos[pair] = ops.eig(dev(M, gpu))[2].cpu().numpy()
    assert (infos[mode] > 0).all()
    np.testing.assert_array_equal(infos["0"], infos[mode])


@pytest.mark.parametrize("pw,big", [((7, 7), False), ((9, 9), False), ((9, 9), True), ((11, 11), False),
                                    ((13, 13), False)])
def test_eig_hessenberg_panel_in_lds_matches_global(gpu, monkeypatch, pw, big):
    """The Hessenberg panel factored in LDS against the same steps on H in
    global memory (KHX_EIG_HESS_LDS=0): same operations in the same order, so
    eigenvalues and eigenvectors agree bit for bit.  Small batches run 512
    threads per matrix (the panel copy may enlarge the launch's LDS), a batch
    larger than the CU count 256 threads (the copy only where it fits)."""
    from khepri_amd import ops
    cus = torch.cuda.get_device_properties(0).multi_processor_count
    pix = W.disc_pixmap()
    C = rcwa.convolution_matrix(pix, pw)
    g = rcwa.g_vectors(pw, np.eye(2))
    rng = np.random.default_rng(sum(pw) + big)
    mats = []
    for i in range(3):
        kx, ky, _ = rcwa.k_vectors(g, tuple(rng.uniform(-0.4, 0.4, 2)), rng.uniform(0.9, 2.2))
        P, Q = rcwa.structured_pq(kx, ky, C)
        mats.append(P @ Q)
    n = mats[0].shape[0]
    mats.append(crand(rng, n, n))
    M = np.stack(mats)
    if big:
        M = M[rng.integers(0, len(M), cus + 5)]
    out = {}
    for flag in ("0", "1"):
        monkeypatch.setenv("KHX_EIG_HESS_LDS", flag)
        lam, Wm, info = ops.eig(dev(M, gpu))
        out[flag] = (lam.cpu().numpy(), Wm.cpu().numpy(), info.cpu().numpy())
    assert (out["1"][2] == 0).all()
    for a, b in zip(out["0"], out["1"]):
        np.testing.assert_array_equal(a, b)
    lam, Wm, _ = out["1"]
    for b in range(4):
        assert np.abs(M[b] @ Wm[b] - Wm[b] * lam[b]).max() / np.abs(M[b]).max() < 1e-11


@pytest.mark.parametrize("n,m", [(162, 64), (162, 324), (250, 100)])
def test_lu_mixed_refined(gpu, monkeypatch, n, m):
    """KHX_LU_MIXED=1: FP64 LU, triangular-solve products on the FP32 matrix
    cores (complex-float operands, error ~1e-7), one FP64 residual refinement
    step; the solution against numpy at 1e-10 of its size, and the FP32
    products alone (before refinement) would miss that by orders of magnitude."""
    from khepri_amd import ops
    monkeypatch.setenv("KHX_LU_MIXED", "1")
    rng = np.random.default_rng(n * m)
    A = crand(rng, 3, n, n) + np.sqrt(n) * np.eye(n)
    A[2] = crand(rng, n, n)            # a general random matrix too
    B = crand(rng, 3, n, m)
    ref = np.linalg.solve(A, B)
    X, info = ops.solve(dev(A, gpu), dev(B, gpu))
    assert (info.cpu().numpy() == 0).all()
    X = X.cpu().numpy()
    for b in range(3):
        err = np.abs(X[b] - ref[b]).max() / np.abs(ref[b]).max()
        assert err < 1e-10, (b, err)
    monkeypatch.setenv("KHX_LU_MIXED", "0")
    X64, _ = ops.solve(dev(A, gpu), dev(B, gpu))
    assert np.abs(X64.cpu().numpy() - ref).max() / np.abs(ref).max() < 1e-11


@pytest.mark.parametrize("n,nb,th", [(97, 16, 256), (162, 16, 256), (250, 16, 256), (338, 16, 256),
                                     (600, 16, 256), (1250, 16, 256), (97, 32, 256), (162, 32, 256),
                                     (250, 32, 256), (338, 32, 256), (97, 16, 64), (162, 16, 64),
                                     (250, 16, 64), (162, 16, 128), (250, 16, 128)])
def test_lu_register_panel_matches_lds_panel(gpu, monkeypatch, n, nb, th):
    """The register-resident panel kernel (16 columns: panel rows in registers
    up to 512 rows on 256 threads, up to 1 536 on 512, and through
    KHX_LU_PANEL_TH up to 192 rows on one wavefront or 256 on 128 threads;
    32 columns: up to 256 rows) and the LDS (or, past 150 KB, global-memory)
    panel kernel make the same pivot choices and the same arithmetic:
    factors, pivots, solutions and status words agree bit for bit, including
    an exactly singular system (zero pivot) and a NaN column."""
    from khepri_amd import ops
    monkeypatch.setenv("KHX_LU_NB", str(nb))
    monkeypatch.setenv("KHX_LU_PANEL_TH", str(th))
    rng = np.random.default_rng(n)
    A = crand(rng, 4, n, n)
    A[1, :, 5] = 0.0                       # singular: zero pivot in column 5
    A[2, 7:, 3] = np.nan                   # a NaN column below the diagonal
    A[3, :, :] = A[3, :, :] + 1e3 * np.eye(n)
    out = {}
    for reg in ("0", "1"):
        monkeypatch.setenv("KHX_LU_PANEL_REG", reg)
        res = []
        for m in (5, 100):
            B = crand(np.random.default_rng(m), 4, n, m)
            X, info = ops.solve(dev(A, gpu), dev(B, gpu))
            res += [X.cpu().numpy(), info.cpu().numpy()]
        lu, piv, info = ops.lu_factor(dev(A, gpu))
        res += [lu.cpu().numpy(), piv.cpu().numpy(), info.cpu().numpy()]
        out[reg] = res
    for a, b in zip(out["0"], out["1"]):
        np.testing.assert_array_equal(a, b)
    info = out["1"][-1]
    assert info[0] == 0 and info[3] == 0 and info[1] > 0
    ref = np.linalg.solve(A[[0, 3]], crand(np.random.default_rng(5), 4, n, 5)[[0, 3]])
    np.testing.assert_allclose(out["1"][0][[0, 3]], ref, atol=1e-10 * np.abs(ref).max())


@pytest.mark.parametrize("env", LU_VARIANTS, ids=lambda e: "-".join(f"{k[4:]}={v}" for k, v in e.items()))
def test_lu_tuning_variants(gpu, monkeypatch, env):
    """Every LU path selectable through the environment: solve and factors."""
    from khepri_amd import ops
    for k, v in env.items():
        monkeypatch.setenv(k, v)
    rng = np.random.default_rng(17)
    for n, m in ((50, 70), (162, 5), (162, 100)):
        A, B = crand(rng, 2, n, n), crand(rng, 2, n, m)
        X, info = ops.solve(dev(A, gpu), dev(B, gpu))
        assert (info.cpu().numpy() == 0).all()
        ref = np.linalg.solve(A, B)
        np.testing.assert_allclose(X.cpu().numpy(), ref, atol=1e-10 * np.abs(ref).max())
        lu, piv, info = ops.lu_factor(dev(A, gpu))
        Xs = ops.lu_solve(lu, piv, dev(B, gpu)).cpu().numpy()
        np.testing.assert_allclose(Xs, ref, atol=1e-10 * np.abs(ref).max())

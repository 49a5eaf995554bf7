"""Adjoint gradient on the GPU: the backward kernels against numpy / torch CPU
autograd, and d(T)/d(pixel eps) against the FP64 CPU autograd reference
(oracle/autograd.py) and central finite differences."""
import numpy as np
import pytest
import torch

from conftest import golden
from oracle import autograd as AG
from khepri_amd import workloads as W

pytestmark = pytest.mark.gpu


def crand(rng, *shape):
    return rng.normal(size=shape) + 1j * rng.normal(size=shape)


def dev(a, gpu):
    return torch.as_tensor(np.ascontiguousarray(a), device=gpu)


def _op(X, o):
    return {"N": X, "T": np.swapaxes(X, -1, -2), "C": np.conj(np.swapaxes(X, -1, -2))}[o]


@pytest.mark.parametrize("M,N,K", [(37, 50, 23), (70, 90, 65)])
def test_zgemm_op_all_variants(gpu, M, N, K):
    from khepri_amd import ops
    rng = np.random.default_rng(M * N)
    for oa in "NTC":
        for ob in "NTC":
            A = crand(rng, 2, *((M, K) if oa == "N" else (K, M)))
            B = crand(rng, 2, *((K, N) if ob == "N" else (N, K)))
            C0 = crand(rng, 2, M, N)
            out = ops.matmul(dev(A, gpu), dev(B, gpu), alpha=0.5 + 1j, beta=-1.0, C=dev(C0, gpu).clone(), opa=oa,
                             opb=ob)
            ref = (0.5 + 1j) * _op(A, oa) @ _op(B, ob) - C0
            np.testing.assert_allclose(out.cpu().numpy(), ref, atol=1e-12 * np.abs(ref).max(), err_msg=oa + ob)


def test_conj_transpose_reduce_dot_axpby(gpu):
    from khepri_amd import ops
    rng = np.random.default_rng(5)
    A, B = crand(rng, 3, 45, 70), crand(rng, 3, 45, 70)
    np.testing.assert_array_equal(ops.conj_transpose(dev(A, gpu)).cpu().numpy(), np.conj(np.swapaxes(A, 1, 2)))
    r1 = ops.reduce_dot(dev(A, gpu), dev(B, gpu), 1).cpu().numpy()
    np.testing.assert_allclose(r1, (A * np.conj(B)).sum(2), atol=1e-12 * 70)
    r0 = ops.reduce_dot(dev(A, gpu), dev(B, gpu), 0)
    ops.reduce_dot(dev(A, gpu), dev(B, gpu), 0, out=r0, accumulate=True)
    np.testing.assert_allclose(r0.cpu().numpy(), 2 * (A * np.conj(B)).sum(1), atol=1e-12 * 90)
    out = ops.axpby(2 - 1j, dev(A, gpu), 0.5j, dev(B, gpu))
    np.testing.assert_allclose(out.cpu().numpy(), (2 - 1j) * A + 0.5j * B, atol=1e-14 * 10)


@pytest.mark.parametrize("pw,shape,cplx", [((3, 3), (16, 16), False), ((5, 7), (24, 20), True),
                                           ((11, 11), (64, 64), False)])
def test_convmat_adjoint_identity(gpu, pw, shape, cplx):
    """C is linear in eps: Re <Cbar, C(eps)> = <epsbar, eps> for every eps."""
    from khepri_amd import ops
    rng = np.random.default_rng(pw[0] * 13 + shape[0])
    N = pw[0] * pw[1]
    Cbar = crand(rng, 2, N, N)
    eps = crand(rng, 2, *shape) if cplx else rng.normal(size=(2,) + shape)
    C = ops.convmat(dev(eps, gpu), pw).cpu().numpy()
    epsbar = ops.convmat_adjoint(dev(Cbar, gpu), pw, shape, cplx).cpu().numpy()
    for b in range(2):
        lhs = np.real(np.vdot(Cbar[b], C[b]))
        rhs = np.real(np.vdot(epsbar[b], eps[b])) if cplx else float(np.sum(epsbar[b] * eps[b]))
        assert abs(lhs - rhs) < 1e-11 * max(1.0, abs(lhs)), (lhs, rhs)


def test_redheffer_bwd_matches_autograd(gpu):
    from khepri_amd import ops
    rng = np.random.default_rng(11)
    n, B = 18, 3
    SA = 0.3 * crand(rng, B, 2, 2, n, n)
    SB = 0.3 * crand(rng, B, 2, 2, n, n)
    G = crand(rng, B, 2, 2, n, n)
    SAb, SBb, info = ops.redheffer_bwd(dev(SA, gpu), dev(SB, gpu), dev(G, gpu))
    assert (info.cpu().numpy() == 0).all()
    for b in range(B):
        ta = torch.tensor(SA[b], requires_grad=True)
        tb = torch.tensor(SB[b], requires_grad=True)
        S = AG.redheffer([[ta[0, 0], ta[0, 1]], [ta[1, 0], ta[1, 1]]], [[tb[0, 0], tb[0, 1]], [tb[1, 0], tb[1, 1]]])
        L = sum(torch.real(torch.sum(torch.conj(torch.tensor(G[b, i, j])) * S[i][j])) for i in range(2)
                for j in range(2))
        L.backward()
        np.testing.assert_allclose(SAb[b].cpu().numpy(), ta.grad.numpy(), atol=1e-10 * np.abs(ta.grad.numpy()).max())
        np.testing.assert_allclose(SBb[b].cpu().numpy(), tb.grad.numpy(), atol=1e-10 * np.abs(tb.grad.numpy()).max())


@pytest.mark.parametrize("ainv", ["1", "0"])
@pytest.mark.parametrize("pw,theta", [((3, 3), 0.0), ((5, 5), 23.0)])
def test_patterned_layer_bwd_matches_autograd(gpu, monkeypatch, pw, theta, ainv):
    """L = Re <Sbar, S(eps)> for a random Sbar: the pixel gradient of the
    layer adjoint (eigen adjoint included) against torch CPU autograd, with
    and without the forward pass's cache, for both ways the forward forms
    A^-1 [X | B] (KHX_LAYER_AINV)."""
    from oracle import rcwa
    from khepri_amd import ops
    monkeypatch.setenv("KHX_LAYER_AINV", ainv)
    rng = np.random.default_rng(pw[0] + int(theta))
    shape = (20, 20)
    eps = rng.uniform(2, 12, size=shape)
    wl, depth = 1 / 0.57, 0.35
    g = rcwa.g_vectors(pw, rcwa.lattice_matrix("square"))
    kp = rcwa.compute_kplanar(1.0, wl, theta, 31.0)
    kx, ky, _ = rcwa.k_vectors(g, kp, wl)
    n = 2 * pw[0] * pw[1]
    Sbar = crand(rng, 1, 2, 2, n, n)
    # torch CPU autograd
    te = torch.tensor(eps, requires_grad=True)
    S = AG.patterned_layer_S(te, depth, kx, ky, 2 * np.pi / wl, pw)
    L = sum(torch.real(torch.sum(torch.conj(torch.tensor(Sbar[0, i, j])) * S[i][j])) for i in range(2)
            for j in range(2))
    L.backward()
    ref = te.grad.numpy()
    # GPU adjoint
    C = ops.convmat(dev(eps[None], gpu), pw)[0]
    IC, info = ops.inverse(C[None])
    kxd, kyd = dev(kx[None], gpu), dev(ky[None], gpu)
    k0d = torch.tensor([2 * np.pi / wl * depth], dtype=torch.float64, device=gpu)
    Sg, (Wm, Vm, lam), info = ops.patterned_layer(C, IC[0], kxd, kyd, k0d, modes=True)
    Sref = np.array([[S[i][j].detach().numpy() for j in range(2)] for i in range(2)])
    np.testing.assert_allclose(Sg[0].cpu().numpy(), Sref, atol=1e-10)
    Cbar, info = ops.patterned_layer_bwd(C, IC[0], kxd, kyd, k0d, Wm, Vm, lam, dev(Sbar, gpu))
    assert (info.cpu().numpy() == 0).all()
    got = ops.convmat_adjoint(Cbar, pw, shape).cpu().numpy()[0]
    np.testing.assert_allclose(got, ref, atol=1e-9 * np.abs(ref).max())
    # the same with the forward pass's intermediates kept (khx_patterned_layer_cached /
    # khx_patterned_layer_bwd_cached; [S11 | S12] then goes through T^-1): S and the
    # gradient to the same tolerances
    cache = torch.empty((1, 9, n, n), dtype=ops.CDT, device=gpu)
    Sc, (Wc, Vc, lamc), info = ops.patterned_layer(C, IC[0], kxd, kyd, k0d, modes=True, cache=cache)
    np.testing.assert_allclose(Sc[0].cpu().numpy(), Sref, atol=1e-10)
    Cbarc, info = ops.patterned_layer_bwd(C, IC[0], kxd, kyd, k0d, Wc, Vc, lamc, dev(Sbar, gpu), cache=cache)
    assert (info.cpu().numpy() == 0).all()
    gotc = ops.convmat_adjoint(Cbarc, pw, shape).cpu().numpy()[0]
    np.testing.assert_allclose(gotc, ref, atol=1e-9 * np.abs(ref).max())


def _crystal(pw, pixmaps, depths, gap=None, device="cuda:0"):
    from khepri_amd import Crystal
    cl = Crystal(pw, device=device)
    names = []
    for j, (p, d) in enumerate(zip(pixmaps, depths)):
        cl.add_layer_pixmap(f"L{j}", p, d)
        names.append(f"L{j}")
        if gap is not None and j == 0:
            cl.add_layer_uniform("G", gap[0], gap[1])
            names.append("G")
    cl.set_device(names, False)
    return cl, [f"L{j}" for j in range(len(pixmaps))]


@pytest.mark.parametrize("pw,theta,phi", [((3, 3), 0.0, 0.0), ((5, 5), 15.0, 60.0)])
def test_flux_gradient_small_vs_autograd(gpu, pw, theta, phi):
    from khepri_amd.adjoint import flux_gradient
    rng = np.random.default_rng(pw[0] * 3 + int(theta))
    pix = [rng.uniform(2, 12, size=(18, 18)) for _ in range(2)]
    depths = [0.3, 0.45]
    wls = 1 / np.array([0.5, 0.56, 0.63])
    cl, names = _crystal(pw, pix, depths, gap=(2.25, 0.2))
    cl.set_source(wls, 1.0, 0.0, theta, phi)
    R, T, grads = flux_gradient(cl, wT=1.0, wR=0.0)
    for b, wl in enumerate(wls):
        ts = [torch.tensor(p, requires_grad=True) for p in pix]
        layers = [("pixmap", ts[0], depths[0]), ("uniform", 2.25, 0.2), ("pixmap", ts[1], depths[1])]
        Rr, Tr = AG.fluxes(layers, pw, wl, theta=theta, phi=phi)
        Tr.backward()
        assert abs(T[b] - float(Tr.detach())) < 1e-11 and abs(R[b] - float(Rr.detach())) < 1e-11
        for j, nm in enumerate(names):
            ref = ts[j].grad.numpy()
            np.testing.assert_allclose(grads[nm][b].cpu().numpy(), ref, atol=1e-9 * np.abs(ref).max())
    # the reflection objective too
    R2, T2, gR = flux_gradient(cl, wT=0.0, wR=1.0, reduce=True)
    acc = [np.zeros((18, 18)) for _ in pix]
    for wl in wls:
        ts = [torch.tensor(p, requires_grad=True) for p in pix]
        layers = [("pixmap", ts[0], depths[0]), ("uniform", 2.25, 0.2), ("pixmap", ts[1], depths[1])]
        Rr, _ = AG.fluxes(layers, pw, wl, theta=theta, phi=phi)
        Rr.backward()
        for j in range(2):
            acc[j] += ts[j].grad.numpy()
    for j, nm in enumerate(names):
        np.testing.assert_allclose(gR[nm].cpu().numpy(), acc[j], atol=1e-9 * np.abs(acc[j]).max())


def _grad_grid_crystal(findex):
    cfg = W.grad_grid()
    cl, names = _crystal(cfg["pw"], cfg["pixmaps"], [cfg["depth"]] * 4)
    cl.set_source(1 / cfg["freqs"][findex], 1.0, 0.0)
    return cfg, cl, names


def test_grad_grid_vs_cpu_autograd(gpu):
    """pw (11, 11), four 64x64 design layers: dT/d eps at all 16 frequencies
    (one batched adjoint solve) against the FP64 CPU autograd reference, and
    T at the two golden frequencies against the reference's values."""
    from khepri_amd.adjoint import flux_gradient
    findex = np.arange(16)
    cfg, cl, names = _grad_grid_crystal(findex)
    R, T, grads = flux_gradient(cl)
    gold = golden("grad_grid.npz")
    np.testing.assert_allclose(T[gold["findex"]], gold["T"], atol=1e-10)
    np.testing.assert_allclose(R[gold["findex"]], gold["R"], atol=1e-10)
    for b, fi in enumerate(findex):
        Tr, ref = AG.transmission_gradient(cfg["pixmaps"], [cfg["depth"]] * 4, cfg["pw"], 1 / cfg["freqs"][fi])
        assert abs(Tr - T[b]) < 1e-10
        for j, nm in enumerate(names):
            got = grads[nm][b].cpu().numpy()
            err = np.abs(got - ref[j]).max() / np.abs(ref[j]).max()
            print(f"f{fi} {nm}: max rel err vs CPU autograd {err:.2e}", flush=True)
            assert err < 1e-8


def test_grad_grid_vs_finite_differences(gpu):
    """Central differences of the GPU forward on all 64 seeded pixels: with
    the step 1e-6 the difference quotient carries ~1e-14 / 2e-6 of rounding
    noise (about 1e-4 of these gradients, the same on the numpy oracle), so it
    is held to 2e-3; with step 1e-3 truncation and rounding are both below 1e-6."""
    from khepri_amd.adjoint import flux_gradient
    from khepri_amd.batch import CrystalBatch
    cfg, cl, names = _grad_grid_crystal(np.array([0]))
    _, T, grads = flux_gradient(cl)
    pix = W.fd_pixels()
    assert len(pix) == 64
    wl = 1 / cfg["freqs"][0]
    for h, tol in ((1e-6, 2e-3), (1e-3, 1e-6)):
        cls = []
        for (l, x, y) in pix:
            for s in (+1, -1):
                pp = [p.copy() for p in cfg["pixmaps"]]
                pp[l][x, y] += s * h
                c, _ = _crystal(cfg["pw"], pp, [cfg["depth"]] * 4)
                c.set_source(wl, 1.0, 0.0)
                cls.append(c)
        _, Tp = CrystalBatch(cls).solve()
        fd = (Tp[0::2] - Tp[1::2]) / (2 * h)
        ad = np.array([grads[names[l]][0, x, y].item() for (l, x, y) in pix])
        rel = np.abs(fd - ad).max() / np.abs(ad).max()
        print(f"h={h:g}: max |FD - adjoint| / max |adjoint| = {rel:.2e}", flush=True)
        assert rel < tol


def test_flux_function_autograd(gpu):
    from khepri_amd.adjoint import FluxFunction, flux_gradient
    rng = np.random.default_rng(3)
    pix = [rng.uniform(2, 12, size=(16, 16)) for _ in range(2)]
    cl, names = _crystal((3, 3), pix, [0.3, 0.4])
    cl.set_source(1 / np.array([0.52, 0.58]), 1.0, 0.0)
    e = [torch.tensor(p, requires_grad=True) for p in pix]
    R, T = FluxFunction.apply(cl, tuple(names), *e)
    (T.sum() + 0.5 * R[1]).backward()
    _, _, gT = flux_gradient(cl, wT=1.0, wR=np.array([0.0, 0.5]), reduce=True)
    for j, nm in enumerate(names):
        np.testing.assert_allclose(e[j].grad.numpy(), gT[nm].cpu().numpy(), rtol=1e-12, atol=1e-20)

"""khx_eig_noalloc (workspace in a static device array, no allocation) against
khx_eig with a caller workspace, through the C-ABI: two calls in a row, the
second with a smaller batch and order, outputs bit-identical (same kernels,
same inputs)."""
import ctypes

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr())


def _eig_ws(lib, A, gpu):
    b, n, _ = A.shape
    Af = A.clone()
    lam = torch.empty((b, n), dtype=torch.complex128, device=gpu)
    W = torch.empty((b, n, n), dtype=torch.complex128, device=gpu)
    info = torch.full((b,), -7, dtype=torch.int32, device=gpu)
    ws = torch.empty(lib.khx_eig_ws_bytes(n, b), dtype=torch.uint8, device=gpu)
    stream = torch.cuda.current_stream().cuda_stream
    assert lib.khx_eig(n, _ptr(Af), b, _ptr(lam), _ptr(W), _ptr(info), _ptr(ws), stream) == 0
    return lam, W, info


def _eig_noalloc(lib, A, gpu):
    b, n, _ = A.shape
    Af = A.clone()
    lam = torch.empty((b, n), dtype=torch.complex128, device=gpu)
    W = torch.empty((b, n, n), dtype=torch.complex128, device=gpu)
    info = torch.full((b,), -7, dtype=torch.int32, device=gpu)
    stream = torch.cuda.current_stream().cuda_stream
    rc = lib.khx_eig_noalloc(n, _ptr(Af), b, _ptr(lam), _ptr(W), _ptr(info), stream)
    assert rc == 0, lib.khx_last_error()
    return lam, W, info


def test_eig_noalloc_matches_eig_with_workspace(gpu):
    from khepri_amd import _lib
    lib = _lib.load()
    limit = lib.khx_eig_noalloc_bytes()
    assert limit == 64 << 20
    rng = np.random.default_rng(7)
    # first call: 30 matrices of order 162 (the layer eigenproblem of 81
    # harmonics, 31.5 MiB of scratch); second call: 5 of order 50
    for b, n in ((30, 162), (5, 50)):
        assert 2 * n * n * 16 * b <= limit
        A = torch.as_tensor(rng.normal(size=(b, n, n)) + 1j * rng.normal(size=(b, n, n)), device=gpu)
        print(f"[no_alloc] batch {b} order {n}", flush=True)
        lam0, W0, info0 = _eig_ws(lib, A, gpu)
        lam1, W1, info1 = _eig_noalloc(lib, A, gpu)
        torch.cuda.synchronize()
        assert (info0.cpu().numpy() == 0).all() and (info1.cpu().numpy() == 0).all()
        assert torch.equal(lam0, lam1)
        assert torch.equal(W0, W1)
        # and the result is an eigen-decomposition (residual against numpy's eigenvalues)
        An, L, Wn = A.cpu().numpy(), lam1.cpu().numpy(), W1.cpu().numpy()
        res = np.abs(An @ Wn - Wn * L[:, None, :]).max()
        assert res < 1e-10 * np.abs(An).max() * n
    # a batch beyond the static array is refused on the host, before any launch
    n = 162
    too_many = limit // (2 * n * n * 16) + 1
    A = torch.zeros((1, n, n), dtype=torch.complex128, device=gpu)
    lam = torch.empty((1, n), dtype=torch.complex128, device=gpu)
    rc = lib.khx_eig_noalloc(n, _ptr(A), too_many, _ptr(lam), _ptr(A), None,
                             torch.cuda.current_stream().cuda_stream)
    assert rc == -1 and b"static workspace" in lib.khx_last_error()

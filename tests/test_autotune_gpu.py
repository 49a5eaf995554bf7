"""GEMM tile autotuner: every tile shape computes the same product, recording
sees the GEMMs of a solve, and an installed tuning table leaves R/T unchanged."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

OPM = {"N": lambda x: x, "T": lambda x: np.swapaxes(x, -1, -2), "C": lambda x: np.conj(np.swapaxes(x, -1, -2))}


def crand(rng, *shape):
    return rng.normal(size=shape) + 1j * rng.normal(size=shape)


@pytest.mark.parametrize("tile", [0, 1, 2, 3, 4, 5, 6, 7])
@pytest.mark.parametrize("M,N,K,opa,opb", [(162, 162, 162, "N", "N"), (37, 101, 29, "T", "N"),
                                           (70, 33, 65, "N", "C"), (5, 130, 81, "C", "T"),
                                           (64, 64, 32, "N", "N"), (1, 1, 1, "C", "N"), (100, 17, 3, "T", "T")])
def test_every_tile_matches_numpy(gpu, tile, M, N, K, opa, opb):
    from khepri_amd import _lib, ops
    lib = _lib.load()
    rng = np.random.default_rng(M * 3 + N + K + tile)
    batch = 3
    A = crand(rng, batch, *((M, K) if opa == "N" else (K, M)))
    B = crand(rng, batch, *((K, N) if opb == "N" else (N, K)))
    C0 = crand(rng, batch, M, N)
    ref = (0.5 + 2j) * OPM[opa](A) @ OPM[opb](B) - 1j * C0
    assert lib.khx_gemm_force(tile) == 0
    try:
        out = ops.matmul(torch.as_tensor(A, device=gpu), torch.as_tensor(B, device=gpu), alpha=0.5 + 2j,
                         beta=-1j, C=torch.as_tensor(C0, device=gpu).clone(), opa=opa, opb=opb)
        torch.cuda.synchronize()
    finally:
        lib.khx_gemm_force(-1)
    np.testing.assert_allclose(out.cpu().numpy(), ref, atol=1e-12 * np.abs(ref).max())


def _crystal(pw=(5, 5)):
    import khepri_amd as kx
    rng = np.random.default_rng(3)
    cl = kx.Crystal(pw, device="cuda:0")
    cl.add_layer_pixmap("A", rng.uniform(2.0, 10.0, size=(32, 32)), 0.4)
    cl.add_layer_uniform("G", 2.25, 0.2)
    cl.add_layer_pixmap("B", rng.uniform(2.0, 10.0, size=(32, 32)), 0.3)
    cl.set_device(["A", "G", "B"], [True, True, True])
    cl.set_source(np.linspace(1.6, 1.9, 6), 1.0, 0.0, 10.0, 20.0)
    return cl


def test_record_tune_install_keeps_results(gpu, tmp_path):
    from khepri_amd import autotune
    cl = _crystal()

    def run():
        cl.solve()
        return [np.asarray(x) for x in cl.poynting_flux_end()]

    R0, T0 = run()
    shapes = autotune.record(run)
    assert shapes, "a solve launches GEMMs"
    assert all(len(s) == 6 and s[0] > 0 and s[1] > 0 and s[2] > 0 for s in shapes)
    table = autotune.tune(shapes, reps=2)
    assert table and all(e["tile"] in autotune.TILES for e in table.values())
    autotune.install(table)
    try:
        R1, T1 = run()
        path = tmp_path / "tiles.json"
        autotune.save(table, path)
        assert autotune.load(path, install_now=False) == json_roundtrip(table)
    finally:
        autotune.clear(table)
    np.testing.assert_allclose(R1, R0, atol=1e-12)
    np.testing.assert_allclose(T1, T0, atol=1e-12)


def json_roundtrip(t):
    import json
    return json.loads(json.dumps(t))

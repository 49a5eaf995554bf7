"""SolveGraph / GradientGraph (khepri_amd.graph): a Crystal solve and a flux
gradient captured as HIP graphs.  Every replay is checked against the eager
call on the same sources, bit for bit (the graph runs the same kernels on the
same inputs), with eager work of another crystal between the replays; new
sources go in through update_sources; the status words recorded during
capture are checked after each replay."""
import numpy as np
import pytest

from khepri_amd import workloads as W

pytestmark = pytest.mark.gpu


def disc(pw=(5, 5), fields=False):
    from khepri_amd import Crystal
    cfg = W.disc_stack_sweep()
    cl = Crystal(pw)
    cl.add_layer_uniform("S1", cfg["eps_gap"], cfg["depth_gap"])
    cl.add_layer_pixmap("Scyl", cfg["pixmap"], cfg["depth_disc"])
    cl.set_device(["Scyl", "S1", "Scyl"], fields)
    return cl, cfg


def other_work():
    cl, cfg = disc((3, 3))
    cl.set_source(1 / cfg["freqs"][:7], 1.0, 0.3)
    cl.solve()
    return cl.poynting_flux_end()


@pytest.mark.parametrize("rt_only", [None, True])
def test_solve_graph_replays_match_eager(gpu, rt_only):
    from khepri_amd.graph import SolveGraph
    cl, cfg = disc()
    fr = cfg["freqs"]
    cl.set_source(1 / fr, 1.0, 0.0)
    cl.solve(rt_only=rt_only)
    R0, T0 = cl.poynting_flux_end()
    g = SolveGraph(cl, rt_only=rt_only)
    for _ in range(3):
        R, T = g.replay()
        np.testing.assert_array_equal(R, R0)
        np.testing.assert_array_equal(T, T0)
        other_work()
    # new sources of the same batch size
    fr2 = fr * 1.01
    g.update_sources(1 / fr2, 0.3, 1.0, 5.0, 20.0)
    R2, T2 = g.replay()
    ref, _ = disc()
    ref.set_source(1 / fr2, 0.3, 1.0, 5.0, 20.0)
    ref.solve(rt_only=rt_only)
    Rr, Tr = ref.poynting_flux_end()
    np.testing.assert_array_equal(R2, Rr)
    np.testing.assert_array_equal(T2, Tr)
    with pytest.raises(ValueError):
        g.update_sources(1 / fr[:10], 1.0, 0.0)


def test_solve_graph_status_checked_after_replay(gpu, monkeypatch):
    """A graph captured with the eigensolver's iteration cap lowered to 2
    (KHX_EIG_ITMAX, read when the launch is recorded) reports non-convergence
    on replay; the graph itself runs to the end either way."""
    from khepri_amd.graph import SolveGraph
    cl, cfg = disc()
    cl.set_source(1 / cfg["freqs"][:16], 1.0, 0.0)
    g = SolveGraph.__new__(SolveGraph)
    g.__dict__.update(cl=cl, batch=cl.batch, rt_only=None)
    cl.solve()
    monkeypatch.setenv("KHX_EIG_ITMAX", "2")
    g._capture()
    monkeypatch.delenv("KHX_EIG_ITMAX")
    with pytest.raises(np.linalg.LinAlgError):
        g.replay()
    g.replay(check=False)


def test_gradient_graph_replays_match_eager(gpu):
    from khepri_amd import Crystal
    from khepri_amd.adjoint import flux_gradient
    from khepri_amd.graph import GradientGraph
    rng = np.random.default_rng(4)
    pix = [rng.uniform(2, 12, size=(16, 16)) for _ in range(2)]
    cl = Crystal((5, 5))
    cl.add_layer_pixmap("L0", pix[0], 0.3)
    cl.add_layer_uniform("G", 2.25, 0.2)
    cl.add_layer_pixmap("L1", pix[1], 0.4)
    cl.set_device(["L0", "G", "L1"], False)
    wl = 1 / np.linspace(0.5, 0.6, 8)
    cl.set_source(wl, 1.0, 0.0)
    R0, T0, g0 = flux_gradient(cl, wT=1.0, wR=0.5)
    gg = GradientGraph(cl, wT=1.0, wR=0.5)
    for _ in range(2):
        R, T, grads = gg.replay()
        np.testing.assert_array_equal(R, R0)
        np.testing.assert_array_equal(T, T0)
        for nm in ("L0", "L1"):
            np.testing.assert_array_equal(grads[nm].cpu().numpy(), g0[nm].cpu().numpy())
        other_work()
    gg.update_sources(wl * 1.02, 1.0, 0.0)
    R, T, grads = gg.replay()
    cl2 = Crystal((5, 5))
    cl2.add_layer_pixmap("L0", pix[0], 0.3)
    cl2.add_layer_uniform("G", 2.25, 0.2)
    cl2.add_layer_pixmap("L1", pix[1], 0.4)
    cl2.set_device(["L0", "G", "L1"], False)
    cl2.set_source(wl * 1.02, 1.0, 0.0)
    Rr, Tr, gr = flux_gradient(cl2, wT=1.0, wR=0.5)
    np.testing.assert_array_equal(T, Tr)
    np.testing.assert_array_equal(grads["L1"].cpu().numpy(), gr["L1"].cpu().numpy())

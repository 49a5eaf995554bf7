"""Crystal solves and flux gradients replayed from HIP graphs.

The reference solves one (wavelength, k_parallel) at a time from Python
(khepri/crystal.py:180-206, then poynting_flux_end, khepri/crystal.py:363-396);
here a whole batch is one sequence of library launches, and for small batches
(tens to hundreds of solves) the host side of that sequence -- Python, ctypes
and the launches themselves -- is a large part of the step.  The classes below
capture the sequence once (torch.cuda.graph, i.e. hipStreamBeginCapture on the
current stream) and replay it:

    g = SolveGraph(crystal)            # eager warm-up solve, then capture
    R, T = g.replay()                  # same sources
    g.update_sources(wavelength=...)   # new sources of the same batch size
    R, T = g.replay()

    gg = GradientGraph(crystal, wT=1.0)   # flux_gradient (khepri_amd.adjoint)
    R, T, grads = gg.replay()

What is captured: the layer solves, the S-matrix chain and the R/T reduction
(and, for GradientGraph, the adjoint sweep down to the pixel gradients),
reading the sources from device buffers owned by the graph.  What stays
outside: uploads of new sources (update_sources copies them into those
buffers before a replay) and the status words of the eigensolver and the LU
factorizations, which the captured sequence records instead of reading;
replay() checks them after the graph has run and raises numpy's LinAlgError
as the eager call does.  Workspaces come from the graph's private memory pool
and stay allocated while the graph object lives.  A replay overwrites the
outputs of the previous one (replay() returns host copies of R and T; the
gradient tensors are the graph's own buffers).
"""
from __future__ import annotations

import torch

from . import ops


class _Captured:
    def __init__(self, crystal):
        self.cl = crystal
        self.batch = crystal.batch

    def _capture(self):
        self.sources = tuple(t.clone() for t in self.cl._sources_dev)
        side = torch.cuda.Stream()
        side.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(side):      # one warm-up on a side stream, as torch.cuda.graph expects
            with ops.deferred_status([]):
                self._run()
        torch.cuda.current_stream().wait_stream(side)
        self.graph = torch.cuda.CUDAGraph()
        self.status = []
        with torch.cuda.graph(self.graph):
            with ops.deferred_status(self.status):
                self.outputs = self._run()

    def _run(self):
        raise NotImplementedError

    def update_sources(self, wavelength, te=1.0, tm=1.0, theta=0.0, phi=0.0, kp=None):
        """New sources for the next replays (Crystal.set_source's arguments);
        the batch size must stay the captured one."""
        cl = self.cl
        cl.set_source(wavelength, te, tm, theta, phi, kp=kp)
        if cl.batch != self.batch:
            raise ValueError(f"the graph was captured for {self.batch} sources, got {cl.batch}")
        for dst, src in zip(self.sources, (cl._kp, cl._wl, cl._te, cl._tm)):
            dst.copy_(torch.as_tensor(src, device=dst.device))

    def _replay(self, check):
        self.graph.replay()
        if check:
            ops.check_status(self.status)


class SolveGraph(_Captured):
    """Crystal.solve + poynting_flux_end as one graph."""

    def __init__(self, crystal, rt_only=None):
        super().__init__(crystal)
        self.rt_only = rt_only
        crystal.solve(rt_only=rt_only)       # eager: caches convolution matrices, tunes GEMM tiles
        crystal.poynting_flux_end()
        self._capture()

    def _run(self):
        cl = self.cl
        cl._static_sources = self.sources
        try:
            cl.solve(rt_only=self.rt_only)
            if cl._RT is not None:
                return cl._RT[0], cl._RT[1]
            kp, wl, te, tm = cl._sources_dev
            R, T, _, _ = ops.poynting(cl.Stot, cl.expansion.g_device(cl._dev()), kp, wl, te, tm,
                                      cl.epsi, cl.epse, orders=False)
            return R, T
        finally:
            cl._static_sources = None

    def replay(self, check=True):
        """Run the captured solve; (R, T) as numpy arrays."""
        self._replay(check)
        R, T = (x.cpu().numpy() for x in self.outputs)
        if self.cl._scalar:
            return R[0], T[0]
        return R, T


class GradientGraph(_Captured):
    """khepri_amd.adjoint.flux_gradient as one graph (weights fixed at capture)."""

    def __init__(self, crystal, wT=1.0, wR=0.0, wrt=None, reduce=False, eta_rel=1e-12):
        from .adjoint import _weights, flux_gradient
        super().__init__(crystal)
        dev = crystal._dev()
        self.kw = dict(wrt=wrt, reduce=reduce, eta_rel=eta_rel)
        flux_gradient(crystal, wT=wT, wR=wR, **self.kw)     # eager warm-up
        self.wT = _weights(wT, crystal.batch, dev)
        self.wR = _weights(wR, crystal.batch, dev)
        crystal.solve(rt_only=True)                          # device sources for the capture buffers
        self._capture()

    def _run(self):
        from .adjoint import _flux_gradient_device
        cl = self.cl
        cl._static_sources = self.sources
        try:
            return _flux_gradient_device(cl, wT=self.wT, wR=self.wR, **self.kw)
        finally:
            cl._static_sources = None

    def replay(self, check=True):
        """Run the captured gradient: (R, T) numpy, grads {layer: device tensor}."""
        self._replay(check)
        R, T, grads = self.outputs
        return R.cpu().numpy(), T.cpu().numpy(), grads

"""Permittivity pixmap rasteriser with the reference's pixel conventions.

Mirrors ``khepri.draw.Drawing`` (khepri/draw.py:21-55, 90-95): pixel centres on
``linspace(-0.5, 0.5, n)`` (both endpoints included), ``meshgrid(indexing="ij")``
so the array is indexed ``eps[x, y]``, discs use a strict ``<`` and rectangles
closed edges.  The canvas is ``ones(shape) * background``, so it is real unless
the background permittivity is complex (a complex island on a real canvas
loses its imaginary part, exactly as in the reference).
"""
from __future__ import annotations

import numpy as np


class Drawing:
    def __init__(self, shape, epsilon_background, lattice=None) -> None:
        if lattice is None:
            lattice = np.array([[1, 0], [0, 1]])
        self.lattice = np.asarray(lattice)
        self.background = epsilon_background
        self._canvas = np.ones(shape) * epsilon_background
        xbar = np.linspace(-0.5, 0.5, shape[0])
        ybar = np.linspace(-0.5, 0.5, shape[1])
        nX, nY = np.meshgrid(xbar, ybar, indexing="ij")
        self.X = self.lattice[0, 0] * nX + self.lattice[1, 0] * nY
        self.Y = self.lattice[0, 1] * nX + self.lattice[1, 1] * nY
        self.geometric_description = []

    def disc(self, xy, radius, epsilon):
        x, y = xy
        mask = np.sqrt((self.X - x) ** 2 + (self.Y - y) ** 2) < radius
        self._canvas[mask] = epsilon
        self.geometric_description.append(
            {"type": "disc", "params": [0.5 + x, 0.5 + y, radius], "epsilon": epsilon})

    def rectangle(self, xy, wh, epsilon):
        x, y = xy
        w, h = wh
        x = x - w / 2
        y = y - h / 2
        mask = (self.X >= x) & (self.X <= x + w) & (self.Y >= y) & (self.Y <= y + h)
        self._canvas[mask] = epsilon
        self.geometric_description.append(
            {"type": "rectangle", "params": [0.5 + x, 0.5 + y, 0.5 + x + w, 0.5 + y + h],
             "epsilon": epsilon})

    def canvas(self):
        return self._canvas.copy()

    def islands(self):
        return self.geometric_description

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
    """A unit cell painted shape by shape; ``canvas()`` is the permittivity
    map, ``islands()`` the same shapes as closed-form island descriptions
    (corner coordinates shifted to the cell's [0, 1] frame)."""

    def __init__(self, shape, epsilon_background, lattice=None) -> None:
        self.lattice = np.eye(2, dtype=int) if lattice is None else np.asarray(lattice)
        self.background = epsilon_background
        self._canvas = epsilon_background * np.ones(shape)
        # pixel centres in lattice coordinates (u along the first vector, v
        # along the second), then in Cartesian coordinates
        u, v = np.meshgrid(np.linspace(-0.5, 0.5, shape[0]), np.linspace(-0.5, 0.5, shape[1]), indexing="ij")
        (a1x, a1y), (a2x, a2y) = self.lattice[0], self.lattice[1]
        self.X = a1x * u + a2x * v
        self.Y = a1y * u + a2y * v
        self.geometric_description = []

    def _paint(self, inside, epsilon, kind, params):
        self._canvas[inside] = epsilon
        self.geometric_description.append({"type": kind, "params": params, "epsilon": epsilon})

    def disc(self, xy, radius, epsilon):
        """Pixels strictly closer than radius to the centre xy."""
        cx, cy = xy
        inside = np.sqrt((self.X - cx) ** 2 + (self.Y - cy) ** 2) < radius
        self._paint(inside, epsilon, "disc", [0.5 + cx, 0.5 + cy, radius])

    def rectangle(self, xy, wh, epsilon):
        """Pixels inside the closed w x h rectangle centred on xy."""
        w, h = wh
        left, bottom = xy[0] - w / 2, xy[1] - h / 2
        right, top = left + w, bottom + h
        inside = (self.X >= left) & (self.X <= right) & (self.Y >= bottom) & (self.Y <= top)
        self._paint(inside, epsilon, "rectangle", [0.5 + left, 0.5 + bottom, 0.5 + right, 0.5 + top])

    def canvas(self):
        return self._canvas.copy()

    def islands(self):
        return self.geometric_description

"""Seeded synthetic workloads of the RCWA solve path.

Five workloads, named by what they are:

* ``disc_stack_sweep``  -- pw (5,5): 128x128 pixmap (eps 12, centred disc r=0.4
  of eps 1) depth 0.55, uniform eps 1 depth 1.1, stack [disc, gap, disc] (the
  same pixmap layer name used twice); 151 frequencies linspace(0.49, 0.6);
  normal incidence, te=1, tm=0.
* ``disc_stack_bz``     -- same stack at pw (9,9); 64 frequencies
  linspace(0.45, 0.65) x a 32x32 cell-centred Bloch grid (65 536 solves).
* ``random_devices``    -- 4096 seeded devices, pw in {5,7,11,13}^2, 1-8 layers of
  128x128 pixmaps of 1-4 discs/rectangles with complex eps.
* ``grad_grid``         -- pw (11,11), 4 layers of 64x64 real eps ~ U[2,12],
  depth 0.3, 16 frequencies linspace(0.5, 0.6), normal incidence.
* ``twisted_bilayer``   -- extended basis (5,5)x(5,5): two disc layers (r=0.3,
  eps 4 in eps 1, depth 0.4) around an eps-1 gap of 0.2, 8 seeded twist
  angles in [2, 30] degrees, 32 frequencies linspace(0.4, 0.7).

Only numpy is used here: these are plain input generators.
"""
from __future__ import annotations

import numpy as np

from .draw import Drawing


# --------------------------------------------------------------------------
def disc_pixmap(n=128, bg=12.0, radius=0.4, eps=1.0, drawing=Drawing):
    d = drawing((n, n), bg)
    d.disc((0.0, 0.0), radius, eps)
    return d.canvas()


def disc_stack_sweep():
    return dict(pw=(5, 5), pixmap=disc_pixmap(), depth_disc=0.55, eps_gap=1.0, depth_gap=1.1,
                freqs=np.linspace(0.49, 0.6, 151))


def bz_grid(shape=(32, 32), a=1.0):
    """Cell-centred Bloch grid (khepri/beams.py:193-207), shape (2, *shape)."""
    b1, b2 = [2 * np.pi / a, 0], [0, 2 * np.pi / a]
    si, sj = 1 / shape[0], 1 / shape[1]
    i, j = np.meshgrid(np.arange(-0.5 + si / 2, 0.5, si),
                       np.arange(-0.5 + sj / 2, 0.5, sj), indexing="ij")
    return np.stack([b1[0] * i + b2[0] * j, b1[1] * i + b2[1] * j])


def disc_stack_bz(nf=64, nk=32):
    freqs = np.linspace(0.45, 0.65, nf)
    kgrid = bz_grid((nk, nk)).reshape(2, -1).T          # (nk*nk, 2)
    # solve index s = f_index * nk^2 + k_index
    f = np.repeat(freqs, kgrid.shape[0])
    kp = np.tile(kgrid, (nf, 1))
    return dict(pw=(9, 9), pixmap=disc_pixmap(), depth_disc=0.55, eps_gap=1.0, depth_gap=1.1,
                freqs=f, kp=kp)


# --------------------------------------------------------------------------
RANDOM_PW = [(5, 5), (7, 7), (11, 11), (13, 13)]


def random_device_params(count=4096, seed=3):
    """Parameter draws of the random-device workload (pixmaps rendered lazily).

    Builder's choices for ranges the workload leaves open: background and
    island eps = U[1,13] + i U[0,0.1]; disc radius U[0.05,0.3]; rectangle
    sides U[0.05,0.5]; centres U[-0.5,0.5]^2; te = tm = 1.
    """
    rng = np.random.default_rng(seed)
    devs = []
    for _ in range(count):
        pw = RANDOM_PW[int(rng.integers(0, 4))]
        nl = int(rng.integers(1, 9))
        layers = []
        for _ in range(nl):
            bg = complex(rng.uniform(1, 13), rng.uniform(0, 0.1))
            shapes = []
            for _ in range(int(rng.integers(1, 5))):
                kind = int(rng.integers(0, 2))
                cx, cy = rng.uniform(-0.5, 0.5, 2)
                if kind == 0:
                    size = (float(rng.uniform(0.05, 0.3)),)
                else:
                    size = tuple(float(v) for v in rng.uniform(0.05, 0.5, 2))
                eps = complex(rng.uniform(1, 13), rng.uniform(0, 0.1))
                shapes.append((kind, (float(cx), float(cy)), size, eps))
            layers.append(dict(bg=bg, shapes=shapes, depth=float(rng.uniform(0.1, 1.0))))
        devs.append(dict(pw=pw, layers=layers, freq=float(rng.uniform(0.3, 0.8)),
                         theta=float(rng.uniform(0, 40)), phi=float(rng.uniform(0, 360))))
    return devs


def render_layer(layer, n=128, drawing=Drawing):
    d = drawing((n, n), layer["bg"])
    for kind, c, size, eps in layer["shapes"]:
        if kind == 0:
            d.disc(c, size[0], eps)
        else:
            d.rectangle(c, size, eps)
    return d.canvas()


# --------------------------------------------------------------------------
def grad_grid(seed=4):
    rng = np.random.default_rng(seed)
    pixmaps = [rng.uniform(2, 12, (64, 64)) for _ in range(4)]
    return dict(pw=(11, 11), pixmaps=pixmaps, depth=0.3, freqs=np.linspace(0.5, 0.6, 16))


def fd_pixels(seed=44, count=64):
    """(layer, x, y) pixel triples for finite-difference checks."""
    rng = np.random.default_rng(seed)
    return np.stack([rng.integers(0, 4, count), rng.integers(0, 64, count),
                     rng.integers(0, 64, count)], axis=1)


# --------------------------------------------------------------------------
def twisted_bilayer(seed=5):
    rng = np.random.default_rng(seed)
    angles_deg = rng.uniform(2, 30, 8)
    return dict(pw=(5, 5), pixmap=disc_pixmap(128, 1.0, 0.3, 4.0), depth=0.4, gap=0.2,
                angles_deg=angles_deg, freqs=np.linspace(0.4, 0.7, 32))

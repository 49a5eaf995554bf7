"""Field reconstruction (khepri/crystal.py:208-343, khepri/fields.py)."""
from __future__ import annotations

import numpy as np


def locate_layer(self, z):
    """Layer containing depth z and the depth relative to it (khepri/crystal.py:208-232)."""
    assert z != np.nan
    layer_index = int(np.searchsorted(self.stack_positions, z)) - 1
    layer_name = self.global_stacking[layer_index]
    zr = z if z <= 0 else z - self.stack_positions[layer_index]
    return self.layers[layer_name], layer_index, zr


def get_source_as_field_vectors(self):
    raise NotImplementedError


def fourier_fields(self, z, incident_fields):
    raise NotImplementedError


def fields_coords_xy(self, x, y, z, incident_fields=None, kp=None, return_fourier=False):
    raise NotImplementedError


def fields_volume(self, x, y, z, incident_fields=None):
    raise NotImplementedError

"""Saving and loading solver state (khepri/__init__.py:12-19).

The reference pickles any object (a Crystal, a list of layers, results) to a
file.  ``save`` / ``load`` keep that contract and the pickle format, with one
change for device data: every torch tensor in the object graph is written as
a host numpy array and comes back as a tensor on the current HIP device
(CPU tensors stay on the CPU).  A file therefore does not depend on the GPU
it was written on, and a solved Crystal can be reloaded and queried
(poynting_flux_end, fields) or solved again.  As with the reference, load
only files you trust: unpickling runs code named in the file.
"""
from __future__ import annotations

import pickle

import numpy as np
import torch


def _restore_tensor(array, on_device, requires_grad=False):
    t = torch.from_numpy(np.array(array, copy=True))
    if on_device:
        from .device import default_device
        t = t.to(default_device())
    return t.requires_grad_(requires_grad) if requires_grad else t


class _Pickler(pickle.Pickler):
    def reducer_override(self, obj):
        if isinstance(obj, torch.Tensor):
            arr = obj.detach().cpu().resolve_conj().resolve_neg().numpy()
            return _restore_tensor, (arr, obj.device.type != "cpu", bool(obj.requires_grad))
        return NotImplemented


def save(path, obj):
    """Pickle obj to path (tensors as host arrays)."""
    with open(path, "wb") as f:
        _Pickler(f, protocol=pickle.HIGHEST_PROTOCOL).dump(obj)


def load(path):
    """Unpickle an object written by save (tensors back on the device)."""
    with open(path, "rb") as f:
        return pickle.load(f)

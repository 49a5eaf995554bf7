"""Device selection: the HIP device the RCWA path runs on."""
from __future__ import annotations

import os

import torch

_default = None


def default_device():
    """cuda:<KHX_DEVICE or 0>; raises when no HIP device is visible (there is
    no CPU execution path)."""
    global _default
    if _default is None:
        if not torch.cuda.is_available():
            raise RuntimeError("khepri_amd needs a HIP device (MI355X); none is visible")
        _default = torch.device("cuda", int(os.environ.get("KHX_DEVICE", "0")))
    return _default


def set_default_device(dev):
    global _default
    _default = torch.device(dev)

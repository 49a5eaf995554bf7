"""Autotuner for the tile shapes of the batched complex GEMM.

libkhx's zgemm comes in eight kernels of 256 threads: FP64 FMA register
tiles (M x N per workgroup) 0 = 64x64, 1 = 32x32, 2 = 64x32, 3 = 32x64, and
the same tiles on the FP64 matrix cores (v_mfma_f64_16x16x4f64) 4 = 64x64,
5 = 32x32, 6 = 64x32, 7 = 32x64.  Without a tuning table the library takes
MFMA 32x32 when either side is at most 96 and MFMA 32x64 otherwise
(``default_tile``).  The tuner

1. runs a workload once with shape recording on (``record``), which collects
   every distinct (M, N, K, opA, opB, batch) the library launches;
2. times each tile shape on every recorded GEMM class with random operands of
   that shape on the current device (``tune``), the batch capped so that one
   operand stays under ``max_bytes``;
3. installs the fastest choice per class in the library (``install``) and can
   save / load the table as JSON so that a later process skips the timing.

The choice is keyed by (M, N, K, opA, opB) and the batch bucket (< 16, < 256,
>= 256), which is also how the library looks it up at launch time.

    from khepri_amd import autotune
    table = autotune.autotune(lambda: crystal.solve())   # record + time + install
    autotune.save(table, "gemm_tiles.json")
    ...
    autotune.load("gemm_tiles.json")                      # in a later process
"""
from __future__ import annotations

import json

import numpy as np
import torch

from . import _lib
from ._lib import KhxError, check

TILES = {0: "64x64", 1: "32x32", 2: "64x32", 3: "32x64",
         4: "mfma64x64", 5: "mfma32x32", 6: "mfma64x32", 7: "mfma32x64"}
_OPNAME = {0: "N", 1: "T", 2: "C"}


def default_tile(M, N):
    """The library's choice for an (M x N) product when no table entry applies."""
    return 5 if min(M, N) <= 96 else 7


def _bucket(batch):
    return 2 if batch >= 256 else (1 if batch >= 16 else 0)


def record(fn, *args, max_shapes=4096, **kwargs):
    """Run fn(*args, **kwargs) with GEMM shape recording on; return the
    recorded shapes as (M, N, K, opA, opB, batch) tuples."""
    lib = _lib.load()
    rc = lib.khx_gemm_record(1, None, 0)
    if rc < 0:
        check(rc, "khx_gemm_record")
    buf = np.zeros((max_shapes, 6), dtype=np.int32)
    try:
        fn(*args, **kwargs)
        torch.cuda.synchronize()
    finally:
        n = lib.khx_gemm_record(0, buf.ctypes.data, max_shapes)
    if n < 0:
        check(n, "khx_gemm_record")
    return [tuple(int(v) for v in r) for r in buf[:n]]


def _operands(M, N, K, oa, ob, batch, device, gen):
    ashape = (batch, M, K) if oa == 0 else (batch, K, M)
    bshape = (batch, K, N) if ob == 0 else (batch, N, K)

    def rnd(shape):
        return torch.complex(torch.randn(shape, dtype=torch.float64, device=device, generator=gen),
                             torch.randn(shape, dtype=torch.float64, device=device, generator=gen))

    return rnd(ashape), rnd(bshape), torch.zeros((batch, M, N), dtype=torch.complex128, device=device)


def time_tiles(shape, reps=5, max_bytes=1 << 30, device="cuda"):
    """Average milliseconds per launch of each tile shape on one GEMM class."""
    from . import ops
    lib = _lib.load()
    M, N, K, oa, ob, batch = shape
    largest = 16 * max(M * K, K * N, M * N)
    bt = max(1, min(batch, max_bytes // largest))
    gen = torch.Generator(device=device)
    gen.manual_seed(1234)
    A, B, C = _operands(M, N, K, oa, ob, bt, device, gen)
    out = {}
    stream = torch.cuda.current_stream()
    try:
        for cfg in TILES:
            check(lib.khx_gemm_force(cfg), "khx_gemm_force")
            ops.matmul(A, B, C=C, opa=_OPNAME[oa], opb=_OPNAME[ob])
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record(stream)
            for _ in range(reps):
                ops.matmul(A, B, C=C, opa=_OPNAME[oa], opb=_OPNAME[ob])
            e1.record(stream)
            e1.synchronize()
            out[cfg] = e0.elapsed_time(e1) / reps
    finally:
        check(lib.khx_gemm_force(-1), "khx_gemm_force")
    return out, bt


def tune(shapes, reps=5, max_bytes=1 << 30, device="cuda"):
    """Time every tile shape on each GEMM class in ``shapes``; return the table
    {key: {"shape": [...], "tile": best, "ms": {tile: ms}, "timed_batch": b}}
    with one entry per (M, N, K, opA, opB, batch bucket)."""
    table = {}
    for sh in shapes:
        M, N, K, oa, ob, batch = sh
        key = f"{M}x{N}x{K}:{_OPNAME[oa]}{_OPNAME[ob]}:b{_bucket(batch)}"
        if key in table and table[key]["shape"][5] >= batch:
            continue
        ms, bt = time_tiles(sh, reps=reps, max_bytes=max_bytes, device=device)
        best = min(ms, key=ms.get)
        table[key] = {"shape": list(sh), "tile": int(best), "ms": {TILES[c]: v for c, v in ms.items()},
                      "timed_batch": int(bt)}
    return table


def install(table):
    """Load a tuning table into the library (replaces the choices it names)."""
    lib = _lib.load()
    for ent in table.values():
        M, N, K, oa, ob, batch = ent["shape"]
        check(lib.khx_gemm_tune_set(M, N, K, oa, ob, batch, int(ent["tile"])), "khx_gemm_tune_set")


def clear(table):
    """Remove the choices a table installed (the built-in rule applies again)."""
    lib = _lib.load()
    for ent in table.values():
        M, N, K, oa, ob, batch = ent["shape"]
        check(lib.khx_gemm_tune_set(M, N, K, oa, ob, batch, -1), "khx_gemm_tune_set")


def autotune(fn, *args, reps=5, max_bytes=1 << 30, **kwargs):
    """record(fn) + tune + install; returns the table."""
    table = tune(record(fn, *args, **kwargs), reps=reps, max_bytes=max_bytes)
    install(table)
    return table


def save(table, path):
    meta = {"device": torch.cuda.get_device_name(0) if torch.cuda.is_available() else None, "tiles": TILES}
    with open(path, "w") as f:
        json.dump({"meta": meta, "table": table}, f, indent=1, sort_keys=True)


def load(path, install_now=True):
    with open(path) as f:
        data = json.load(f)
    table = data["table"]
    for ent in table.values():
        if len(ent["shape"]) != 6 or int(ent["tile"]) not in TILES:
            raise KhxError(f"{path}: malformed tuning entry {ent}")
    if install_now:
        install(table)
    return table

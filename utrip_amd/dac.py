"""DAC files -- the CANS binary output format (cansio/dacdef*.f,
dacopnr*.f, read by idl/dacget*.pro).

A DAC file is a Fortran unformatted-sequential file (each record framed by
4-byte little-endian length markers) whose first five records are

    1            (endian probe, INTEGER*4)
    0            (version)
    mtype        (4 = INTEGER*4, 5 = REAL*4, 6 = REAL*8)
    ndim         (number of dimension entries that follow)
    dims...      (one record; a trailing -1 marks an unlimited time axis)

followed by one record per snapshot (the whole array, column-major).  The
static "d" files (dacdef1d ...) have no time axis; the time-series "s" files
(dacdef1s ...) end the dimension list with -1.
"""
from __future__ import annotations

import struct
from pathlib import Path
from typing import List, Sequence

import numpy as np

_DT = {4: np.int32, 5: np.float32, 6: np.float64}
_MT = {np.dtype(np.int32): 4, np.dtype(np.float32): 5, np.dtype(np.float64): 6}


def _record(f, payload: bytes) -> None:
    n = struct.pack("<i", len(payload))
    f.write(n + payload + n)


def _read_record(f) -> bytes:
    head = f.read(4)
    if len(head) < 4:
        raise EOFError
    (n,) = struct.unpack("<i", head)
    data = f.read(n)
    tail = f.read(4)
    if len(data) != n or tail != head:
        raise ValueError("corrupt DAC record")
    return data


def write(path, snapshots: Sequence[np.ndarray], series: bool = True) -> None:
    """Write a DAC file: `series` = a time-series ("s") file with one record
    per snapshot, otherwise a static ("d") file holding one array."""
    snaps = [np.asarray(s) for s in snapshots]
    if not snaps:
        raise ValueError("no data")
    dt = snaps[0].dtype
    if dt not in _MT:
        raise TypeError(f"DAC supports int32/float32/float64, not {dt}")
    dims = list(snaps[0].shape) if snaps[0].ndim else []
    if series:
        dims = dims + [-1]
    with open(path, "wb") as f:
        _record(f, struct.pack("<i", 1))
        _record(f, struct.pack("<i", 0))
        _record(f, struct.pack("<i", _MT[dt]))
        _record(f, struct.pack("<i", len(dims)))
        _record(f, struct.pack("<%di" % len(dims), *dims))
        for s in snaps:
            if s.dtype != dt or list(s.shape) != list(snaps[0].shape):
                raise ValueError("snapshots must share dtype and shape")
            _record(f, np.asfortranarray(s).tobytes(order="F"))


def read(path) -> tuple:
    """(snapshots, header) of a DAC file; arrays are Fortran-ordered."""
    with open(path, "rb") as f:
        (probe,) = struct.unpack("<i", _read_record(f))
        if probe != 1:
            raise ValueError("not a little-endian DAC file")
        (ver,) = struct.unpack("<i", _read_record(f))
        (mtype,) = struct.unpack("<i", _read_record(f))
        (ndim,) = struct.unpack("<i", _read_record(f))
        raw = _read_record(f)
        dims = list(struct.unpack("<%di" % ndim, raw)) if ndim else []
        series = bool(dims) and dims[-1] == -1
        shape = tuple(dims[:-1] if series else dims)
        dtype = _DT[mtype]
        snaps: List[np.ndarray] = []
        while True:
            try:
                data = _read_record(f)
            except EOFError:
                break
            arr = np.frombuffer(data, dtype=dtype)
            snaps.append(arr.reshape(shape, order="F").copy(order="F"))
    return snaps, dict(version=ver, mtype=mtype, dims=dims, series=series)


def write_state(directory, fields: dict, grid: dict | None = None) -> None:
    """One time-series file per field (<name>.dac) and one static file per
    grid array, the layout a CANS main program writes and reads back with
    mcont=1 (cans1d/md_mhdshktb/hdmlw/main.f:77-102)."""
    d = Path(directory)
    d.mkdir(parents=True, exist_ok=True)
    for k, v in fields.items():
        write(d / f"{k}.dac", [v], series=True)
    for k, v in (grid or {}).items():
        write(d / f"{k}.dac", [v], series=False)


def read_last(path) -> np.ndarray:
    """Last snapshot of a DAC file (what a restart keeps)."""
    snaps, _ = read(path)
    if not snaps:
        raise ValueError(f"{path}: no snapshot")
    return snaps[-1]

"""HomelTab collapse-table file IO (host side).

Format of read_homeltab / write_homeltab (Huntian.jl src/EllipsoidalCollapse/CollapseTable.jl:78-115):
a 36-byte header of 3 Int32 (Nx, Ny, Nz) and 6 Float32 (X1, X2, Y1, Y2, Z1, Z2), then the
Float32 table in column-major order.  The axis bounds are widened to Float64 on read.  The
table value is 1 + z_vir on the regular grid (log10 F, e, p/e); out-of-box lookups give -1.
"""
from __future__ import annotations

import os

import numpy as np

DATA_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data")
WEBSKY_TABLE = os.path.join(DATA_DIR, "HomelTab_websky.dat")


def read_homeltab(path: str) -> dict:
    with open(path, "rb") as fh:
        raw = fh.read()
    nx, ny, nz = (int(v) for v in np.frombuffer(raw, "<i4", 3, 0))
    bounds = np.frombuffer(raw, "<f4", 6, 12).astype(np.float64)
    n = nx * ny * nz
    if len(raw) != 36 + 4 * n:
        raise ValueError(f"{path}: expected {36 + 4 * n} bytes, found {len(raw)}")
    table = np.frombuffer(raw, "<f4", n, 36).reshape((nx, ny, nz), order="F").copy(order="F")
    x1, x2, y1, y2, z1, z2 = (float(v) for v in bounds)
    return dict(nx=nx, ny=ny, nz=nz, x1=x1, x2=x2, y1=y1, y2=y2, z1=z1, z2=z2,
                out_val=-1.0, table=table)


def write_homeltab(path: str, ct: dict) -> None:
    hdr = np.array([ct["nx"], ct["ny"], ct["nz"]], "<i4").tobytes()
    hdr += np.array([ct[k] for k in ("x1", "x2", "y1", "y2", "z1", "z2")], "<f4").tobytes()
    body = np.asarray(ct["table"], np.float32).ravel(order="F").astype("<f4").tobytes()
    with open(path, "wb") as fh:
        fh.write(hdr + body)


def make_ctab(table, x1, x2, y1, y2, z1, z2, out_val=-1.0) -> dict:
    """Collapse-table dict from an in-memory (nx, ny, nz) array (as CollapseTableParams)."""
    t = np.asfortranarray(np.asarray(table, np.float32))
    nx, ny, nz = t.shape
    return dict(nx=nx, ny=ny, nz=nz, x1=float(x1), x2=float(x2), y1=float(y1), y2=float(y2),
                z1=float(z1), z2=float(z2), out_val=float(out_val), table=t)

"""Halo catalogs in the reference's .pksc layout (src/IO/Catalog.jl:4-140).

Header: Int32 Nhalo, Float32 RTHLmax, Float32 z_out; then Nhalo records of 11 Float32
(HaloRecord: x y z, vx vy vz, RTHL, vx2 vy2 vz2, overdensity) or 33 Float32 (ExtHaloRecord:
those, then e_v p_v, strain 11 22 33 23 13 12, d2F, zform, grad xyz, gradf xyz, Rf,
FcollvRf, d2FRf, gradrf xyz).  Little-endian, no padding (Fortran STREAM)."""
from __future__ import annotations

import numpy as np

BASIC_FIELDS = ("x", "y", "z", "vx", "vy", "vz", "RTHL", "vx2", "vy2", "vz2", "overdensity")
EXT_FIELDS = BASIC_FIELDS + (
    "e_v", "p_v", "strain_11", "strain_22", "strain_33", "strain_23", "strain_13",
    "strain_12", "d2F", "zform", "grad_x", "grad_y", "grad_z", "gradf_x", "gradf_y",
    "gradf_z", "Rf", "FcollvRf", "d2FRf", "gradrf_x", "gradrf_y", "gradrf_z")


def write_pksc(path: str, records, RTHLmax: float, z_out: float) -> None:
    """write_pksc (Catalog.jl:53-93): records is an (N, 11) or (N, 33) float32 array."""
    rec = np.ascontiguousarray(np.asarray(records, dtype="<f4"))
    if rec.ndim != 2 or rec.shape[1] not in (11, 33):
        raise ValueError(f"records must be (N, 11) or (N, 33), got {rec.shape}")
    with open(path, "wb") as fh:
        np.array([rec.shape[0]], dtype="<i4").tofile(fh)
        np.array([RTHLmax, z_out], dtype="<f4").tofile(fh)
        rec.tofile(fh)


def read_pksc(path: str):
    """read_pksc (Catalog.jl:100-120): (records (N, 11 or 33) float32, RTHLmax, z_out); the
    record size follows from the file size, as in the reference."""
    with open(path, "rb") as fh:
        nhalo = int(np.fromfile(fh, dtype="<i4", count=1)[0])
        RTHLmax, z_out = (float(v) for v in np.fromfile(fh, dtype="<f4", count=2))
        data = np.fromfile(fh, dtype="<f4")
    nfields = data.size // nhalo if nhalo > 0 else 11
    return data[:nhalo * nfields].reshape(nhalo, nfields), RTHLmax, z_out


def as_table(records) -> dict:
    """Records as a dict of named columns (HaloRecord / ExtHaloRecord field names)."""
    rec = np.asarray(records)
    names = EXT_FIELDS if rec.shape[1] == 33 else BASIC_FIELDS
    return {k: rec[:, q] for q, k in enumerate(names)}

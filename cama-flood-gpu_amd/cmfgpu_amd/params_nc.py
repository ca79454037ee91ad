"""Parameter files in the reference's parameters.nc schema (NetCDF classic).

The reference writes its map parameters with netCDF4 (HDF5 storage,
cmfgpu/params/merit_map.py:1085-1240): dimensions ``catchment``,
``flood_level``, ``bifurcation_path``, ``bifurcation_level``, ``levee``,
``reservoir`` (cmfgpu/params/schema.py:14-22), one variable per parameter
field under the field's own name, connectivity as int64 ids
(``downstream_id`` etc., a mouth pointing at itself, merit_map.py:449-452).

Neither netCDF4 nor h5py is part of this stack, so this module reads and
writes the NetCDF *classic* encodings instead: CDF-1, CDF-2 (64-bit offsets)
and CDF-5 (64-bit data, which keeps the int64 ids).  A reference file
converts without loss with ``nccopy -k cdf5 parameters.nc parameters_cdf5.nc``.
The reader derives the ``*_idx`` index fields the model takes from the ids
(the reference's ReferenceIndexField, cmfgpu/modules/base.py:323).

Format: the published NetCDF classic format specification (big-endian header
of dimensions, global attributes and variables, each non-record variable's
data contiguous at its ``begin`` offset, padded to 4 bytes).
"""
from __future__ import annotations

import struct
from pathlib import Path

import numpy as np

# nc_type codes and their big-endian numpy dtypes
_TYPES = {1: ">i1", 2: "S1", 3: ">i2", 4: ">i4", 5: ">f4", 6: ">f8", 7: ">u1", 8: ">u2", 9: ">u4",
          10: ">i8", 11: ">u8"}
_CODES = {np.dtype(np.int8): 1, np.dtype(np.int16): 3, np.dtype(np.int32): 4, np.dtype(np.float32): 5,
          np.dtype(np.float64): 6, np.dtype(np.uint8): 7, np.dtype(np.uint16): 8, np.dtype(np.uint32): 9,
          np.dtype(np.int64): 10, np.dtype(np.uint64): 11}
_NC_DIMENSION, _NC_VARIABLE, _NC_ATTRIBUTE = 10, 11, 12

# id field -> the index field derived from it (ids looked up in catchment_id)
ID_FIELDS = {
    "downstream_id": "downstream_idx",
    "bifurcation_catchment_id": "bifurcation_catchment_idx",
    "bifurcation_downstream_id": "bifurcation_downstream_idx",
    "levee_catchment_id": "levee_catchment_idx",
    "reservoir_catchment_id": "reservoir_catchment_idx",
}
_RESERVOIR_FIELDS = ("conservation_volume", "emergency_volume", "adjustment_volume", "normal_outflow",
                     "effective_normal_outflow", "adjustment_outflow", "flood_control_outflow")

# Dimensions of the schema's fields by name: the map metadata
# (cmfgpu/params/schema.py:44-73) and the module fields (schema.py:13-22,
# modules/base.py:92-270, bifurcation.py:75-150, levee.py:84-110,
# reservoir.py:76-120).  A name not listed here is placed by its shape.
_C, _P, _PL = ("catchment",), ("bifurcation_path",), ("bifurcation_path", "bifurcation_level")
FIELD_DIMS: dict[str, tuple[str, ...]] = {
    "nx": (), "ny": (), "num_basins": (),
    "basin_sizes": ("basin",),
    "output_catchment_id": ("saved_points",),
    "flood_depth_table": ("catchment", "flood_level"),
    **{k: _C for k in (
        "catchment_id", "catchment_basin_id", "downstream_id", "river_width", "river_length", "river_height",
        "catchment_elevation", "catchment_area", "runoff", "downstream_distance", "river_manning", "flood_manning",
        "river_storage", "flood_storage", "protected_storage", "protected_depth", "river_depth", "flood_depth",
        "river_outflow", "flood_outflow", "river_cross_section_depth", "flood_cross_section_depth",
        "flood_cross_section_area", "flood_fraction", "catchment_x", "catchment_y", "catchment_mainstem_basin_id",
        "upstream_area", "river_mouth_id", "longitude", "latitude", "satellite_width")},
    **{k: _P for k in ("bifurcation_path_id", "bifurcation_catchment_id", "bifurcation_downstream_id",
                       "bifurcation_length", "bifurcation_catchment_x", "bifurcation_downstream_x",
                       "bifurcation_catchment_y", "bifurcation_downstream_y")},
    **{k: _PL for k in ("bifurcation_manning", "bifurcation_width", "bifurcation_elevation", "bifurcation_outflow",
                        "bifurcation_cross_section_depth")},
    **{k: ("levee",) for k in ("levee_id", "levee_catchment_id", "levee_crown_height", "levee_fraction",
                               "levee_base_height", "levee_catchment_x", "levee_catchment_y")},
    **{k: ("reservoir",) for k in ("reservoir_id", "reservoir_catchment_id", "reservoir_capacity",
                                   "reservoir_area") + _RESERVOIR_FIELDS},
    **{k: ("gauge",) for k in ("gauge_catchment_id", "gauge_station_id", "gauge_reported_area_km2",
                               "gauge_allocated_area_km2", "gauge_alloc_error")},
}
# dimension order of a written file
_DIM_ORDER = ("catchment", "flood_level", "bifurcation_path", "bifurcation_level", "levee", "reservoir", "basin",
              "gauge", "saved_points")


class ParameterFileError(ValueError):
    pass


def _pad4(n: int) -> int:
    return (4 - n % 4) % 4


class _Reader:
    def __init__(self, buf: bytes):
        self.buf, self.pos = buf, 0
        if buf[:4] == b"\x89HDF":
            raise ParameterFileError(
                "this is a NETCDF4 (HDF5) file; netCDF4/h5py are not available here: convert it with "
                "`nccopy -k cdf5 <in> <out>` (NetCDF classic, 64-bit data keeps the int64 ids)")
        if buf[:3] != b"CDF" or buf[3] not in (1, 2, 5):
            raise ParameterFileError("not a NetCDF classic file (CDF-1, CDF-2 or CDF-5)")
        self.version = buf[3]
        self.pos = 4

    def i32(self) -> int:
        v = struct.unpack_from(">i", self.buf, self.pos)[0]
        self.pos += 4
        return v

    def i64(self) -> int:
        v = struct.unpack_from(">q", self.buf, self.pos)[0]
        self.pos += 8
        return v

    def count(self) -> int:          # NON_NEG: 8 bytes in CDF-5
        return self.i64() if self.version == 5 else self.i32()

    def offset(self) -> int:         # OFFSET: 8 bytes in CDF-2 and CDF-5
        return self.i32() if self.version == 1 else self.i64()

    def name(self) -> str:
        n = self.count()
        s = self.buf[self.pos:self.pos + n].decode("utf-8")
        self.pos += n + _pad4(n)
        return s

    def values(self, code: int, n: int):
        dt = np.dtype(_TYPES[code])
        raw = self.buf[self.pos:self.pos + n * dt.itemsize]
        self.pos += n * dt.itemsize + _pad4(n * dt.itemsize)
        if code == 2:
            return raw.decode("utf-8", "replace").rstrip("\0")
        return np.frombuffer(raw, dt).astype(dt.newbyteorder("="))

    def attrs(self) -> dict:
        tag, n = self.i32(), self.count()
        if tag not in (0, _NC_ATTRIBUTE):
            raise ParameterFileError("malformed attribute list")
        out = {}
        for _ in range(n):
            key = self.name()
            code = self.i32()
            out[key] = self.values(code, self.count())
        return out


def read_nc(path) -> tuple[dict, dict, dict]:
    """Every variable of a NetCDF classic file: (arrays, dimensions, global
    attributes).  Record (unlimited) variables are not part of the schema and
    are refused."""
    buf = Path(path).read_bytes()
    r = _Reader(buf)
    r.count()                        # numrecs: the schema has no record variables
    tag, n = r.i32(), r.count()
    if tag not in (0, _NC_DIMENSION):
        raise ParameterFileError("malformed dimension list")
    dims = []
    for _ in range(n):
        dims.append((r.name(), r.count()))
    gattrs = r.attrs()
    tag, n = r.i32(), r.count()
    if tag not in (0, _NC_VARIABLE):
        raise ParameterFileError("malformed variable list")
    arrays = {}
    for _ in range(n):
        name = r.name()
        ids = [r.count() for _ in range(r.count())]
        r.attrs()
        code = r.i32()
        r.count()                    # vsize (recomputed from the shape)
        begin = r.offset()
        shape = tuple(dims[d][1] for d in ids)
        if any(dims[d][1] == 0 for d in ids):
            raise ParameterFileError(f"{name}: record variables are not part of the parameter schema")
        if code not in _TYPES or code == 2:
            raise ParameterFileError(f"{name}: unsupported type code {code}")
        if code > 6 and r.version != 5:
            raise ParameterFileError(f"{name}: type code {code} exists only in CDF-5 files (CDF-{r.version} "
                                     f"allows codes 1-6)")
        dt = np.dtype(_TYPES[code])
        count = int(np.prod(shape, dtype=np.int64)) if shape else 1
        if begin + count * dt.itemsize > len(buf):
            raise ParameterFileError(f"{name}: data runs past the end of the file")
        a = np.frombuffer(buf, dt, count, begin).reshape(shape)
        arrays[name] = a.astype(dt.newbyteorder("="))
    return arrays, dict(dims), gattrs


def _index_of(ids: np.ndarray, table: np.ndarray, what: str) -> np.ndarray:
    order = np.argsort(table, kind="stable")
    pos = np.searchsorted(table[order], ids)
    pos = np.minimum(pos, table.shape[0] - 1)
    idx = order[pos]
    bad = table[idx] != ids
    if bad.any():
        raise ParameterFileError(f"{what}: {int(bad.sum())} ids not among catchment_id "
                                 f"(first {ids[bad][:5].tolist()})")
    return idx.astype(np.int32)


def read_parameters(path) -> dict:
    """Parameter fields from a parameters.nc in NetCDF classic encoding, with
    the index fields derived from the ids; ready for ``CaMaFlood(src, ...)``."""
    arrays, dims, _ = read_nc(path)
    if "catchment_id" not in arrays or "downstream_id" not in arrays:
        raise ParameterFileError("catchment_id and downstream_id are required")
    f = dict(arrays)
    cid = np.asarray(f["catchment_id"], np.int64)
    if np.unique(cid).size != cid.size:
        raise ParameterFileError("catchment_id holds duplicates")
    for id_name, idx_name in ID_FIELDS.items():
        if id_name in f:
            f[id_name] = np.asarray(f[id_name], np.int64)
            f[idx_name] = _index_of(f[id_name], cid, id_name)
    for k in ("catchment_id", "catchment_basin_id", "bifurcation_path_id", "reservoir_id"):
        if k in f:
            f[k] = np.asarray(f[k], np.int64)
    return f


def _dims_of(name: str, a: np.ndarray, n: int) -> tuple[str, ...]:
    if name in FIELD_DIMS:
        dims = FIELD_DIMS[name]
        if len(dims) != a.ndim:
            raise ParameterFileError(f"{name}: the schema gives it dimensions {dims}, got shape {a.shape}")
        return dims
    if name.startswith("bifurcation_"):
        return ("bifurcation_path", "bifurcation_level")[:a.ndim]
    if name.startswith("levee_"):
        return ("levee",)
    if name.startswith("reservoir_") or name in _RESERVOIR_FIELDS:
        return ("reservoir",)
    if a.ndim == 1 and a.shape[0] == n:
        return ("catchment",)
    raise ParameterFileError(f"{name}: no dimension of the parameter schema fits shape {a.shape}")


def write_parameters(path, fields: dict, version: int = 5, title: str = "CaMa-Flood parameters") -> None:
    """Write parameter fields in the parameters.nc schema.  Index fields
    (``*_idx``) are left out (the reader derives them from the ids).
    version 5 (CDF-5) keeps int64; versions 1 and 2 store int64 fields as
    int32 and refuse values outside that range."""
    if version not in (1, 2, 5):
        raise ValueError("version must be 1, 2 or 5")
    n = int(np.asarray(fields["catchment_id"]).shape[0])
    items = []
    dim_len: dict[str, int] = {}
    for name, a in fields.items():
        if name.endswith("_idx"):
            continue
        a = np.asarray(a)
        if a.dtype == np.bool_:
            # NC_UBYTE (code 7) exists only in CDF-5: CDF-1/2 files get NC_BYTE
            a = a.astype(np.uint8 if version == 5 else np.int8)
        if version != 5 and a.dtype == np.uint8:
            if a.size and a.max() > 127:
                raise ParameterFileError(f"{name}: uint8 values above 127 need CDF-5 (version=5)")
            a = a.astype(np.int8)
        if version != 5 and a.dtype == np.int64:
            if a.size and (a.min() < -2 ** 31 or a.max() >= 2 ** 31):
                raise ParameterFileError(f"{name}: int64 values need CDF-5 (version=5)")
            a = a.astype(np.int32)
        if version != 5 and a.dtype in (np.uint16, np.uint32, np.uint64):
            raise ParameterFileError(f"{name}: unsigned wide types need CDF-5")
        if a.dtype not in _CODES:
            raise ParameterFileError(f"{name}: unsupported dtype {a.dtype}")
        dn = _dims_of(name, a, n)
        for d, size in zip(dn, a.shape):
            if dim_len.setdefault(d, size) != size:
                raise ParameterFileError(f"{name}: dimension {d} is {size}, elsewhere {dim_len[d]}")
        items.append((name, dn, a))
    order = [d for d in _DIM_ORDER if d in dim_len]
    dim_id = {d: k for k, d in enumerate(order)}

    def count(v: int) -> bytes:
        return struct.pack(">q" if version == 5 else ">i", v)

    def offset(v: int) -> bytes:
        return struct.pack(">i" if version == 1 else ">q", v)

    def name_bytes(s: str) -> bytes:
        b = s.encode("utf-8")
        return count(len(b)) + b + b"\0" * _pad4(len(b))

    def header(begins: list[int]) -> bytes:
        h = b"CDF" + bytes([version]) + count(0)
        h += struct.pack(">i", _NC_DIMENSION) + count(len(order))
        for d in order:
            h += name_bytes(d) + count(dim_len[d])
        t = title.encode("utf-8")
        h += struct.pack(">i", _NC_ATTRIBUTE) + count(1) + name_bytes("title") + struct.pack(">i", 2) \
            + count(len(t)) + t + b"\0" * _pad4(len(t))
        h += struct.pack(">i", _NC_VARIABLE) + count(len(items))
        for (name, dn, a), b in zip(items, begins):
            vsize = a.nbytes + _pad4(a.nbytes)
            if version != 5 and vsize >= 2 ** 31:
                raise ParameterFileError(f"{name}: {vsize} bytes need CDF-5 (version=5)")
            h += name_bytes(name) + count(len(dn)) + b"".join(count(dim_id[d]) for d in dn)
            h += struct.pack(">i", 0) + count(0)                 # no variable attributes
            h += struct.pack(">i", _CODES[a.dtype]) + count(vsize)
            h += offset(b)
        return h

    size = len(header([0] * len(items)))
    begins, pos = [], size
    for _, _, a in items:
        begins.append(pos)
        pos += a.nbytes + _pad4(a.nbytes)
    if version == 1 and pos >= 2 ** 31:
        raise ParameterFileError("file too large for CDF-1; use version 2 or 5")
    with open(path, "wb") as fh:
        fh.write(header(begins))
        for _, _, a in items:
            raw = np.ascontiguousarray(a, a.dtype.newbyteorder(">")).tobytes()
            fh.write(raw + b"\0" * _pad4(len(raw)))

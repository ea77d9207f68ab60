"""Minimal Exodus II mesh reader (netCDF classic files, CDF-1 and CDF-2).

Mesh ingestion for txa_mesh_upload: the counterpart of the reference's
STK_ExodusReaderFactory (adapters-stk/src/stk_interface/Panzer_STK_ExodusReaderFactory.cpp),
which reads through SEACAS/stk_io (not in this image).  This reader follows the
published formats: the netCDF classic header (magic "CDF" + version, dimension,
attribute and variable lists, big-endian, names and values padded to 4 bytes)
and the Exodus II variables coordx / coordy / coordz (or coord[num_dim][num_nodes])
and connect<k>[num_el_in_blk<k>][num_nod_per_el<k>] with its elem_type attribute,
1-based node numbers.  Only non-record variables are read (coordinates and
connectivity are not time-dependent).
"""
from __future__ import annotations

import struct

import numpy as np

_NC_DIMENSION, _NC_VARIABLE, _NC_ATTRIBUTE = 0x0A, 0x0B, 0x0C
_TYPES = {1: ("i1", 1), 2: ("S1", 1), 3: (">i2", 2), 4: (">i4", 4), 5: (">f4", 4), 6: (">f8", 8)}


class _Reader:
    def __init__(self, buf: bytes):
        self.b, self.p = buf, 0

    def i32(self):
        v = struct.unpack_from(">i", self.b, self.p)[0]
        self.p += 4
        return v

    def i64(self):
        v = struct.unpack_from(">q", self.b, self.p)[0]
        self.p += 8
        return v

    def name(self):
        n = self.i32()
        s = self.b[self.p:self.p + n].decode("utf-8")
        self.p += (n + 3) // 4 * 4
        return s

    def values(self, nc_type, n):
        dt, size = _TYPES[nc_type]
        raw = self.b[self.p:self.p + n * size]
        self.p += (n * size + 3) // 4 * 4
        if nc_type == 2:
            return raw.split(b"\0", 1)[0].decode("utf-8", "replace")
        return np.frombuffer(raw, dtype=dt)

    def attrs(self):
        tag, n = self.i32(), self.i32()
        if tag == 0 and n == 0:
            return {}
        if tag != _NC_ATTRIBUTE:
            raise ValueError("netCDF: attribute list expected")
        out = {}
        for _ in range(n):
            nm = self.name()
            t = self.i32()
            out[nm] = self.values(t, self.i32())
        return out


def read_netcdf(path):
    """(dims {name: length}, global attributes, variables {name: ndarray}) of a classic netCDF file
    (non-record variables only)."""
    buf = open(path, "rb").read()
    if buf[:3] != b"CDF" or buf[3] not in (1, 2):
        raise ValueError(f"{path}: not a netCDF classic file")
    r = _Reader(buf)
    r.p = 4
    offset64 = buf[3] == 2
    r.i32()                                   # numrecs
    dims = []
    tag, n = r.i32(), r.i32()
    if tag == _NC_DIMENSION:
        for _ in range(n):
            dims.append((r.name(), r.i32()))
    elif (tag, n) != (0, 0):
        raise ValueError("netCDF: dimension list expected")
    gatts = r.attrs()
    variables = {}
    tag, n = r.i32(), r.i32()
    if tag not in (_NC_VARIABLE, 0):
        raise ValueError("netCDF: variable list expected")
    for _ in range(n if tag == _NC_VARIABLE else 0):
        nm = r.name()
        dimids = [r.i32() for _ in range(r.i32())]
        vatts = r.attrs()
        t = r.i32()
        r.i32()                               # vsize
        begin = r.i64() if offset64 else r.i32()
        if dimids and dims[dimids[0]][1] == 0:
            continue                          # record variable (unlimited first dimension): not read
        shape = [dims[d][1] for d in dimids]
        dt, size = _TYPES[t]
        count = int(np.prod(shape)) if shape else 1
        data = np.frombuffer(buf, dtype=dt, count=count, offset=begin).reshape(shape)
        variables[nm] = (data, vatts)
    return dict(dims), gatts, variables


def read_exodus(path):
    """Coordinates [nodes][3] (float64) and element blocks [(elem_type, connectivity [elements][nodes]
    0-based int64)] of an Exodus II file."""
    dims, _, v = read_netcdf(path)
    nn = dims["num_nodes"]
    ndim = dims.get("num_dim", 3)
    xyz = np.zeros((nn, 3))
    if "coord" in v:
        c = v["coord"][0]
        xyz[:, :ndim] = np.asarray(c, np.float64).reshape(ndim, nn).T
    else:
        for d, nm in enumerate(("coordx", "coordy", "coordz")[:ndim]):
            xyz[:, d] = np.asarray(v[nm][0], np.float64)
    blocks = []
    k = 1
    while f"connect{k}" in v:
        conn, att = v[f"connect{k}"]
        etype = att.get("elem_type", "")
        blocks.append((str(etype).upper(), np.asarray(conn, np.int64) - 1))
        k += 1
    return xyz, blocks

"""NetCDF versions of binary input cases, for the reader tests.

CDF-1 and CDF-2 files are written with scipy.io.netcdf_file, an independent
implementation of the NetCDF classic formats; CDF-5 files (which scipy
cannot write) with `write_cdf5` below, from the published CDF-5 layout.
Variable names and shapes are those the reference's NetCDF reader asks for
(swe_code/io/input.c:14-186): the weights on the unit sphere and the ids
1-based, as nc2bin.c expects them.
"""
import struct

import numpy as np
from scipy.io import netcdf_file

GRAV = 9.80616


def case_vars(raw, types=None):
    """(name, dims, numpy array) of every variable, in file order."""
    types = types or {}
    N, n = raw.n_nodes, raw.nnbr
    out = [("Nnodes", (), np.int32(N)), ("Nnbr", (), np.int32(n)),
           ("a", (), np.float64(raw.a)), ("g", (), np.float64(GRAV)),
           ("gh0", (), np.float64(raw.gh0)), ("timestep", (), np.float64(raw.dt)),
           ("gamma", (), np.float64(raw.gamma))]
    for k, nm in enumerate(("x", "y", "z")):
        out.append((nm, ("nodes",), raw.X[:, k]))
    out += [("f", ("nodes",), raw.f), ("ghm", ("nodes",), raw.ghm)]
    for nm in ("p_u", "p_v", "p_w", "gradghm"):
        out.append((nm, ("nodes", "three"), getattr(raw, nm)))
    for nm, arr in (("DPx", raw.Dx), ("DPy", raw.Dy), ("DPz", raw.Dz), ("L", raw.L)):
        out.append((nm, ("nodes", "nbr"), arr))
    out.append(("idx", ("nodes", "nbr"), (raw.idx + 1).astype(np.int32)))
    out += [("H", ("nodes", "four"), raw.H_init), ("H_100", ("nodes", "four"), raw.H_final)]
    return [(nm, dims, np.asarray(a).astype(types.get(nm, np.asarray(a).dtype)))
            for nm, dims, a in out]


def write_case_nc(raw, path, version=1, types=None, record_state=False, drop=()):
    """scipy writer: version 1 (classic) or 2 (64-bit offset); `types` maps a
    variable to the numpy dtype it is stored in; `record_state` stores H and
    H_100 as record variables of one record ("nodes": one record per node);
    `drop` leaves variables out."""
    f = netcdf_file(str(path), "w", version=version)
    if record_state:   # the record dimension is the first one scipy accepts
        f.createDimension("time", None)
    f.createDimension("nodes", raw.n_nodes)
    f.createDimension("nbr", raw.nnbr)
    f.createDimension("three", 3)
    f.createDimension("four", 4)
    for nm, dims, a in case_vars(raw, types):
        if nm in drop:
            continue
        if record_state == "nodes" and nm in ("H", "H_100"):
            # one record per node: the records of H and H_100 interleave
            v = f.createVariable(nm, a.dtype, ("time", "four"))
            v[:] = a
            continue
        if record_state and nm in ("H", "H_100"):
            v = f.createVariable(nm, a.dtype, ("time",) + dims)
            v[0] = a
            continue
        v = f.createVariable(nm, a.dtype, dims)
        v[...] = a
    f.close()


_CDF5_TYPE = {np.dtype(np.int8): 1, np.dtype(np.int16): 3, np.dtype(np.int32): 4,
              np.dtype(np.float32): 5, np.dtype(np.float64): 6, np.dtype(np.uint8): 7,
              np.dtype(np.uint16): 8, np.dtype(np.uint32): 9, np.dtype(np.int64): 10,
              np.dtype(np.uint64): 11}


def write_cdf5(raw, path, types=None):
    """CDF-5: magic CDF\\x05, 64-bit counts, dimension lengths, dimension ids,
    vsize and begin; data big-endian, each variable padded to 4 bytes."""
    vars_ = case_vars(raw, types)
    dims = [("nodes", raw.n_nodes), ("nbr", raw.nnbr), ("three", 3), ("four", 4)]
    dim_id = {nm: k for k, (nm, _) in enumerate(dims)}

    def name(s):
        b = s.encode()
        return struct.pack(">q", len(b)) + b + b"\0" * (-len(b) % 4)

    def header(begins):
        h = b"CDF\x05" + struct.pack(">q", 0)
        h += struct.pack(">Iq", 0x0A, len(dims))
        for nm, ln in dims:
            h += name(nm) + struct.pack(">q", ln)
        h += struct.pack(">Iq", 0, 0)   # no global attributes
        h += struct.pack(">Iq", 0x0B, len(vars_))
        for (nm, dd, a), b in zip(vars_, begins):
            h += name(nm) + struct.pack(">q", len(dd))
            h += b"".join(struct.pack(">q", dim_id[d]) for d in dd)
            h += struct.pack(">Iq", 0, 0)   # no attributes
            vsize = a.nbytes + (-a.nbytes % 4)
            h += struct.pack(">Iqq", _CDF5_TYPE[a.dtype], vsize, b)
        return h

    sizes = [a.nbytes + (-a.nbytes % 4) for _, _, a in vars_]
    hlen = len(header([0] * len(vars_)))
    begins, off = [], hlen
    for s in sizes:
        begins.append(off)
        off += s
    with open(path, "wb") as fp:
        fp.write(header(begins))
        for (_, _, a), s in zip(vars_, sizes):
            b = np.ascontiguousarray(a).astype(a.dtype.newbyteorder(">")).tobytes()
            fp.write(b + b"\0" * (s - len(b)))

"""hmap::Array interchange formats for numpy arrays of shape (ny, nx).

Mirrors HighMap/src/array/io.cpp:28-66,111-131:

* raw float32 (`to_file` / `from_file`): the values in memory order, index
  j*nx + i, no header; reading needs the shape;
* `.npy` (`to_numpy` / `from_numpy`): written as float32 of shape (nx, ny) in
  Fortran order, which is the same memory order, so `numpy.load` of such a
  file gives the transpose of the (ny, nx) array; read back in either order,
  a C-order (nx, ny) file being transposed on the way in.

The .npy reader and writer are this module's own (format version 1.0 is
written; 1.0, 2.0 and 3.0 are read); only little-endian float32 is accepted,
as HighMap reads its arrays as float.
"""
import ast
import struct

import numpy as np

_MAGIC = b"\x93NUMPY"


def to_file(path, z):
    """Array::to_file: raw float32, index j*nx + i."""
    np.ascontiguousarray(z, np.float32).tofile(path)


def from_file(path, shape):
    """Array::from_file: nx*ny raw float32 values into shape (nx, ny);
    returns a (ny, nx) array."""
    nx, ny = int(shape[0]), int(shape[1])
    a = np.fromfile(path, np.float32, count=nx * ny)
    if a.size != nx * ny:
        raise ValueError("%s holds %d floats, %d expected" % (path, a.size, nx * ny))
    return a.reshape(ny, nx)


def _header(nx, ny):
    d = "{'descr': '<f4', 'fortran_order': True, 'shape': (%d, %d), }" % (nx, ny)
    total = (10 + len(d) + 1 + 63) // 64 * 64
    return (d + " " * (total - 10 - len(d) - 1) + "\n").encode("latin1")


def to_numpy(path, z):
    """Array::to_numpy: float32, shape (nx, ny), Fortran order."""
    z = np.ascontiguousarray(z, np.float32)
    if z.ndim != 2:
        raise ValueError("z must be 2-D (ny, nx)")
    ny, nx = z.shape
    h = _header(nx, ny)
    with open(path, "wb") as f:
        f.write(_MAGIC + bytes([1, 0]) + struct.pack("<H", len(h)) + h)
        f.write(z.astype("<f4", copy=False).tobytes())


def from_numpy(path):
    """Array::from_numpy: the (nx, ny) array of the file, in either order;
    returns it as a (ny, nx) array."""
    with open(path, "rb") as f:
        pre = f.read(8)
        if len(pre) != 8 or pre[:6] != _MAGIC:
            raise ValueError("%s is not a .npy file" % path)
        if pre[6] == 1:
            (hl,) = struct.unpack("<H", f.read(2))
        elif pre[6] in (2, 3):
            (hl,) = struct.unpack("<I", f.read(4))
        else:
            raise ValueError("unknown .npy version %d" % pre[6])
        h = ast.literal_eval(f.read(hl).decode("latin1"))
        if h.get("descr") != "<f4":
            raise ValueError("%s: dtype %r is not little-endian float32" % (path, h.get("descr")))
        shape = tuple(h["shape"])
        if len(shape) != 2:
            raise ValueError("%s: shape %r is not 2-D" % (path, shape))
        nx, ny = shape
        data = np.fromfile(f, "<f4", count=nx * ny)
    if data.size != nx * ny:
        raise ValueError("%s: short data" % path)
    if h["fortran_order"]:
        return data.reshape(ny, nx).astype(np.float32, copy=False)
    return np.ascontiguousarray(data.reshape(nx, ny).T, np.float32)

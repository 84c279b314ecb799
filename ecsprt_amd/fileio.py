"""The reference's on-disk formats (src/file-io.c) through the library's C entry points.

HDF5 fluid files (OutputFluidHDF5: datasets "density", "x flux", "y flux", "z flux",
"rhs", "potential"), the HDF5 density input of LoadFluidQuantities ("density-kji") and
PETSc binary vectors (OutputSwarmBinary).  HDF5 is loaded by the library at run time;
`hdf5_available()` says whether it was found.
"""
import ctypes
import os

import numpy as np

from . import _lib

FLUID_FIELDS = ("density", "x flux", "y flux", "z flux")


class FileIOError(RuntimeError):
    pass


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _check(rc):
    if rc != _lib.OK:
        msg = _lib.load().ecsprt_io_last_error()
        raise FileIOError(f"code {rc}: {msg.decode() if msg else ''}")


def hdf5_available():
    """(True, "") when the HDF5 library loads, else (False, reason)."""
    buf = ctypes.create_string_buffer(512)
    ok = _lib.load().ecsprt_io_hdf5_available(buf, 512)
    return bool(ok), buf.value.decode()


def write_fluid_hdf5(path, moments, rhs, phi):
    """OutputFluidHDF5's file from host arrays: moments [.., 4] on the grid shape of
    rhs and phi ([Nz, Ny, Nx], or [Ny, Nx] in two dimensions)."""
    rhs = np.ascontiguousarray(rhs, dtype=np.float64)
    phi = np.ascontiguousarray(phi, dtype=np.float64)
    m = np.ascontiguousarray(moments, dtype=np.float64)
    if rhs.ndim not in (2, 3) or phi.shape != rhs.shape or m.shape != rhs.shape + (4,):
        raise ValueError("moments must be [grid..., 4] and rhs, phi the grid shape")
    shape = rhs.shape if rhs.ndim == 3 else (1,) + rhs.shape
    nz, ny, nx = shape
    _check(_lib.load().ecsprt_io_write_fluid_hdf5(os.fsencode(path), rhs.ndim, nx, ny, nz, _dp(m), _dp(rhs), _dp(phi)))


def write_dataset(path, name, data):
    """A new HDF5 file holding one dataset (e.g. a "density-kji" input field)."""
    data = np.ascontiguousarray(data, dtype=np.float64)
    dims = (ctypes.c_long * data.ndim)(*data.shape)
    _check(_lib.load().ecsprt_io_write_hdf5_dataset(os.fsencode(path), name.encode(), data.ndim, dims, _dp(data)))


def read_dataset(path, name):
    """Dataset `name` of an HDF5 file as a float64 array of its shape."""
    L = _lib.load()
    nd = ctypes.c_int()
    dims = (ctypes.c_long * 4)()
    _check(L.ecsprt_io_read_hdf5_dataset(os.fsencode(path), name.encode(), ctypes.byref(nd), dims, None, 0))
    shape = tuple(dims[q] for q in range(nd.value))
    out = np.empty(shape, dtype=np.float64)
    _check(L.ecsprt_io_read_hdf5_dataset(os.fsencode(path), name.encode(), ctypes.byref(nd), dims, _dp(out),
                                         out.size))
    return out


def read_fluid_hdf5(path):
    """The datasets of a fluid file: {"density": .., "x flux": .., ..., "rhs": .., "potential": ..}."""
    return {name: read_dataset(path, name) for name in FLUID_FIELDS + ("rhs", "potential")}


def write_vec_binary(path, values, bs=1):
    """A PETSc binary vector (VecView layout) with block size bs in "<path>.info"."""
    v = np.ascontiguousarray(values, dtype=np.float64).reshape(-1)
    _check(_lib.load().ecsprt_io_write_vec_binary(os.fsencode(path), v.size, bs, _dp(v)))


def read_vec_binary(path):
    """(values, block size) of a PETSc binary vector."""
    L = _lib.load()
    n, bs = ctypes.c_long(), ctypes.c_int()
    _check(L.ecsprt_io_read_vec_binary(os.fsencode(path), None, 0, ctypes.byref(n), ctypes.byref(bs)))
    out = np.empty(n.value, dtype=np.float64)
    _check(L.ecsprt_io_read_vec_binary(os.fsencode(path), _dp(out), out.size, ctypes.byref(n), ctypes.byref(bs)))
    return out, bs.value

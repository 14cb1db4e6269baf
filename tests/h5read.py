"""Minimal HDF5 reader over ctypes (libhdf5 of the image, /opt/conda/lib),
independent of the library's own I/O module: dataset names, shapes and
values of a file, for the tests of the Data / Resume writers."""
import ctypes as C
import os

import numpy as np

H5_LIB = os.environ.get("PV_HDF5_LIB", "/opt/conda/lib/libhdf5.so")


class H5File:
    def __init__(self, path):
        h5 = C.CDLL(H5_LIB)
        h5.H5open()
        h5.H5Fopen.restype = C.c_int64
        h5.H5Fopen.argtypes = [C.c_char_p, C.c_uint, C.c_int64]
        h5.H5Dopen2.restype = C.c_int64
        h5.H5Dopen2.argtypes = [C.c_int64, C.c_char_p, C.c_int64]
        h5.H5Dget_space.restype = C.c_int64
        h5.H5Dget_space.argtypes = [C.c_int64]
        h5.H5Sget_simple_extent_ndims.argtypes = [C.c_int64]
        h5.H5Sget_simple_extent_dims.argtypes = [C.c_int64, C.POINTER(C.c_uint64), C.c_void_p]
        h5.H5Dread.argtypes = [C.c_int64, C.c_int64, C.c_int64, C.c_int64, C.c_int64, C.c_void_p]
        h5.H5Dclose.argtypes = [C.c_int64]
        h5.H5Sclose.argtypes = [C.c_int64]
        h5.H5Fclose.argtypes = [C.c_int64]
        h5.H5Lexists.argtypes = [C.c_int64, C.c_char_p, C.c_int64]
        self.h5 = h5
        self.f = h5.H5Fopen(path.encode(), 0, 0)          # H5F_ACC_RDONLY, H5P_DEFAULT
        assert self.f >= 0, path
        self.t_double = C.c_int64.in_dll(h5, "H5T_NATIVE_DOUBLE_g").value
        self.t_int = C.c_int64.in_dll(h5, "H5T_NATIVE_INT_g").value

    def has(self, name):
        parts = name.strip("/").split("/")
        for k in range(1, len(parts) + 1):
            if self.h5.H5Lexists(self.f, ("/" + "/".join(parts[:k])).encode(), 0) <= 0:
                return False
        return True

    def shape(self, name):
        d = self.h5.H5Dopen2(self.f, name.encode(), 0)
        assert d >= 0, name
        sp = self.h5.H5Dget_space(d)
        nd = self.h5.H5Sget_simple_extent_ndims(sp)
        dims = (C.c_uint64 * max(nd, 1))()
        self.h5.H5Sget_simple_extent_dims(sp, dims, None)
        self.h5.H5Sclose(sp)
        self.h5.H5Dclose(d)
        return tuple(int(dims[k]) for k in range(nd))

    def read(self, name, integer=False):
        shape = self.shape(name)
        a = np.zeros(shape, dtype=np.int32 if integer else np.float64)
        d = self.h5.H5Dopen2(self.f, name.encode(), 0)
        rc = self.h5.H5Dread(d, self.t_int if integer else self.t_double, 0, 0, 0, a.ctypes.data)
        self.h5.H5Dclose(d)
        assert rc >= 0, name
        return a

    def close(self):
        if self.f >= 0:
            self.h5.H5Fclose(self.f)
            self.f = -1

"""A minimal HDF5 reader over the image's libhdf5 (ctypes), independent of
the library under test: lists a file's datasets and reads datasets and
attributes with their stored type class and size."""
import ctypes as C
from pathlib import Path

import numpy as np

HDF5_LIB = Path("/opt/conda/lib/libhdf5.so.103")
hid_t = C.c_int64
hsize_t = C.c_uint64


class GInfo(C.Structure):
    _fields_ = [("storage_type", C.c_int), ("nlinks", hsize_t), ("max_corder", C.c_int64), ("mounted", C.c_bool)]


_H = None


def lib():
    global _H
    if _H is None:
        H = C.CDLL(str(HDF5_LIB))
        H.H5open()
        for f in ("H5Fopen",):
            getattr(H, f).restype = hid_t
        H.H5Fopen.argtypes = [C.c_char_p, C.c_uint, hid_t]
        H.H5Fclose.argtypes = [hid_t]
        H.H5Dopen2.restype = hid_t
        H.H5Dopen2.argtypes = [hid_t, C.c_char_p, hid_t]
        H.H5Dclose.argtypes = [hid_t]
        H.H5Dget_space.restype = hid_t
        H.H5Dget_space.argtypes = [hid_t]
        H.H5Dget_type.restype = hid_t
        H.H5Dget_type.argtypes = [hid_t]
        H.H5Sget_simple_extent_ndims.argtypes = [hid_t]
        H.H5Sget_simple_extent_dims.argtypes = [hid_t, C.POINTER(hsize_t), C.POINTER(hsize_t)]
        H.H5Sclose.argtypes = [hid_t]
        H.H5Tget_class.argtypes = [hid_t]
        H.H5Tget_size.restype = C.c_size_t
        H.H5Tget_size.argtypes = [hid_t]
        H.H5Tclose.argtypes = [hid_t]
        H.H5Dread.argtypes = [hid_t, hid_t, hid_t, hid_t, hid_t, C.c_void_p]
        H.H5Aopen.restype = hid_t
        H.H5Aopen.argtypes = [hid_t, C.c_char_p, hid_t]
        H.H5Aget_space.restype = hid_t
        H.H5Aget_space.argtypes = [hid_t]
        H.H5Aget_type.restype = hid_t
        H.H5Aget_type.argtypes = [hid_t]
        H.H5Aread.argtypes = [hid_t, hid_t, C.c_void_p]
        H.H5Aclose.argtypes = [hid_t]
        H.H5Gget_info.argtypes = [hid_t, C.POINTER(GInfo)]
        H.H5Lget_name_by_idx.restype = C.c_ssize_t
        H.H5Lget_name_by_idx.argtypes = [hid_t, C.c_char_p, C.c_int, C.c_int, hsize_t, C.c_char_p, C.c_size_t, hid_t]
        _H = H
    return _H


def _native(name):
    return hid_t.in_dll(lib(), name).value


TYPE_CLASS = {0: "int", 1: "float"}


class File:
    def __init__(self, path):
        self.H = lib()
        self.f = self.H.H5Fopen(str(path).encode(), 0, 0)  # read only, default properties
        if self.f < 0:
            raise OSError(f"cannot open {path}")

    def close(self):
        if self.f >= 0:
            self.H.H5Fclose(self.f)
            self.f = -1

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def names(self):
        info = GInfo()
        assert self.H.H5Gget_info(self.f, C.byref(info)) >= 0
        out = []
        for i in range(info.nlinks):
            buf = C.create_string_buffer(256)
            self.H.H5Lget_name_by_idx(self.f, b".", 0, 0, i, buf, 256, 0)
            out.append(buf.value.decode())
        return sorted(out)

    def _read(self, obj, space_of, type_of, read):
        H = self.H
        sp = space_of(obj)
        nd = H.H5Sget_simple_extent_ndims(sp)
        dims = (hsize_t * max(nd, 1))()
        H.H5Sget_simple_extent_dims(sp, dims, None)
        H.H5Sclose(sp)
        shape = tuple(int(dims[i]) for i in range(nd))
        t = type_of(obj)
        cls, size = TYPE_CLASS.get(H.H5Tget_class(t), "other"), int(H.H5Tget_size(t))
        H.H5Tclose(t)
        if cls == "float":
            arr, mem = np.zeros(shape, np.float64), _native("H5T_NATIVE_DOUBLE_g")
        else:
            arr, mem = np.zeros(shape, np.int64), _native("H5T_NATIVE_LLONG_g")
        if arr.size:
            assert read(obj, mem, arr.ctypes.data) >= 0
        return arr, cls, size

    def dataset(self, name):
        """(values, type class, stored size in bytes)"""
        H = self.H
        d = H.H5Dopen2(self.f, name.encode(), 0)
        if d < 0:
            raise KeyError(name)
        try:
            return self._read(d, H.H5Dget_space, H.H5Dget_type, lambda o, m, p: H.H5Dread(o, m, 0, 0, 0, p))
        finally:
            H.H5Dclose(d)

    def attribute(self, dset, name):
        H = self.H
        d = H.H5Dopen2(self.f, dset.encode(), 0)
        a = H.H5Aopen(d, name.encode(), 0)
        try:
            if a < 0:
                raise KeyError(f"{dset}@{name}")
            return self._read(a, H.H5Aget_space, H.H5Aget_type, lambda o, m, p: H.H5Aread(o, m, p))
        finally:
            if a >= 0:
                H.H5Aclose(a)
            H.H5Dclose(d)

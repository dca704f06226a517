"""A minimal HDF5 reader for the tests, calling libhdf5's C API directly
through ctypes: it shares no code with the writer under test.

    with H5File(path) as f:
        f.groups("/")            -> child group names
        f.attrs("/B0_0_")        -> {name: numpy array (1-D)}
        f.attr_types("/B0_0_")   -> {name: (class, size in bytes)}
        f.dataset("/B0_0_/field_density") -> numpy array of the file's shape
"""
import ctypes
import os

import numpy as np

HDF5_LIB = "/opt/conda/lib/libhdf5.so"
H5T_INTEGER, H5T_FLOAT = 0, 1


def available():
    return os.path.exists(HDF5_LIB)


_h5 = None


def lib():
    global _h5
    if _h5 is None:
        L = ctypes.CDLL(HDF5_LIB)
        hid, herr, hsize = ctypes.c_int64, ctypes.c_int, ctypes.c_ulonglong
        L.H5open.restype = herr
        L.H5Eset_auto2.argtypes = [hid, ctypes.c_void_p, ctypes.c_void_p]
        L.H5Fopen.argtypes = [ctypes.c_char_p, ctypes.c_uint, hid]
        L.H5Fopen.restype = hid
        L.H5Fclose.argtypes = [hid]
        L.H5Gopen2.argtypes = [hid, ctypes.c_char_p, hid]
        L.H5Gopen2.restype = hid
        L.H5Gclose.argtypes = [hid]
        L.H5Lget_name_by_idx.argtypes = [hid, ctypes.c_char_p, ctypes.c_int, ctypes.c_int,
                                         hsize, ctypes.c_char_p, ctypes.c_size_t, hid]
        L.H5Lget_name_by_idx.restype = ctypes.c_ssize_t
        L.H5Aget_num_attrs.argtypes = [hid]
        L.H5Aget_num_attrs.restype = ctypes.c_int
        L.H5Aopen_idx.argtypes = [hid, ctypes.c_uint]
        L.H5Aopen_idx.restype = hid
        L.H5Aget_name.argtypes = [hid, ctypes.c_size_t, ctypes.c_char_p]
        L.H5Aget_name.restype = ctypes.c_ssize_t
        L.H5Aget_space.argtypes = [hid]
        L.H5Aget_space.restype = hid
        L.H5Aget_type.argtypes = [hid]
        L.H5Aget_type.restype = hid
        L.H5Aread.argtypes = [hid, hid, ctypes.c_void_p]
        L.H5Aclose.argtypes = [hid]
        L.H5Dopen2.argtypes = [hid, ctypes.c_char_p, hid]
        L.H5Dopen2.restype = hid
        L.H5Dget_space.argtypes = [hid]
        L.H5Dget_space.restype = hid
        L.H5Dget_type.argtypes = [hid]
        L.H5Dget_type.restype = hid
        L.H5Dread.argtypes = [hid, hid, hid, hid, hid, ctypes.c_void_p]
        L.H5Dclose.argtypes = [hid]
        L.H5Sget_simple_extent_ndims.argtypes = [hid]
        L.H5Sget_simple_extent_dims.argtypes = [hid, ctypes.POINTER(hsize), ctypes.c_void_p]
        L.H5Sclose.argtypes = [hid]
        L.H5Tget_class.argtypes = [hid]
        L.H5Tget_size.argtypes = [hid]
        L.H5Tget_size.restype = ctypes.c_size_t
        L.H5Tclose.argtypes = [hid]
        L.H5open()
        L.H5Eset_auto2(0, None, None)
        _h5 = L
    return _h5


def _native(L, cls, size):
    name = {(H5T_FLOAT, 8): "H5T_NATIVE_DOUBLE_g", (H5T_INTEGER, 4): "H5T_NATIVE_INT_g",
            (H5T_INTEGER, 8): "H5T_NATIVE_LLONG_g", (H5T_INTEGER, 1): "H5T_NATIVE_SCHAR_g"}
    dt = {(H5T_FLOAT, 8): np.float64, (H5T_INTEGER, 4): np.int32,
          (H5T_INTEGER, 8): np.int64, (H5T_INTEGER, 1): np.int8}
    return ctypes.c_int64.in_dll(L, name[(cls, size)]).value, dt[(cls, size)]


def _dims(L, space):
    nd = L.H5Sget_simple_extent_ndims(space)
    d = (ctypes.c_ulonglong * max(1, nd))()
    L.H5Sget_simple_extent_dims(space, d, None)
    return tuple(int(d[i]) for i in range(nd))


class H5File:
    def __init__(self, path):
        self.L = lib()
        self.fid = self.L.H5Fopen(path.encode(), 0, 0)
        if self.fid < 0:
            raise IOError("cannot open %s" % path)

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.L.H5Fclose(self.fid)

    def groups(self, path="/"):
        out, i = [], 0
        buf = ctypes.create_string_buffer(256)
        while self.L.H5Lget_name_by_idx(self.fid, path.encode(), 0, 0, i, buf, 256, 0) > 0:
            out.append(buf.value.decode())
            i += 1
        return out

    def _open(self, path):
        return self.L.H5Gopen2(self.fid, path.encode(), 0)

    def _each_attr(self, path):
        L = self.L
        loc = self._open(path)
        assert loc >= 0, path
        try:
            for i in range(L.H5Aget_num_attrs(loc)):
                a = L.H5Aopen_idx(loc, i)
                buf = ctypes.create_string_buffer(256)
                L.H5Aget_name(a, 256, buf)
                t, s = L.H5Aget_type(a), L.H5Aget_space(a)
                yield buf.value.decode(), a, L.H5Tget_class(t), int(L.H5Tget_size(t)), _dims(L, s)
                L.H5Tclose(t)
                L.H5Sclose(s)
                L.H5Aclose(a)
        finally:
            L.H5Gclose(loc)

    def attr_types(self, path="/"):
        return {n: (c, z, d) for n, _, c, z, d in self._each_attr(path)}

    def attrs(self, path="/"):
        out = {}
        for name, a, cls, size, dims in self._each_attr(path):
            native, dt = _native(self.L, cls, size)
            v = np.zeros(dims, dtype=dt)
            assert self.L.H5Aread(a, native, v.ctypes.data) >= 0
            out[name] = v
        return out

    def dataset(self, path):
        L = self.L
        ds = L.H5Dopen2(self.fid, path.encode(), 0)
        if ds < 0:
            raise KeyError(path)
        t, s = L.H5Dget_type(ds), L.H5Dget_space(ds)
        native, dt = _native(L, L.H5Tget_class(t), int(L.H5Tget_size(t)))
        v = np.zeros(_dims(L, s), dtype=dt)
        assert L.H5Dread(ds, native, 0, 0, 0, v.ctypes.data) >= 0
        L.H5Tclose(t)
        L.H5Sclose(s)
        L.H5Dclose(ds)
        return v

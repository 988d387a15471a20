"""A reader of the HDF5 files the diagnostics write, on the HDF5 C library
of the image (ctypes; test infrastructure): datasets as numpy arrays,
attributes as str / float / int, and the objects of a group; and a writer
of plain double datasets (the particle files of "file" beams)."""
import ctypes as C
import os

import numpy as np

HDF5 = os.path.join(os.environ.get("QPAD_HDF5_ROOT", "/opt/conda"), "lib", "libhdf5.so")
_h = None
hid_t, herr_t, hsize_t = C.c_int64, C.c_int, C.c_uint64
H5T_FLOAT, H5T_INTEGER, H5T_STRING = 1, 0, 3


def lib():
    global _h
    if _h is None:
        h = C.CDLL(HDF5)
        h.H5open()
        for name, res, args in (
                ("H5Fopen", hid_t, [C.c_char_p, C.c_uint, hid_t]), ("H5Fclose", herr_t, [hid_t]),
                ("H5Dopen2", hid_t, [hid_t, C.c_char_p, hid_t]), ("H5Dclose", herr_t, [hid_t]),
                ("H5Dget_space", hid_t, [hid_t]), ("H5Dread", herr_t, [hid_t, hid_t, hid_t, hid_t, hid_t, C.c_void_p]),
                ("H5Sget_simple_extent_ndims", C.c_int, [hid_t]),
                ("H5Sget_simple_extent_dims", C.c_int, [hid_t, C.POINTER(hsize_t), C.POINTER(hsize_t)]),
                ("H5Sclose", herr_t, [hid_t]), ("H5Oopen", hid_t, [hid_t, C.c_char_p, hid_t]),
                ("H5Oclose", herr_t, [hid_t]), ("H5Aopen", hid_t, [hid_t, C.c_char_p, hid_t]),
                ("H5Aclose", herr_t, [hid_t]), ("H5Aget_type", hid_t, [hid_t]),
                ("H5Aread", herr_t, [hid_t, hid_t, C.c_void_p]), ("H5Tget_class", C.c_int, [hid_t]),
                ("H5Tget_size", C.c_size_t, [hid_t]), ("H5Tget_strpad", C.c_int, [hid_t]),
                ("H5Tclose", herr_t, [hid_t]), ("H5Aget_num_attrs", C.c_int, [hid_t]),
                ("H5Gget_num_objs", herr_t, [hid_t, C.POINTER(hsize_t)]),
                ("H5Gget_objname_by_idx", C.c_ssize_t, [hid_t, hsize_t, C.c_char_p, C.c_size_t]),
                ("H5Gopen2", hid_t, [hid_t, C.c_char_p, hid_t]), ("H5Gclose", herr_t, [hid_t]),
                ("H5Fcreate", hid_t, [C.c_char_p, C.c_uint, hid_t, hid_t]),
                ("H5Screate_simple", hid_t, [C.c_int, C.POINTER(hsize_t), C.POINTER(hsize_t)]),
                ("H5Dcreate2", hid_t, [hid_t, C.c_char_p, hid_t, hid_t, hid_t, hid_t, hid_t]),
                ("H5Dwrite", herr_t, [hid_t, hid_t, hid_t, hid_t, hid_t, C.c_void_p])):
            f = getattr(h, name)
            f.restype, f.argtypes = res, args
        _h = h
    return _h


def _native(name):
    return hid_t.in_dll(lib(), name).value


def write_datasets(path, arrays):
    """a new file at path with one float64 dataset per item of arrays (the
    array's own shape)"""
    h = lib()
    f = h.H5Fcreate(path.encode(), 2, 0, 0)      # H5F_ACC_TRUNC
    assert f >= 0, path
    dbl = _native("H5T_NATIVE_DOUBLE_g")
    for name, a in arrays.items():
        a = np.ascontiguousarray(a, dtype=np.float64)
        dims = (hsize_t * a.ndim)(*a.shape)
        sp = h.H5Screate_simple(a.ndim, dims, None)
        d = h.H5Dcreate2(f, name.encode(), _native("H5T_IEEE_F64LE_g"), sp, 0, 0, 0)
        assert d >= 0 and h.H5Dwrite(d, dbl, 0, 0, 0, a.ctypes.data_as(C.c_void_p)) >= 0, name
        h.H5Dclose(d)
        h.H5Sclose(sp)
    h.H5Fclose(f)


class File:
    def __init__(self, path):
        self.fid = lib().H5Fopen(path.encode(), 0, 0)   # H5F_ACC_RDONLY, H5P_DEFAULT
        assert self.fid >= 0, path

    def close(self):
        lib().H5Fclose(self.fid)

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def dataset(self, name):
        h = lib()
        d = h.H5Dopen2(self.fid, name.encode(), 0)
        assert d >= 0, name
        sp = h.H5Dget_space(d)
        nd = h.H5Sget_simple_extent_ndims(sp)
        dims = (hsize_t * nd)()
        h.H5Sget_simple_extent_dims(sp, dims, None)
        out = np.zeros(tuple(int(x) for x in dims), dtype=np.float64)
        assert h.H5Dread(d, _native("H5T_NATIVE_DOUBLE_g"), 0, 0, 0, out.ctypes.data_as(C.c_void_p)) >= 0
        h.H5Sclose(sp)
        h.H5Dclose(d)
        return out

    def attr(self, obj, name):
        """(value, raw length for strings)"""
        h = lib()
        o = h.H5Oopen(self.fid, obj.encode(), 0)
        a = h.H5Aopen(o, name.encode(), 0)
        assert a >= 0, (obj, name)
        t = h.H5Aget_type(a)
        cls = h.H5Tget_class(t)
        if cls == H5T_STRING:
            n = h.H5Tget_size(t)
            buf = C.create_string_buffer(n)
            h.H5Aread(a, t, buf)
            val = (buf.raw.decode(), n, h.H5Tget_strpad(t))
        elif cls == H5T_FLOAT:
            v = C.c_double()
            h.H5Aread(a, _native("H5T_NATIVE_DOUBLE_g"), C.byref(v))
            val = v.value
        else:
            v = C.c_int()
            h.H5Aread(a, _native("H5T_NATIVE_INT_g"), C.byref(v))
            val = v.value
        h.H5Tclose(t)
        h.H5Aclose(a)
        h.H5Oclose(o)
        return val

    def names(self, group="/"):
        h = lib()
        g = h.H5Gopen2(self.fid, group.encode(), 0)
        n = hsize_t()
        h.H5Gget_num_objs(g, C.byref(n))
        out = []
        for i in range(n.value):
            buf = C.create_string_buffer(256)
            h.H5Gget_objname_by_idx(g, i, buf, 256)
            out.append(buf.value.decode())
        h.H5Gclose(g)
        return sorted(out)

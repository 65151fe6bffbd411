"""HDF5 files in the layout of the k-Wave C++ binaries, through the system HDF5 library (ctypes).

The reference writes these files with h5py (kwave/utils/io.py:52-190, ``write_matrix``
and ``write_attributes``); this module calls the HDF5 C library that h5py wraps,
found as ``KW_LIBHDF5`` (a path), ``ctypes.util.find_library("hdf5")`` or a
few standard install locations.  Without it the functions raise ImportError,
as the reference does without h5py.

Layout (the C++ binary's, which is MATLAB's column-major view):

* every variable is a 3-D dataset; a C-ordered array ``a`` with shape
  (X, Y, Z) is stored with HDF5 dimensions [Z, Y, X] holding ``a.T``, a 2-D
  (X, Y) one as [1, Y, X], scalars and vectors of n values as [1, 1, n]
  (1-D arrays keep their order);
* element types are float32 (attribute data_type "float") or uint64
  (data_type "long"), with domain_type "real";
* the file carries the attributes created_by, creation_date,
  file_description, file_type ("input" / "output"), major_version "1" and
  minor_version "2".
"""
from __future__ import annotations

import ctypes as C
import ctypes.util
import datetime
import os
import platform

import numpy as np

_CANDIDATES = ("/opt/conda/lib/libhdf5.so", "/usr/lib/x86_64-linux-gnu/hdf5/serial/libhdf5.so",
               "/usr/lib/x86_64-linux-gnu/libhdf5_serial.so", "/usr/lib64/libhdf5.so", "/usr/local/lib/libhdf5.so")

hid_t = C.c_int64
herr_t = C.c_int
hsize_t = C.c_uint64

H5F_ACC_RDONLY, H5F_ACC_RDWR, H5F_ACC_TRUNC, H5F_ACC_EXCL = 0x0, 0x1, 0x2, 0x4
H5P_DEFAULT = 0
H5S_ALL = 0
H5S_SCALAR = 0
H5T_INTEGER, H5T_FLOAT, H5T_STRING = 0, 1, 3
H5T_STR_NULLPAD = 1

_lib = None


class _GInfo(C.Structure):
    _fields_ = [("storage_type", C.c_int), ("nlinks", hsize_t), ("max_corder", C.c_int64), ("mounted", C.c_uint)]


def lib():
    """The HDF5 C library with the signatures used here (loaded once)."""
    global _lib
    if _lib is not None:
        return _lib
    paths = [os.environ.get("KW_LIBHDF5"), ctypes.util.find_library("hdf5"), *_CANDIDATES]
    err = None
    for p in paths:
        if not p:
            continue
        try:
            L = C.CDLL(p)
            break
        except OSError as e:
            err = e
    else:
        raise ImportError(f"the HDF5 C library (libhdf5) was not found; set KW_LIBHDF5 to its path ({err})")
    sig = {
        "H5open": ([], herr_t), "H5Eset_auto2": ([hid_t, C.c_void_p, C.c_void_p], herr_t),
        "H5Fcreate": ([C.c_char_p, C.c_uint, hid_t, hid_t], hid_t), "H5Fopen": ([C.c_char_p, C.c_uint, hid_t], hid_t),
        "H5Fclose": ([hid_t], herr_t),
        "H5Screate_simple": ([C.c_int, C.POINTER(hsize_t), C.POINTER(hsize_t)], hid_t), "H5Screate": ([C.c_int], hid_t),
        "H5Sclose": ([hid_t], herr_t), "H5Sget_simple_extent_ndims": ([hid_t], C.c_int),
        "H5Sget_simple_extent_dims": ([hid_t, C.POINTER(hsize_t), C.POINTER(hsize_t)], C.c_int),
        "H5Dcreate2": ([hid_t, C.c_char_p, hid_t, hid_t, hid_t, hid_t, hid_t], hid_t),
        "H5Dopen2": ([hid_t, C.c_char_p, hid_t], hid_t), "H5Dclose": ([hid_t], herr_t),
        "H5Dwrite": ([hid_t, hid_t, hid_t, hid_t, hid_t, C.c_void_p], herr_t),
        "H5Dread": ([hid_t, hid_t, hid_t, hid_t, hid_t, C.c_void_p], herr_t),
        "H5Dget_space": ([hid_t], hid_t), "H5Dget_type": ([hid_t], hid_t),
        "H5Tcopy": ([hid_t], hid_t), "H5Tset_size": ([hid_t, C.c_size_t], herr_t), "H5Tget_size": ([hid_t], C.c_size_t),
        "H5Tset_strpad": ([hid_t, C.c_int], herr_t), "H5Tget_class": ([hid_t], C.c_int), "H5Tclose": ([hid_t], herr_t),
        "H5Tget_sign": ([hid_t], C.c_int),
        "H5Acreate2": ([hid_t, C.c_char_p, hid_t, hid_t, hid_t, hid_t], hid_t), "H5Aopen": ([hid_t, C.c_char_p, hid_t], hid_t),
        "H5Awrite": ([hid_t, hid_t, C.c_void_p], herr_t), "H5Aread": ([hid_t, hid_t, C.c_void_p], herr_t),
        "H5Aget_type": ([hid_t], hid_t), "H5Aclose": ([hid_t], herr_t), "H5Aexists": ([hid_t, C.c_char_p], C.c_int),
        "H5Gget_info": ([hid_t, C.POINTER(_GInfo)], herr_t),
        "H5Lget_name_by_idx": ([hid_t, C.c_char_p, C.c_int, C.c_int, hsize_t, C.c_char_p, C.c_size_t, hid_t], C.c_ssize_t),
        "H5Lexists": ([hid_t, C.c_char_p, hid_t], C.c_int),
    }
    for name, (args, ret) in sig.items():
        fn = getattr(L, name)
        fn.argtypes, fn.restype = args, ret
    if L.H5open() < 0:
        raise ImportError("H5open failed")
    L.H5Eset_auto2(0, None, None)  # report failures through return codes, not the error stack printer
    _lib = L
    return L


def _gid(name):
    return hid_t.in_dll(lib(), name).value


def _ok(v, what):
    if v < 0:
        raise OSError(f"HDF5: {what} failed")
    return v


class _File:
    def __init__(self, path, mode):
        L = lib()
        p = os.fsencode(path)
        if mode == "w":
            self.h = _ok(L.H5Fcreate(p, H5F_ACC_EXCL, H5P_DEFAULT, H5P_DEFAULT), f"creating {path}")
        else:
            self.h = _ok(L.H5Fopen(p, H5F_ACC_RDWR if mode == "a" else H5F_ACC_RDONLY, H5P_DEFAULT), f"opening {path}")

    def __enter__(self):
        return self.h

    def __exit__(self, *exc):
        lib().H5Fclose(self.h)


def _str_attr(obj, name, value):
    L = lib()
    data = value.encode() if isinstance(value, str) else bytes(value)
    t = _ok(L.H5Tcopy(_gid("H5T_C_S1_g")), "string type")
    L.H5Tset_size(t, max(1, len(data)))
    L.H5Tset_strpad(t, H5T_STR_NULLPAD)
    s = _ok(L.H5Screate(H5S_SCALAR), "scalar space")
    a = _ok(L.H5Acreate2(obj, name.encode(), t, s, H5P_DEFAULT, H5P_DEFAULT), f"attribute {name}")
    buf = C.create_string_buffer(data, max(1, len(data)))
    _ok(L.H5Awrite(a, t, buf), f"writing attribute {name}")
    L.H5Aclose(a)
    L.H5Sclose(s)
    L.H5Tclose(t)


def _read_str_attr(obj, name):
    L = lib()
    if L.H5Aexists(obj, name.encode()) <= 0:
        return None
    a = _ok(L.H5Aopen(obj, name.encode(), H5P_DEFAULT), f"attribute {name}")
    t = L.H5Aget_type(a)
    n = L.H5Tget_size(t)
    buf = C.create_string_buffer(n + 1)
    _ok(L.H5Aread(a, t, buf), f"reading attribute {name}")
    L.H5Tclose(t)
    L.H5Aclose(a)
    return buf.value.decode(errors="replace")


def kwave_layout(matrix):
    """(HDF5 dimensions, C-contiguous data) of an array in the binaries' layout (see module docstring)."""
    a = np.asarray(matrix)
    if a.ndim == 3:
        return [a.shape[2], a.shape[1], a.shape[0]], np.ascontiguousarray(a.transpose(2, 1, 0))
    if a.ndim == 2:
        return [1, a.shape[1], a.shape[0]], np.ascontiguousarray(a.T)
    if a.ndim <= 1:
        return [1, 1, max(a.size, 1)], np.ascontiguousarray(a.reshape(-1))
    raise ValueError("Input matrix must have 1, 2 or 3 dimensions.")


def write_matrix(filename, matrix, matrix_name):
    """Add one variable (float32 or uint64) to the file, creating the file if needed (utils/io.py:52-190)."""
    L = lib()
    a = np.asarray(matrix)
    if a.dtype == np.float32:
        mem_t, file_t, ctype = _gid("H5T_NATIVE_FLOAT_g"), _gid("H5T_IEEE_F32LE_g"), "float"
    elif a.dtype == np.uint64:
        mem_t, file_t, ctype = _gid("H5T_NATIVE_ULLONG_g"), _gid("H5T_STD_U64LE_g"), "long"
    else:
        raise ValueError("Input matrix must be of type single or uint64.")
    if np.iscomplexobj(a):
        raise NotImplementedError("complex matrices are not written by the solver path")
    dims, data = kwave_layout(a)
    mode = "a" if os.path.exists(filename) else "w"
    with _File(filename, mode) as f:
        if L.H5Lexists(f, matrix_name.encode(), H5P_DEFAULT) > 0:
            raise ValueError(f"{matrix_name!r} already exists in {filename}")
        d3 = (hsize_t * 3)(*dims)
        s = _ok(L.H5Screate_simple(3, d3, None), "dataspace")
        d = _ok(L.H5Dcreate2(f, matrix_name.encode(), file_t, s, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT), f"dataset {matrix_name}")
        _ok(L.H5Dwrite(d, mem_t, H5S_ALL, H5S_ALL, H5P_DEFAULT, data.ctypes.data_as(C.c_void_p)), f"writing {matrix_name}")
        _str_attr(d, "data_type", ctype)
        _str_attr(d, "domain_type", "real")
        L.H5Dclose(d)
        L.H5Sclose(s)


def write_attributes(filename, file_type="input", file_description=None, created_by="kwave_amd (MI355X k-space solver)"):
    """File-level attributes of the binaries' files (utils/io.py:194-243)."""
    if file_description is None:
        file_description = f"{file_type.capitalize()} data created on {platform.system()} by {created_by}"
    attrs = {"major_version": "1", "minor_version": "2", "created_by": created_by, "file_description": file_description,
             "file_type": file_type, "creation_date": datetime.datetime.now().strftime("%d-%b-%Y-%H-%M-%S")}
    with _File(filename, "a") as f:
        for k, v in attrs.items():
            _str_attr(f, k, v)


def read_file(filename):
    """Every variable of a file as {name: array}, with HDF5 dimensions as stored (h5py's view)."""
    L = lib()
    out = {}
    with _File(filename, "r") as f:
        info = _GInfo()
        _ok(L.H5Gget_info(f, C.byref(info)), "group info")
        for i in range(info.nlinks):
            n = L.H5Lget_name_by_idx(f, b".", 0, 0, i, None, 0, H5P_DEFAULT)
            buf = C.create_string_buffer(n + 1)
            L.H5Lget_name_by_idx(f, b".", 0, 0, i, buf, n + 1, H5P_DEFAULT)
            name = buf.value.decode()
            d = _ok(L.H5Dopen2(f, name.encode(), H5P_DEFAULT), f"dataset {name}")
            s = L.H5Dget_space(d)
            nd = L.H5Sget_simple_extent_ndims(s)
            dims = (hsize_t * max(nd, 1))()
            L.H5Sget_simple_extent_dims(s, dims, None)
            shape = tuple(int(dims[k]) for k in range(nd))
            t = L.H5Dget_type(d)
            cls, size, sign = L.H5Tget_class(t), L.H5Tget_size(t), L.H5Tget_sign(t)
            if cls == H5T_FLOAT and size == 4:
                arr, mem_t = np.empty(shape, np.float32), _gid("H5T_NATIVE_FLOAT_g")
            elif cls == H5T_FLOAT and size == 8:
                arr, mem_t = np.empty(shape, np.float64), _gid("H5T_NATIVE_DOUBLE_g")
            elif cls == H5T_INTEGER and size == 8:
                arr, mem_t = (np.empty(shape, np.uint64), _gid("H5T_NATIVE_ULLONG_g")) if sign == 0 else \
                             (np.empty(shape, np.int64), _gid("H5T_NATIVE_LLONG_g"))
            else:
                raise ValueError(f"{name}: unsupported element type (class {cls}, {size} bytes)")
            _ok(L.H5Dread(d, mem_t, H5S_ALL, H5S_ALL, H5P_DEFAULT, arr.ctypes.data_as(C.c_void_p)), f"reading {name}")
            L.H5Tclose(t)
            L.H5Sclose(s)
            L.H5Dclose(d)
            out[name] = arr
    return out


def read_attributes(filename, dataset=None):
    """String attributes of the file (or of one dataset)."""
    L = lib()
    names = ("data_type", "domain_type") if dataset else ("major_version", "minor_version", "created_by",
                                                          "file_description", "file_type", "creation_date")
    with _File(filename, "r") as f:
        obj = L.H5Dopen2(f, dataset.encode(), H5P_DEFAULT) if dataset else f
        try:
            return {n: v for n in names if (v := _read_str_attr(obj, n)) is not None}
        finally:
            if dataset:
                L.H5Dclose(obj)


def from_kwave_layout(arr, ndim):
    """Inverse of kwave_layout for a variable stored from an ndim-dimensional C-ordered array."""
    a = np.asarray(arr)
    if ndim == 3:
        return np.ascontiguousarray(a.transpose(2, 1, 0))
    if ndim == 2:
        return np.ascontiguousarray(a.reshape(a.shape[-2], a.shape[-1]).T)
    return a.reshape(-1)

"""A small HDF5 binding over the system's libhdf5 (C API, through ctypes).

Only what the result files need: files, groups, datasets (optionally gzip-chunked) and
attributes of float64, int64, bool, complex128 and UTF-8 strings, written with the type
conventions h5py uses so that h5py readers see the same values -- numpy bool as an int8
enum {FALSE, TRUE}, complex128 as the compound {r: f64, i: f64}, Python strings as
variable-length UTF-8 strings, numeric data in little-endian IEEE / two's-complement types.

The library is taken from ``AMT_HDF5_LIB``, else the first of the usual names and the conda
location that loads; :func:`available` reports whether one did.
"""

from __future__ import annotations

import ctypes
import ctypes.util
import os
from pathlib import Path

import numpy as np

_hid = ctypes.c_int64
_herr = ctypes.c_int
_hsize = ctypes.c_ulonglong
_SEARCH = ("libhdf5.so.103", "libhdf5.so", "/opt/conda/lib/libhdf5.so.103", "/opt/conda/lib/libhdf5.so")

H5F_ACC_RDONLY, H5F_ACC_TRUNC = 0x0000, 0x0002
H5P_DEFAULT = 0
H5S_ALL = 0
H5S_SCALAR = 0
H5T_VARIABLE = ctypes.c_size_t(-1).value
H5T_CSET_UTF8 = 1
H5T_STR_NULLTERM = 0
H5T_INTEGER, H5T_FLOAT, H5T_STRING, H5T_COMPOUND, H5T_ENUM = 0, 1, 3, 6, 8

_LIB = None
_SIGS = {
    "H5open": (_herr, []),
    "H5Eset_auto2": (_herr, [_hid, ctypes.c_void_p, ctypes.c_void_p]),
    "H5Fcreate": (_hid, [ctypes.c_char_p, ctypes.c_uint, _hid, _hid]),
    "H5Fopen": (_hid, [ctypes.c_char_p, ctypes.c_uint, _hid]),
    "H5Fclose": (_herr, [_hid]),
    "H5Gcreate2": (_hid, [_hid, ctypes.c_char_p, _hid, _hid, _hid]),
    "H5Gopen2": (_hid, [_hid, ctypes.c_char_p, _hid]),
    "H5Gclose": (_herr, [_hid]),
    "H5Lexists": (ctypes.c_int, [_hid, ctypes.c_char_p, _hid]),
    "H5Screate": (_hid, [ctypes.c_int]),
    "H5Screate_simple": (_hid, [ctypes.c_int, ctypes.POINTER(_hsize), ctypes.POINTER(_hsize)]),
    "H5Sget_simple_extent_ndims": (ctypes.c_int, [_hid]),
    "H5Sget_simple_extent_dims": (ctypes.c_int, [_hid, ctypes.POINTER(_hsize), ctypes.POINTER(_hsize)]),
    "H5Sclose": (_herr, [_hid]),
    "H5Pcreate": (_hid, [_hid]),
    "H5Pset_chunk": (_herr, [_hid, ctypes.c_int, ctypes.POINTER(_hsize)]),
    "H5Pset_deflate": (_herr, [_hid, ctypes.c_uint]),
    "H5Pset_create_intermediate_group": (_herr, [_hid, ctypes.c_uint]),
    "H5Pclose": (_herr, [_hid]),
    "H5Dcreate2": (_hid, [_hid, ctypes.c_char_p, _hid, _hid, _hid, _hid, _hid]),
    "H5Dopen2": (_hid, [_hid, ctypes.c_char_p, _hid]),
    "H5Dwrite": (_herr, [_hid, _hid, _hid, _hid, _hid, ctypes.c_void_p]),
    "H5Dread": (_herr, [_hid, _hid, _hid, _hid, _hid, ctypes.c_void_p]),
    "H5Dget_space": (_hid, [_hid]),
    "H5Dget_type": (_hid, [_hid]),
    "H5Dvlen_reclaim": (_herr, [_hid, _hid, _hid, ctypes.c_void_p]),
    "H5Dclose": (_herr, [_hid]),
    "H5Acreate2": (_hid, [_hid, ctypes.c_char_p, _hid, _hid, _hid, _hid]),
    "H5Aopen": (_hid, [_hid, ctypes.c_char_p, _hid]),
    "H5Awrite": (_herr, [_hid, _hid, ctypes.c_void_p]),
    "H5Aread": (_herr, [_hid, _hid, ctypes.c_void_p]),
    "H5Aget_type": (_hid, [_hid]),
    "H5Aget_space": (_hid, [_hid]),
    "H5Aclose": (_herr, [_hid]),
    "H5Adelete": (_herr, [_hid, ctypes.c_char_p]),
    "H5Aexists": (ctypes.c_int, [_hid, ctypes.c_char_p]),
    "H5Tcopy": (_hid, [_hid]),
    "H5Tcreate": (_hid, [ctypes.c_int, ctypes.c_size_t]),
    "H5Tinsert": (_herr, [_hid, ctypes.c_char_p, ctypes.c_size_t, _hid]),
    "H5Tenum_create": (_hid, [_hid]),
    "H5Tenum_insert": (_herr, [_hid, ctypes.c_char_p, ctypes.c_void_p]),
    "H5Tset_size": (_herr, [_hid, ctypes.c_size_t]),
    "H5Tset_cset": (_herr, [_hid, ctypes.c_int]),
    "H5Tset_strpad": (_herr, [_hid, ctypes.c_int]),
    "H5Tget_class": (ctypes.c_int, [_hid]),
    "H5Tget_size": (ctypes.c_size_t, [_hid]),
    "H5Tis_variable_str": (ctypes.c_int, [_hid]),
    "H5Tclose": (_herr, [_hid]),
}
_ATTR_OP = ctypes.CFUNCTYPE(_herr, _hid, ctypes.c_char_p, ctypes.c_void_p, ctypes.c_void_p)


class HDF5Error(RuntimeError):
    pass


def _load():
    global _LIB
    if _LIB is not None:
        return _LIB
    names = [os.environ["AMT_HDF5_LIB"]] if os.environ.get("AMT_HDF5_LIB") else []
    found = ctypes.util.find_library("hdf5")
    names += ([found] if found else []) + list(_SEARCH)
    err = None
    for name in names:
        try:
            lib = ctypes.CDLL(name)
            break
        except OSError as e:
            err = e
    else:
        raise HDF5Error(f"no libhdf5 could be loaded (tried {names}): {err}")
    for fname, (res, args) in _SIGS.items():
        fn = getattr(lib, fname)
        fn.restype, fn.argtypes = res, args
    lib.H5Aiterate2.restype = _herr
    lib.H5Aiterate2.argtypes = [_hid, ctypes.c_int, ctypes.c_int, ctypes.POINTER(_hsize), _ATTR_OP, ctypes.c_void_p]
    if lib.H5open() < 0:
        raise HDF5Error("H5open failed")
    lib.H5Eset_auto2(0, None, None)  # errors are reported through return codes
    _LIB = lib
    return lib


def available() -> bool:
    try:
        _load()
        return True
    except HDF5Error:
        return False


def _g(name: str) -> int:
    return _hid.in_dll(_load(), name).value


def _ok(value, what: str):
    if value < 0:
        raise HDF5Error(f"HDF5 call failed: {what}")
    return value


def _dims(shape):
    return (_hsize * max(1, len(shape)))(*[int(x) for x in shape]) if len(shape) else None


def _space(shape) -> int:
    lib = _load()
    if len(shape) == 0:
        return _ok(lib.H5Screate(H5S_SCALAR), "H5Screate")
    return _ok(lib.H5Screate_simple(len(shape), _dims(shape), None), "H5Screate_simple")


def _string_type() -> int:
    lib = _load()
    t = _ok(lib.H5Tcopy(_g("H5T_C_S1_g")), "H5Tcopy")
    lib.H5Tset_size(t, H5T_VARIABLE)
    lib.H5Tset_cset(t, H5T_CSET_UTF8)
    lib.H5Tset_strpad(t, H5T_STR_NULLTERM)
    return t


def _bool_type() -> int:
    lib = _load()
    t = _ok(lib.H5Tenum_create(_g("H5T_STD_I8LE_g")), "H5Tenum_create")
    for name, v in ((b"FALSE", 0), (b"TRUE", 1)):
        val = ctypes.c_int8(v)
        lib.H5Tenum_insert(t, name, ctypes.byref(val))
    return t


def _complex_type() -> int:
    lib = _load()
    t = _ok(lib.H5Tcreate(H5T_COMPOUND, 16), "H5Tcreate")
    lib.H5Tinsert(t, b"r", 0, _g("H5T_IEEE_F64LE_g"))
    lib.H5Tinsert(t, b"i", 8, _g("H5T_IEEE_F64LE_g"))
    return t


def _prepare(value):
    """(array, file type id, memory type id, owned type ids, keep-alive) for a value."""
    if isinstance(value, (str, bytes)) or (isinstance(value, np.ndarray) and value.dtype.kind in "USO"):
        arr = np.asarray(value, dtype=object)
        enc = [(s.decode() if isinstance(s, bytes) else str(s)).encode("utf-8") for s in arr.ravel()]
        ptrs = (ctypes.c_char_p * max(1, len(enc)))(*enc)
        t = _string_type()
        return arr.shape, ptrs, t, t, [t], enc
    arr = np.asarray(value)
    if arr.dtype == np.bool_:
        data = np.ascontiguousarray(arr.astype(np.int8))
        t = _bool_type()
        return arr.shape, data, t, t, [t], None
    if arr.dtype.kind == "c":
        data = np.ascontiguousarray(arr.astype(np.complex128))
        t = _complex_type()
        return arr.shape, data, t, t, [t], None
    if arr.dtype.kind in "iu":
        data = np.ascontiguousarray(arr.astype(np.int64))
        return arr.shape, data, _g("H5T_STD_I64LE_g"), _g("H5T_NATIVE_LLONG_g"), [], None
    if arr.dtype.kind == "f":
        data = np.ascontiguousarray(arr.astype(np.float64))
        return arr.shape, data, _g("H5T_IEEE_F64LE_g"), _g("H5T_NATIVE_DOUBLE_g"), [], None
    raise TypeError(f"cannot store values of dtype {arr.dtype} in HDF5 here")


def _buffer(data):
    return ctypes.cast(data, ctypes.c_void_p) if isinstance(data, ctypes.Array) else data.ctypes.data_as(ctypes.c_void_p)


def _read(obj_id: int, type_fn, space_fn, read_fn):
    """Read a whole dataset or attribute into numpy / Python values."""
    lib = _load()
    ftype = _ok(type_fn(obj_id), "get_type")
    space = _ok(space_fn(obj_id), "get_space")
    try:
        nd = lib.H5Sget_simple_extent_ndims(space)
        shape = ()
        if nd > 0:
            d = (_hsize * nd)()
            lib.H5Sget_simple_extent_dims(space, d, None)
            shape = tuple(int(x) for x in d)
        count = int(np.prod(shape)) if shape else 1
        cls = lib.H5Tget_class(ftype)
        if cls == H5T_STRING:
            if lib.H5Tis_variable_str(ftype) > 0:
                mem = _string_type()
                ptrs = (ctypes.c_char_p * count)()
                _ok(read_fn(obj_id, mem, ptrs), "read strings")
                out = [p.decode("utf-8") if p is not None else "" for p in ptrs]
                # the library allocated the strings: give them back
                lib.H5Dvlen_reclaim(mem, space, H5P_DEFAULT, ctypes.cast(ptrs, ctypes.c_void_p))
                lib.H5Tclose(mem)
            else:
                size = lib.H5Tget_size(ftype)
                buf = ctypes.create_string_buffer(size * count)
                _ok(read_fn(obj_id, ftype, buf), "read fixed strings")
                raw = buf.raw
                out = [raw[i * size:(i + 1) * size].split(b"\0", 1)[0].decode("utf-8") for i in range(count)]
            return out[0] if not shape else np.array(out, dtype=object).reshape(shape)
        if cls == H5T_ENUM:
            data = np.empty(shape or (1,), dtype=np.int8)
            mem = _bool_type()
            _ok(read_fn(obj_id, mem, data.ctypes.data_as(ctypes.c_void_p)), "read bool")
            lib.H5Tclose(mem)
            data = data.astype(np.bool_)
        elif cls == H5T_COMPOUND:
            data = np.empty(shape or (1,), dtype=np.complex128)
            mem = _complex_type()
            _ok(read_fn(obj_id, mem, data.ctypes.data_as(ctypes.c_void_p)), "read complex")
            lib.H5Tclose(mem)
        elif cls == H5T_INTEGER:
            data = np.empty(shape or (1,), dtype=np.int64)
            _ok(read_fn(obj_id, _g("H5T_NATIVE_LLONG_g"), data.ctypes.data_as(ctypes.c_void_p)), "read int")
        elif cls == H5T_FLOAT:
            data = np.empty(shape or (1,), dtype=np.float64)
            _ok(read_fn(obj_id, _g("H5T_NATIVE_DOUBLE_g"), data.ctypes.data_as(ctypes.c_void_p)), "read float")
        else:
            raise HDF5Error(f"unsupported HDF5 type class {cls}")
        return data.reshape(shape)[()] if not shape else data.reshape(shape)
    finally:
        lib.H5Sclose(space)
        lib.H5Tclose(ftype)


class _Attrs:
    def __init__(self, oid: int):
        self._id = oid

    def __setitem__(self, name: str, value) -> None:
        lib = _load()
        key = name.encode()
        if lib.H5Aexists(self._id, key) > 0:
            lib.H5Adelete(self._id, key)
        shape, data, ftype, mtype, owned, keep = _prepare(value)
        space = _space(shape)
        try:
            aid = _ok(lib.H5Acreate2(self._id, key, ftype, space, H5P_DEFAULT, H5P_DEFAULT), f"attribute {name}")
            try:
                _ok(lib.H5Awrite(aid, mtype, _buffer(data)), f"write attribute {name}")
            finally:
                lib.H5Aclose(aid)
        finally:
            lib.H5Sclose(space)
            for t in owned:
                lib.H5Tclose(t)
        del keep

    def __getitem__(self, name: str):
        lib = _load()
        aid = lib.H5Aopen(self._id, name.encode(), H5P_DEFAULT)
        if aid < 0:
            raise KeyError(name)
        try:
            return _read(aid, lib.H5Aget_type, lib.H5Aget_space, lib.H5Aread)
        finally:
            lib.H5Aclose(aid)

    def __contains__(self, name: str) -> bool:
        return _load().H5Aexists(self._id, name.encode()) > 0

    def keys(self) -> list:
        lib = _load()
        names = []

        def collect(_loc, name, _info, _data):
            names.append(name.decode())
            return 0

        op = _ATTR_OP(collect)
        idx = _hsize(0)
        _ok(lib.H5Aiterate2(self._id, 0, 0, ctypes.byref(idx), op, None), "H5Aiterate2")
        return names

    def to_dict(self) -> dict:
        return {k: self[k] for k in self.keys()}


class Group:
    def __init__(self, oid: int, close):
        self.id = oid
        self._close = close
        self.attrs = _Attrs(oid)

    def close(self) -> None:
        if self.id >= 0:
            self._close(self.id)
            self.id = -1

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __contains__(self, path: str) -> bool:
        lib = _load()
        parts = path.strip("/").split("/")
        for i in range(1, len(parts) + 1):
            if lib.H5Lexists(self.id, "/".join(parts[:i]).encode(), H5P_DEFAULT) <= 0:
                return False
        return True

    def create_group(self, name: str) -> "Group":
        lib = _load()
        gid = _ok(lib.H5Gcreate2(self.id, name.encode(), H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT), f"group {name}")
        return Group(gid, lib.H5Gclose)

    def group(self, name: str) -> "Group":
        lib = _load()
        return Group(_ok(lib.H5Gopen2(self.id, name.encode(), H5P_DEFAULT), f"open group {name}"), lib.H5Gclose)

    def create_dataset(self, name: str, data, *, compression: str | None = None, attrs: dict | None = None) -> None:
        """Write ``data`` as dataset ``name`` (intermediate groups created), optionally gzip
        compressed in chunks of at most ~1 MB, and attach ``attrs``."""
        lib = _load()
        shape, buf, ftype, mtype, owned, keep = _prepare(data)
        space = _space(shape)
        lcpl = _ok(lib.H5Pcreate(_g("H5P_CLS_LINK_CREATE_ID_g")), "H5Pcreate link")
        lib.H5Pset_create_intermediate_group(lcpl, 1)
        dcpl = H5P_DEFAULT
        try:
            if compression == "gzip" and len(shape) > 0 and int(np.prod(shape)) > 0:
                dcpl = _ok(lib.H5Pcreate(_g("H5P_CLS_DATASET_CREATE_ID_g")), "H5Pcreate dataset")
                chunk = list(shape)
                item = 16 if isinstance(buf, np.ndarray) and buf.dtype.kind == "c" else 8
                while int(np.prod(chunk)) * item > (1 << 20) and max(chunk) > 1:
                    axis = int(np.argmax(chunk))
                    chunk[axis] = (chunk[axis] + 1) // 2
                _ok(lib.H5Pset_chunk(dcpl, len(chunk), _dims(chunk)), "H5Pset_chunk")
                _ok(lib.H5Pset_deflate(dcpl, 4), "H5Pset_deflate")
            did = _ok(lib.H5Dcreate2(self.id, name.encode(), ftype, space, lcpl, dcpl, H5P_DEFAULT), f"dataset {name}")
            try:
                _ok(lib.H5Dwrite(did, mtype, H5S_ALL, H5S_ALL, H5P_DEFAULT, _buffer(buf)), f"write {name}")
                for k, v in (attrs or {}).items():
                    _Attrs(did)[k] = v
            finally:
                lib.H5Dclose(did)
        finally:
            if dcpl != H5P_DEFAULT:
                lib.H5Pclose(dcpl)
            lib.H5Pclose(lcpl)
            lib.H5Sclose(space)
            for t in owned:
                lib.H5Tclose(t)
        del keep

    def dataset_attrs(self, name: str) -> dict:
        lib = _load()
        did = _ok(lib.H5Dopen2(self.id, name.encode(), H5P_DEFAULT), f"open {name}")
        try:
            return _Attrs(did).to_dict()
        finally:
            lib.H5Dclose(did)

    def read(self, name: str):
        lib = _load()
        did = _ok(lib.H5Dopen2(self.id, name.encode(), H5P_DEFAULT), f"open {name}")
        try:
            return _read(did, lib.H5Dget_type, lib.H5Dget_space,
                         lambda oid, mem, buf: lib.H5Dread(oid, mem, H5S_ALL, H5S_ALL, H5P_DEFAULT,
                                                           ctypes.cast(buf, ctypes.c_void_p)))
        finally:
            lib.H5Dclose(did)


def create(path) -> Group:
    lib = _load()
    fid = _ok(lib.H5Fcreate(str(Path(path)).encode(), H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT), f"create {path}")
    return Group(fid, lib.H5Fclose)


def open_read(path) -> Group:
    lib = _load()
    fid = _ok(lib.H5Fopen(str(Path(path)).encode(), H5F_ACC_RDONLY, H5P_DEFAULT), f"open {path}")
    return Group(fid, lib.H5Fclose)


__all__ = ["Group", "HDF5Error", "available", "create", "open_read"]

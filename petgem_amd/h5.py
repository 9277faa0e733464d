"""Minimal HDF5 reading and writing through the HDF5 C library (ctypes).

The reference reads receivers and conductivity files with h5py
(``petgem/preprocessing.py:248-254,387-393``) and writes its results file with
h5py (``petgem/postprocessing.py:341-461``).  h5py is not installed for this
interpreter, so this module binds ``libhdf5`` directly; when h5py *is*
importable it is used instead.  Layout conventions follow h5py's so files are
interchangeable: complex datasets are compound ``{r: f8, i: f8}``, strings are
fixed-length byte strings, booleans are int8.
"""
import ctypes
import ctypes.util
import os

import numpy as np

try:  # pragma: no cover - depends on the environment
    import h5py as _h5py
except Exception:  # noqa: BLE001
    _h5py = None

_CANDIDATES = [os.environ.get("PETGEM_HDF5_LIB", ""), "/opt/conda/lib/libhdf5.so",
               ctypes.util.find_library("hdf5") or ""]

H5F_ACC_RDONLY, H5F_ACC_TRUNC = 0, 2
H5P_DEFAULT = 0
H5T_COMPOUND = 6
H5S_SCALAR = 0

_lib = None


class HDF5Unavailable(RuntimeError):
    pass


def _L():
    global _lib
    if _lib is None:
        for c in _CANDIDATES:
            if c and os.path.exists(c):
                try:
                    h = ctypes.CDLL(c)
                    break
                except OSError:
                    continue
        else:
            raise HDF5Unavailable("no HDF5 C library found (set PETGEM_HDF5_LIB) and no h5py")
        h.H5open()
        i64, P = ctypes.c_int64, ctypes.c_void_p
        sig = {
            "H5Fopen": (i64, [ctypes.c_char_p, ctypes.c_uint, i64]),
            "H5Fcreate": (i64, [ctypes.c_char_p, ctypes.c_uint, i64, i64]),
            "H5Fclose": (ctypes.c_int, [i64]),
            "H5Dopen2": (i64, [i64, ctypes.c_char_p, i64]),
            "H5Dcreate2": (i64, [i64, ctypes.c_char_p, i64, i64, i64, i64, i64]),
            "H5Dread": (ctypes.c_int, [i64, i64, i64, i64, i64, P]),
            "H5Dwrite": (ctypes.c_int, [i64, i64, i64, i64, i64, P]),
            "H5Dget_space": (i64, [i64]),
            "H5Dget_type": (i64, [i64]),
            "H5Dclose": (ctypes.c_int, [i64]),
            "H5Gcreate2": (i64, [i64, ctypes.c_char_p, i64, i64, i64]),
            "H5Gclose": (ctypes.c_int, [i64]),
            "H5Screate_simple": (i64, [ctypes.c_int, P, P]),
            "H5Screate": (i64, [ctypes.c_int]),
            "H5Sget_simple_extent_ndims": (ctypes.c_int, [i64]),
            "H5Sget_simple_extent_dims": (ctypes.c_int, [i64, P, P]),
            "H5Sclose": (ctypes.c_int, [i64]),
            "H5Tcreate": (i64, [ctypes.c_int, ctypes.c_size_t]),
            "H5Tinsert": (ctypes.c_int, [i64, ctypes.c_char_p, ctypes.c_size_t, i64]),
            "H5Tcopy": (i64, [i64]),
            "H5Tset_size": (ctypes.c_int, [i64, ctypes.c_size_t]),
            "H5Tget_class": (ctypes.c_int, [i64]),
            "H5Tclose": (ctypes.c_int, [i64]),
        }
        for name, (res, args) in sig.items():
            f = getattr(h, name)
            f.restype, f.argtypes = res, args
        _lib = h
    return _lib


def _gid(name):
    return ctypes.c_int64.in_dll(_L(), name).value


def _ok(v, what):
    if v < 0:
        raise OSError(f"HDF5 call failed: {what}")
    return v


def read_dataset(path, name="data"):
    """Read a numeric dataset as float64 (the receivers / conductivity files)."""
    if _h5py is not None:
        with _h5py.File(path, "r") as f:
            return np.asarray(f[name][()], dtype=np.float64)
    L = _L()
    fid = _ok(L.H5Fopen(path.encode(), H5F_ACC_RDONLY, H5P_DEFAULT), f"open {path}")
    try:
        did = _ok(L.H5Dopen2(fid, name.encode(), H5P_DEFAULT), f"open dataset {name}")
        sid = L.H5Dget_space(did)
        nd = L.H5Sget_simple_extent_ndims(sid)
        dims = (ctypes.c_uint64 * max(nd, 1))()
        L.H5Sget_simple_extent_dims(sid, dims, None)
        shape = tuple(int(d) for d in dims[:nd])
        out = np.empty(shape, dtype=np.float64)
        _ok(L.H5Dread(did, _gid("H5T_NATIVE_DOUBLE_g"), 0, 0, H5P_DEFAULT,
                      out.ctypes.data_as(ctypes.c_void_p)), "read")
        L.H5Sclose(sid)
        L.H5Dclose(did)
    finally:
        L.H5Fclose(fid)
    return out


class Writer:
    """Write groups and datasets into a new HDF5 file (h5py-compatible layout)."""

    def __init__(self, path):
        self.path = path
        if _h5py is not None:
            self._f = _h5py.File(path, "w")
            self._fid = None
        else:
            self._f = None
            self._fid = _ok(_L().H5Fcreate(path.encode(), H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT),
                            f"create {path}")
        self._groups = {}

    def _parent(self, name):
        head, _, leaf = name.rpartition("/")
        return head, leaf

    def group(self, name):
        if self._f is not None:
            self._f.require_group(name)
            return
        parent = self._fid
        acc = ""
        for part in name.strip("/").split("/"):
            acc = f"{acc}/{part}" if acc else part
            if acc not in self._groups:
                self._groups[acc] = _ok(_L().H5Gcreate2(parent, part.encode(), H5P_DEFAULT,
                                                        H5P_DEFAULT, H5P_DEFAULT), f"group {acc}")
            parent = self._groups[acc]

    def dataset(self, name, value):
        if self._f is not None:
            self._f.create_dataset(name, data=value)
            return
        L = _L()
        head, leaf = self._parent(name)
        if head:
            self.group(head)
        loc = self._groups[head] if head else self._fid
        tid_close = False
        if isinstance(value, (str, bytes)):
            raw = value.encode() if isinstance(value, str) else value
            arr = np.frombuffer(raw + b"\0", dtype=np.uint8)
            tid = L.H5Tcopy(_gid("H5T_C_S1_g"))
            L.H5Tset_size(tid, max(1, len(raw)))
            tid_close = True
            buf = np.frombuffer(raw if raw else b"\0", dtype=np.uint8).copy()
            sid = L.H5Screate(H5S_SCALAR)
            mem = tid
            data = buf
            del arr
        else:
            a = np.asarray(value)
            if a.dtype == np.bool_:
                a = a.astype(np.int8)
            if np.iscomplexobj(a):
                a = np.array(a, dtype=np.complex128, order="C")
                tid = L.H5Tcreate(H5T_COMPOUND, 16)
                L.H5Tinsert(tid, b"r", 0, _gid("H5T_NATIVE_DOUBLE_g"))
                L.H5Tinsert(tid, b"i", 8, _gid("H5T_NATIVE_DOUBLE_g"))
                tid_close = True
            elif np.issubdtype(a.dtype, np.integer):
                a = np.array(a, dtype=np.int64, order="C")
                tid = _gid("H5T_NATIVE_LLONG_g")
            else:
                a = np.array(a, dtype=np.float64, order="C")
                tid = _gid("H5T_NATIVE_DOUBLE_g")
            if a.ndim == 0:
                sid = L.H5Screate(H5S_SCALAR)
            else:
                dims = (ctypes.c_uint64 * a.ndim)(*a.shape)
                sid = L.H5Screate_simple(a.ndim, dims, None)
            mem = tid
            data = a
        did = _ok(L.H5Dcreate2(loc, leaf.encode(), tid, sid, H5P_DEFAULT, H5P_DEFAULT,
                               H5P_DEFAULT), f"dataset {name}")
        _ok(L.H5Dwrite(did, mem, 0, 0, H5P_DEFAULT, data.ctypes.data_as(ctypes.c_void_p)),
            f"write {name}")
        L.H5Dclose(did)
        L.H5Sclose(sid)
        if tid_close:
            L.H5Tclose(tid)

    def close(self):
        if self._f is not None:
            self._f.close()
            return
        L = _L()
        for g in sorted(self._groups, key=len, reverse=True):
            L.H5Gclose(self._groups[g])
        L.H5Fclose(self._fid)
        self._groups = {}

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


def write_data(path, array, name="data"):
    """Write one float64 dataset (the layout of the receivers / conductivity files)."""
    with Writer(path) as w:
        w.dataset(name, np.asarray(array, dtype=np.float64))

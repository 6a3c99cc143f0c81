"""Reading H5MD files back in tests, through the HDF5 C library (ctypes).

Test helper only: the product writes files (dmd_amd.trajectory) and never
reads them."""
import ctypes as C
import os

import numpy as np

HDF5_LIB = os.path.join(os.environ.get("DMDX_HDF5_ROOT", "/opt/conda"), "lib", "libhdf5.so.103")

_h = None
_hid = C.c_int64  # hid_t (HDF5 1.10)
H5P_DEFAULT = 0
H5S_ALL = 0
H5F_ACC_RDONLY = 0


def available() -> bool:
    return os.path.exists(HDF5_LIB)


def _lib():
    global _h
    if _h is None:
        h = C.CDLL(HDF5_LIB)
        h.H5open()
        for f in ("H5Fopen",):
            getattr(h, f).restype = _hid
        h.H5Fopen.argtypes = [C.c_char_p, C.c_uint, _hid]
        h.H5Dopen2.restype = _hid
        h.H5Dopen2.argtypes = [_hid, C.c_char_p, _hid]
        h.H5Dget_space.restype = _hid
        h.H5Dget_space.argtypes = [_hid]
        h.H5Dget_type.restype = _hid
        h.H5Dget_type.argtypes = [_hid]
        h.H5Dget_create_plist.restype = _hid
        h.H5Dget_create_plist.argtypes = [_hid]
        h.H5Sget_simple_extent_ndims.argtypes = [_hid]
        h.H5Sget_simple_extent_dims.argtypes = [_hid, C.POINTER(C.c_uint64),
                                                C.POINTER(C.c_uint64)]
        h.H5Pget_chunk.argtypes = [_hid, C.c_int, C.POINTER(C.c_uint64)]
        h.H5Tget_class.argtypes = [_hid]
        h.H5Tget_size.argtypes = [_hid]
        h.H5Tget_size.restype = C.c_size_t
        h.H5Dread.argtypes = [_hid, _hid, _hid, _hid, _hid, C.c_void_p]
        h.H5Lexists.argtypes = [_hid, C.c_char_p, _hid]
        for f in ("H5Dclose", "H5Sclose", "H5Tclose", "H5Pclose", "H5Fclose"):
            getattr(h, f).argtypes = [_hid]
        _h = h
    return _h


class File:
    """Read-only view of an HDF5 file."""

    def __init__(self, path):
        h = _lib()
        self.f = h.H5Fopen(os.fsencode(path), H5F_ACC_RDONLY, H5P_DEFAULT)
        if self.f < 0:
            raise OSError(f"cannot open {path}")

    def close(self):
        if self.f >= 0:
            _lib().H5Fclose(self.f)
            self.f = -1

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def exists(self, name: str) -> bool:
        """Every link along the path exists."""
        h = _lib()
        parts = [p for p in name.split("/") if p]
        for k in range(1, len(parts) + 1):
            if h.H5Lexists(self.f, ("/" + "/".join(parts[:k])).encode(), H5P_DEFAULT) <= 0:
                return False
        return True

    def info(self, name: str) -> dict:
        """shape, maxshape (-1 = unlimited), chunk shape, type class and size."""
        h = _lib()
        ds = h.H5Dopen2(self.f, name.encode(), H5P_DEFAULT)
        if ds < 0:
            raise KeyError(name)
        sp = h.H5Dget_space(ds)
        nd = h.H5Sget_simple_extent_ndims(sp)
        dims = (C.c_uint64 * nd)()
        maxd = (C.c_uint64 * nd)()
        h.H5Sget_simple_extent_dims(sp, dims, maxd)
        pl = h.H5Dget_create_plist(ds)
        ch = (C.c_uint64 * nd)()
        h.H5Pget_chunk(pl, nd, ch)
        ty = h.H5Dget_type(ds)
        cls, size = h.H5Tget_class(ty), h.H5Tget_size(ty)
        for f, x in ((h.H5Tclose, ty), (h.H5Pclose, pl), (h.H5Sclose, sp), (h.H5Dclose, ds)):
            f(x)
        unl = (1 << 64) - 1
        return {"shape": tuple(int(v) for v in dims),
                "maxshape": tuple(-1 if v == unl else int(v) for v in maxd),
                "chunk": tuple(int(v) for v in ch),
                "class": {0: "int", 1: "float"}.get(cls, cls), "size": int(size)}

    def read(self, name: str) -> np.ndarray:
        h = _lib()
        inf = self.info(name)
        if inf["class"] == "float":
            out = np.empty(inf["shape"], np.float64)
            mem = C.c_int64.in_dll(h, "H5T_NATIVE_DOUBLE_g").value
        else:
            out = np.empty(inf["shape"], np.int32)
            mem = C.c_int64.in_dll(h, "H5T_NATIVE_INT_g").value
        ds = h.H5Dopen2(self.f, name.encode(), H5P_DEFAULT)
        if out.size:
            rc = h.H5Dread(ds, mem, H5S_ALL, H5S_ALL, H5P_DEFAULT, out.ctypes.data)
            if rc < 0:
                h.H5Dclose(ds)
                raise OSError(f"cannot read {name}")
        h.H5Dclose(ds)
        return out

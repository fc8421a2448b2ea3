"""Minimal HDF5 file access over libhdf5's C API (ctypes), for checkpoints.

The reference writes a walker's checkpoint through HDF5.jl into an HDF5 group
(src/sse.jl:89-107); this module gives the Python host side the same file
format: groups, scalar datasets (H5S_SCALAR, as HDF5.jl writes Julia scalars),
n-dimensional numeric datasets and one-member compound datasets (HDF5.jl's
mapping of an isbits struct such as OperCode, whose only field is `code`).

The library is found through $SSE_HDF5_LIB, the loader's search path, or the
usual install locations; it is loaded on first use, so the rest of the package
never depends on it.  Every failing HDF5 call raises H5Error.
"""
import ctypes as C
import ctypes.util
import os

import numpy as np

hid_t = C.c_int64
herr_t = C.c_int
htri_t = C.c_int
hsize_t = C.c_ulonglong

H5F_ACC_RDONLY = 0x0000
H5F_ACC_RDWR = 0x0001
H5F_ACC_TRUNC = 0x0002
H5P_DEFAULT = 0
H5S_ALL = 0
H5S_SCALAR = 0
H5S_SIMPLE = 1
H5T_INTEGER = 0
H5T_FLOAT = 1
H5T_COMPOUND = 6
H5T_SGN_NONE = 0

_CANDIDATES = ("libhdf5.so", "/opt/conda/lib/libhdf5.so.103", "/opt/conda/lib/libhdf5.so",
               "/usr/lib/x86_64-linux-gnu/hdf5/serial/libhdf5.so",
               "/usr/lib/x86_64-linux-gnu/libhdf5_serial.so")


class H5Error(RuntimeError):
    pass


_lib = None
_native = {}


def _sig(lib, name, restype, *argtypes):
    f = getattr(lib, name)
    f.restype = restype
    f.argtypes = list(argtypes)


def library():
    """Load libhdf5 once and declare the calls this module uses."""
    global _lib
    if _lib is not None:
        return _lib
    names = [os.environ.get("SSE_HDF5_LIB"), ctypes.util.find_library("hdf5")] + list(_CANDIDATES)
    errors = []
    for nm in names:
        if not nm:
            continue
        try:
            lib = C.CDLL(nm)
            break
        except OSError as e:
            errors.append(f"{nm}: {e}")
    else:
        raise H5Error("libhdf5 not found (set SSE_HDF5_LIB): " + "; ".join(errors))
    vp, cp = C.c_void_p, C.c_char_p
    _sig(lib, "H5open", herr_t)
    _sig(lib, "H5get_libversion", herr_t, C.POINTER(C.c_uint), C.POINTER(C.c_uint),
         C.POINTER(C.c_uint))
    _sig(lib, "H5Eset_auto2", herr_t, hid_t, vp, vp)
    _sig(lib, "H5free_memory", herr_t, vp)
    _sig(lib, "H5Fcreate", hid_t, cp, C.c_uint, hid_t, hid_t)
    _sig(lib, "H5Fopen", hid_t, cp, C.c_uint, hid_t)
    _sig(lib, "H5Fclose", herr_t, hid_t)
    _sig(lib, "H5Gcreate2", hid_t, hid_t, cp, hid_t, hid_t, hid_t)
    _sig(lib, "H5Gopen2", hid_t, hid_t, cp, hid_t)
    _sig(lib, "H5Gclose", herr_t, hid_t)
    _sig(lib, "H5Lexists", htri_t, hid_t, cp, hid_t)
    _sig(lib, "H5Ldelete", herr_t, hid_t, cp, hid_t)
    _sig(lib, "H5Screate", hid_t, C.c_int)
    _sig(lib, "H5Screate_simple", hid_t, C.c_int, C.POINTER(hsize_t), C.POINTER(hsize_t))
    _sig(lib, "H5Sclose", herr_t, hid_t)
    _sig(lib, "H5Sget_simple_extent_type", C.c_int, hid_t)
    _sig(lib, "H5Sget_simple_extent_ndims", C.c_int, hid_t)
    _sig(lib, "H5Sget_simple_extent_dims", C.c_int, hid_t, C.POINTER(hsize_t), C.POINTER(hsize_t))
    _sig(lib, "H5Dcreate2", hid_t, hid_t, cp, hid_t, hid_t, hid_t, hid_t, hid_t)
    _sig(lib, "H5Dopen2", hid_t, hid_t, cp, hid_t)
    _sig(lib, "H5Dclose", herr_t, hid_t)
    _sig(lib, "H5Dwrite", herr_t, hid_t, hid_t, hid_t, hid_t, hid_t, vp)
    _sig(lib, "H5Dread", herr_t, hid_t, hid_t, hid_t, hid_t, hid_t, vp)
    _sig(lib, "H5Dget_space", hid_t, hid_t)
    _sig(lib, "H5Dget_type", hid_t, hid_t)
    _sig(lib, "H5Tcreate", hid_t, C.c_int, C.c_size_t)
    _sig(lib, "H5Tinsert", herr_t, hid_t, cp, C.c_size_t, hid_t)
    _sig(lib, "H5Tclose", herr_t, hid_t)
    _sig(lib, "H5Tget_class", C.c_int, hid_t)
    _sig(lib, "H5Tget_size", C.c_size_t, hid_t)
    _sig(lib, "H5Tget_sign", C.c_int, hid_t)
    _sig(lib, "H5Tget_nmembers", C.c_int, hid_t)
    _sig(lib, "H5Tget_member_type", hid_t, hid_t, C.c_uint)
    _sig(lib, "H5Tget_member_name", vp, hid_t, C.c_uint)
    if lib.H5open() < 0:
        raise H5Error("H5open failed")
    lib.H5Eset_auto2(0, None, None)  # errors come back as return codes only
    for key, sym in (("u1", "H5T_NATIVE_UINT8_g"), ("i1", "H5T_NATIVE_INT8_g"),
                     ("u2", "H5T_NATIVE_UINT16_g"), ("i2", "H5T_NATIVE_INT16_g"),
                     ("u4", "H5T_NATIVE_UINT32_g"), ("i4", "H5T_NATIVE_INT32_g"),
                     ("u8", "H5T_NATIVE_UINT64_g"), ("i8", "H5T_NATIVE_INT64_g"),
                     ("f4", "H5T_NATIVE_FLOAT_g"), ("f8", "H5T_NATIVE_DOUBLE_g")):
        _native[key] = hid_t.in_dll(lib, sym).value
    _lib = lib
    return lib


def available():
    try:
        library()
        return True
    except H5Error:
        return False


def version():
    lib = library()
    v = [C.c_uint(), C.c_uint(), C.c_uint()]
    lib.H5get_libversion(*[C.byref(x) for x in v])
    return tuple(x.value for x in v)


def _ok(ret, what):
    if ret < 0:
        raise H5Error(what)
    return ret


def _native_for(dtype):
    dt = np.dtype(dtype)
    key = f"{dt.kind}{dt.itemsize}"
    if key not in _native:
        raise H5Error(f"no HDF5 native type for numpy dtype {dt}")
    return _native[key]


def _dtype_of(lib, tid):
    """numpy dtype of an HDF5 integer/float datatype."""
    cls = lib.H5Tget_class(tid)
    size = lib.H5Tget_size(tid)
    if cls == H5T_INTEGER:
        signed = lib.H5Tget_sign(tid) != H5T_SGN_NONE
        return np.dtype(f"{'i' if signed else 'u'}{size}")
    if cls == H5T_FLOAT and size in (4, 8):
        return np.dtype(f"f{size}")
    raise H5Error(f"unsupported HDF5 datatype class {cls} (size {size})")


class Group:
    """An open HDF5 group (the file's root group included)."""

    def __init__(self, hid, owned=True):
        self.hid = hid
        self._owned = owned

    def close(self):
        if self._owned and self.hid is not None:
            library().H5Gclose(self.hid)
        self.hid = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __contains__(self, name):
        lib = library()
        return _ok(lib.H5Lexists(self.hid, name.encode(), H5P_DEFAULT), f"H5Lexists {name}") > 0

    def create_group(self, name):
        """Create the group `name` (or open it if it exists)."""
        lib = library()
        if name in self:
            return self.group(name)
        return Group(_ok(lib.H5Gcreate2(self.hid, name.encode(), H5P_DEFAULT, H5P_DEFAULT,
                                        H5P_DEFAULT), f"create group {name}"))

    def group(self, name):
        lib = library()
        if name not in self:
            raise KeyError(name)
        return Group(_ok(lib.H5Gopen2(self.hid, name.encode(), H5P_DEFAULT), f"open group {name}"))

    def write(self, name, value, compound_member=None):
        """Write a dataset, replacing one of the same name.

        Python ints become int64 scalars and floats float64 scalars (scalar
        dataspace); numpy arrays keep their dtype and shape (C order).  With
        `compound_member`, an integer array is stored as a compound type with
        that one member."""
        lib = library()
        if isinstance(value, (bool, np.bool_)):
            raise H5Error("booleans are not stored")
        if isinstance(value, (int, np.integer)) and not isinstance(value, np.ndarray):
            arr = np.array(value, dtype=np.int64 if isinstance(value, int) else value.dtype)
            scalar = True
        elif isinstance(value, (float, np.floating)):
            arr = np.array(value, dtype=np.float64)
            scalar = True
        else:
            arr = np.asarray(value)
            scalar = arr.ndim == 0  # (ascontiguousarray would make it 1-d)
            if not scalar:
                arr = np.ascontiguousarray(arr)
        if name in self:
            _ok(lib.H5Ldelete(self.hid, name.encode(), H5P_DEFAULT), f"replace {name}")
        if scalar:
            space = _ok(lib.H5Screate(H5S_SCALAR), "H5Screate")
        else:
            dims = (hsize_t * max(1, arr.ndim))(*arr.shape)
            space = _ok(lib.H5Screate_simple(arr.ndim, dims, None), "H5Screate_simple")
        mtype = _native_for(arr.dtype)
        ctype = None
        try:
            if compound_member is not None:
                if arr.dtype.kind not in "iu":
                    raise H5Error("compound datasets hold one integer member here")
                ctype = _ok(lib.H5Tcreate(H5T_COMPOUND, arr.dtype.itemsize), "H5Tcreate")
                _ok(lib.H5Tinsert(ctype, compound_member.encode(), 0, mtype), "H5Tinsert")
                mtype = ctype
            dset = _ok(lib.H5Dcreate2(self.hid, name.encode(), mtype, space, H5P_DEFAULT,
                                      H5P_DEFAULT, H5P_DEFAULT), f"create dataset {name}")
            try:
                _ok(lib.H5Dwrite(dset, mtype, H5S_ALL, H5S_ALL, H5P_DEFAULT,
                                 arr.ctypes.data_as(C.c_void_p)), f"write {name}")
            finally:
                lib.H5Dclose(dset)
        finally:
            if ctype is not None:
                lib.H5Tclose(ctype)
            lib.H5Sclose(space)

    def read(self, name):
        """Read a dataset: a Python scalar for a scalar dataspace, else a numpy
        array.  A compound dataset with one integer member reads as an array of
        that member."""
        lib = library()
        if name not in self:
            raise KeyError(name)
        dset = _ok(lib.H5Dopen2(self.hid, name.encode(), H5P_DEFAULT), f"open dataset {name}")
        ftype = space = mtype = None
        try:
            ftype = _ok(lib.H5Dget_type(dset), "H5Dget_type")
            space = _ok(lib.H5Dget_space(dset), "H5Dget_space")
            cls = lib.H5Tget_class(ftype)
            member = None
            if cls == H5T_COMPOUND:
                if lib.H5Tget_nmembers(ftype) != 1:
                    raise H5Error(f"{name}: compound datasets with one member only")
                sub = _ok(lib.H5Tget_member_type(ftype, 0), "member type")
                try:
                    dt = _dtype_of(lib, sub)
                finally:
                    lib.H5Tclose(sub)
                p = lib.H5Tget_member_name(ftype, 0)
                member = C.cast(p, C.c_char_p).value
                lib.H5free_memory(p)
                mtype = _ok(lib.H5Tcreate(H5T_COMPOUND, dt.itemsize), "H5Tcreate")
                _ok(lib.H5Tinsert(mtype, member, 0, _native_for(dt)), "H5Tinsert")
                rtype = mtype
            else:
                dt = _dtype_of(lib, ftype)
                rtype = _native_for(dt)
            kind = lib.H5Sget_simple_extent_type(space)
            if kind == H5S_SCALAR:
                shape = ()
            elif kind == H5S_SIMPLE:
                nd = lib.H5Sget_simple_extent_ndims(space)
                dims = (hsize_t * max(1, nd))()
                _ok(lib.H5Sget_simple_extent_dims(space, dims, None), "extent dims")
                shape = tuple(int(dims[i]) for i in range(nd))
            else:
                raise H5Error(f"{name}: empty (null) dataspace")
            out = np.empty(shape, dtype=dt)
            if out.size:
                _ok(lib.H5Dread(dset, rtype, H5S_ALL, H5S_ALL, H5P_DEFAULT,
                                out.ctypes.data_as(C.c_void_p)), f"read {name}")
            return out.item() if shape == () else out
        finally:
            if mtype is not None:
                lib.H5Tclose(mtype)
            if space is not None:
                lib.H5Sclose(space)
            if ftype is not None:
                lib.H5Tclose(ftype)
            lib.H5Dclose(dset)

    def dataset_info(self, name):
        """(class, element size, compound member name or None, shape) of a dataset:
        lets tests check the on-disk layout."""
        lib = library()
        dset = _ok(lib.H5Dopen2(self.hid, name.encode(), H5P_DEFAULT), f"open dataset {name}")
        try:
            ftype = lib.H5Dget_type(dset)
            space = lib.H5Dget_space(dset)
            cls, size = lib.H5Tget_class(ftype), lib.H5Tget_size(ftype)
            member = None
            if cls == H5T_COMPOUND and lib.H5Tget_nmembers(ftype) >= 1:
                p = lib.H5Tget_member_name(ftype, 0)
                member = C.cast(p, C.c_char_p).value.decode()
                lib.H5free_memory(p)
            if lib.H5Sget_simple_extent_type(space) == H5S_SCALAR:
                shape = ()
            else:
                nd = lib.H5Sget_simple_extent_ndims(space)
                dims = (hsize_t * max(1, nd))()
                lib.H5Sget_simple_extent_dims(space, dims, None)
                shape = tuple(int(dims[i]) for i in range(nd))
            lib.H5Tclose(ftype)
            lib.H5Sclose(space)
            return cls, size, member, shape
        finally:
            lib.H5Dclose(dset)


class File(Group):
    """An HDF5 file; mode "r" (read), "w" (create or truncate), "a" (read and
    write, created when missing).  Acts as its root group."""

    def __init__(self, path, mode="r"):
        lib = library()
        p = os.fspath(path).encode()
        if mode == "w" or (mode == "a" and not os.path.exists(path)):
            fid = lib.H5Fcreate(p, H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT)
        elif mode in ("r", "a"):
            fid = lib.H5Fopen(p, H5F_ACC_RDONLY if mode == "r" else H5F_ACC_RDWR, H5P_DEFAULT)
        else:
            raise ValueError(f"mode {mode!r}")
        self.fid = _ok(fid, f"open {path}")
        root = _ok(lib.H5Gopen2(self.fid, b"/", H5P_DEFAULT), "open root group")
        super().__init__(root)

    def close(self):
        super().close()
        if getattr(self, "fid", None) is not None:
            library().H5Fclose(self.fid)
            self.fid = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

"""TEST INFRASTRUCTURE ONLY -- ctypes driver of the CPU oracle.

The oracle (siddon_oracle.c) restates tomograpy/C_siddon.c.template in plain
C/OpenMP.  This module feeds it numpy arrays and the same per-view header
records the product library consumes.  Only tests/, __graft_entry__.smoke()
and bench.py's cpu_baseline leg may import it; the product never does.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libsiddon_oracle.so")

PJ, BPJ, PJT, BPJT = 0, 1, 2, 3

_i32 = ctypes.c_int32
_i64 = ctypes.c_int64
_dbl = ctypes.c_double
_vp = ctypes.c_void_p


class Map(ctypes.Structure):
    _fields_ = [("cdelt", _dbl * 3), ("pshape", _dbl * 3), ("mmin", _dbl * 3),
                ("mmax", _dbl * 3), ("n", _i32 * 3), ("pad", _i32), ("s", _i64 * 4)]


class Data(ctypes.Structure):
    _fields_ = [("n1", _i32), ("n2", _i32), ("nv", _i32), ("pad", _i32),
                ("s", _i64 * 3), ("ms", _i64 * 3), ("hdr", _vp), ("mask", _vp),
                ("tindex", _vp)]


_lib = None


def build():
    """Compiles the oracle with its Makefile (gcc, OpenMP)."""
    subprocess.run(["make", "-C", HERE, "-s"], check=True)
    return LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        l = ctypes.CDLL(LIB_PATH)
        for sfx in ("f32", "f64"):
            f = getattr(l, "tgo_siddon_" + sfx)
            f.restype = ctypes.c_int
            f.argtypes = [ctypes.POINTER(Map), ctypes.POINTER(Data), _vp, _vp, ctypes.c_int,
                          ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                          ctypes.POINTER(ctypes.c_longlong)]
            f = getattr(l, "tgo_ray_trace_" + sfx)
            f.restype = ctypes.c_longlong
            f.argtypes = [ctypes.POINTER(Map), _vp, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                          _vp, _vp, _vp, ctypes.c_longlong]
            f = getattr(l, "tgo_full_unit_vector_" + sfx)
            f.restype = None
            f.argtypes = [ctypes.POINTER(Data), _vp, _vp]
            f = getattr(l, "tgo_full_intersection_parameters_" + sfx)
            f.restype = None
            f.argtypes = [ctypes.POINTER(Map), ctypes.POINTER(Data), _vp, _vp, _vp]
            f = getattr(l, "tgo_ray_steps_" + sfx)
            f.restype = None
            f.argtypes = [ctypes.POINTER(Map), ctypes.POINTER(Data), ctypes.c_int, _vp]
        l.tgo_num_threads.restype = ctypes.c_int
        _lib = l
    return _lib


def _sfx(dtype):
    name = np.dtype(dtype).name
    if name == "float32":
        return "f32"
    if name == "float64":
        return "f64"
    raise KeyError(name)


def make_map(map_rec, shape, strides_elems=None, nt=1):
    """Map struct from the 12-value record (CDELT, PSHAPE, MMIN, MMAX) and shape."""
    m = Map()
    for k in range(3):
        m.cdelt[k] = map_rec[k]
        m.pshape[k] = map_rec[3 + k]
        m.mmin[k] = map_rec[6 + k]
        m.mmax[k] = map_rec[9 + k]
        m.n[k] = int(shape[k])
    if strides_elems is not None:
        for k, s in enumerate(strides_elems):
            m.s[k] = int(s)
    return m


def _elem_strides(a):
    return [s // a.itemsize for s in a.strides]


def siddon(kind, records, map_rec, cube, data, sun=False, nthread=0, t0=0, t1=-1, mask=None,
           tindex=None):
    """Runs the oracle projector (kind PJ/PJT) or backprojector (BPJ/BPJT) in place.

    records: (nv, 18) float64 view records; cube/data: numpy arrays of the same
    float dtype with any strides.  Returns the number of Siddon steps taken.
    """
    assert cube.dtype == data.dtype
    records = np.ascontiguousarray(records, dtype=np.float64)
    m = make_map(map_rec, cube.shape, _elem_strides(cube))
    d = Data()
    d.n1, d.n2, d.nv = data.shape
    for k, s in enumerate(_elem_strides(data)):
        d.s[k] = s
    d.hdr = records.ctypes.data
    keep = [records]
    if mask is not None:
        mk = (np.asarray(mask) != 0).astype(np.uint8)
        keep.append(mk)
        d.mask = mk.ctypes.data
        for k, s in enumerate(_elem_strides(mk)):
            d.ms[k] = s
    if tindex is not None:
        ti = np.ascontiguousarray(tindex, dtype=np.int32)
        keep.append(ti)
        d.tindex = ti.ctypes.data
    steps = ctypes.c_longlong(0)
    f = getattr(lib(), "tgo_siddon_" + _sfx(data.dtype))
    rc = f(ctypes.byref(m), ctypes.byref(d), cube.ctypes.data, data.ctypes.data, int(kind),
           int(bool(sun)), int(nthread), int(t0), int(t1), ctypes.byref(steps))
    if rc != 0:
        raise RuntimeError("oracle returned %d" % rc)
    return steps.value


def ray_trace(map_rec, shape, record, i, j, sun=False, dtype=np.float64, cap=None):
    """Voxel indices (k, 3), lengths (k,) and unit vector of one ray."""
    cap = cap or int(sum(shape)) + 8
    m = make_map(map_rec, shape)
    rec = np.ascontiguousarray(record, dtype=np.float64)
    iv = np.zeros((cap, 3), dtype=np.int32)
    ln = np.zeros(cap, dtype=dtype)
    u = np.zeros(3, dtype=dtype)
    f = getattr(lib(), "tgo_ray_trace_" + _sfx(dtype))
    n = f(ctypes.byref(m), rec.ctypes.data, int(i), int(j), int(bool(sun)), iv.ctypes.data,
          ln.ctypes.data, u.ctypes.data, cap)
    if n < 0:
        raise RuntimeError("NaN ray state")
    n = min(n, cap)
    return iv[:n].copy(), ln[:n].copy(), u


def _data_struct(records, n1, n2, mask=None):
    d = Data()
    d.n1, d.n2, d.nv = n1, n2, records.shape[0]
    d.s[0], d.s[1], d.s[2] = n2 * records.shape[0], records.shape[0], 1
    d.hdr = records.ctypes.data
    return d


def full_unit_vector(records, n1, n2, dtype=np.float64, data=None):
    """u of shape (n1, n2, nv, 3), skipping NaN data pixels when data is given."""
    records = np.ascontiguousarray(records, dtype=np.float64)
    nv = records.shape[0]
    d = _data_struct(records, n1, n2)
    dd = None
    if data is not None:
        dd = np.ascontiguousarray(data, dtype=dtype)
    u = np.zeros((n1, n2, nv, 3), dtype=dtype)
    getattr(lib(), "tgo_full_unit_vector_" + _sfx(dtype))(
        ctypes.byref(d), None if dd is None else dd.ctypes.data, u.ctypes.data)
    return u


def full_intersection_parameters(map_rec, records, u, dtype=np.float64):
    records = np.ascontiguousarray(records, dtype=np.float64)
    u = np.ascontiguousarray(u, dtype=dtype)
    n1, n2, nv, _ = u.shape
    d = _data_struct(records, n1, n2)
    m = make_map(map_rec, (1, 1, 1))
    a1 = np.zeros_like(u)
    an = np.zeros_like(u)
    getattr(lib(), "tgo_full_intersection_parameters_" + _sfx(dtype))(
        ctypes.byref(m), ctypes.byref(d), u.ctypes.data, a1.ctypes.data, an.ctypes.data)
    return a1, an


def ray_steps(map_rec, shape, records, n1, n2, sun=False, dtype=np.float64, mask=None):
    """Per-ray step counts, shape (nv, n1, n2) uint32."""
    records = np.ascontiguousarray(records, dtype=np.float64)
    m = make_map(map_rec, shape)
    d = _data_struct(records, n1, n2)
    keep = None
    if mask is not None:
        keep = (np.asarray(mask) != 0).astype(np.uint8)
        d.mask = keep.ctypes.data
        for k, s in enumerate(_elem_strides(keep)):
            d.ms[k] = s
    out = np.zeros((records.shape[0], n1, n2), dtype=np.uint32)
    getattr(lib(), "tgo_ray_steps_" + _sfx(dtype))(ctypes.byref(m), ctypes.byref(d),
                                                   int(bool(sun)), out.ctypes.data)
    return out


def num_threads():
    return lib().tgo_num_threads()

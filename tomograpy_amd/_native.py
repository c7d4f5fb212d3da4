"""ctypes binding of the C ABI declared in include/tomograpy_amd.h.

The library ``libtomograpy_amd.so`` is built in-tree by tomograpy_amd/build.py
(``__graft_entry__.build()``).  There is no fallback: importing this module
without the library raises, and every call checks the returned status.
"""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libtomograpy_amd.so")

TG_F32, TG_F64 = 0, 1
TG_PJ, TG_BPJ, TG_PJT, TG_BPJT = 0, 1, 2, 3
TG_OBSTACLE_SUN = 1
TG_DETERMINISTIC = 2
TG_ZERO_OUT = 4
TG_ACC_F64 = 8
TG_HDR = 18
TG_CG_SCAL = 16 + 1024

TG_OK, TG_EINVAL, TG_EHIP, TG_EWORKSPACE, TG_EGEOMETRY = 0, -1, -2, -3, -4

_vp = ctypes.c_void_p
_i32 = ctypes.c_int32
_i64 = ctypes.c_int64
_u32 = ctypes.c_uint32
_sz = ctypes.c_size_t
_dbl = ctypes.c_double


class TgMap(ctypes.Structure):
    _fields_ = [("cdelt", _dbl * 3), ("pshape", _dbl * 3), ("mmin", _dbl * 3),
                ("mmax", _dbl * 3), ("n", _i32 * 3), ("nt", _i32), ("stride", _i64 * 4)]


class TgCall(ctypes.Structure):
    _fields_ = [("dtype", _i32), ("kind", _i32), ("flags", _u32), ("tile_j", _i32),
                ("map", TgMap),
                ("n1", _i32), ("n2", _i32), ("nv", _i32), ("pad0", _i32),
                ("data_stride", _i64 * 3),
                ("views", _vp), ("cube", _vp), ("data", _vp),
                ("pixmask", _vp), ("mask_stride", _i64 * 3),
                ("tindex", _vp), ("pix_weight", _vp), ("vox_weight", _vp),
                ("voxmask", _vp), ("vox_stride", _i64 * 3),
                ("status", _vp), ("workspace", _vp), ("workspace_bytes", _sz),
                ("acc", _vp), ("ray_steps", _vp)]


class TgCgSeg(ctypes.Structure):
    _fields_ = [("x", _vp), ("r", _vp), ("p", _vp), ("q", _vp), ("voxmask", _vp), ("weight", _vp),
                ("scal", _vp),
                ("n", _i64), ("block0", _i64), ("shape", _i32 * 4), ("axes", _u32),
                ("identity", _i32), ("lam", _dbl)]


(TG_CG_OP_REGULARIZE, TG_CG_OP_DOT_PQ, TG_CG_OP_UPDATE, TG_CG_OP_DIRECTION, TG_CG_OP_INIT,
 TG_CG_OP_ZERO_Q, TG_CG_OP_REGULARIZE_DOT) = range(7)


class TgBpjPlan(ctypes.Structure):
    _fields_ = [("brick", _i32 * 3), ("chunk_records", _i32), ("nb_axis", _i32 * 3),
                ("nbricks", _i32), ("order", _i32), ("trace_tile", _i32),
                ("tie_spread", _i32), ("pairs_ready", _i32), ("records", _i64), ("chunks", _i64),
                ("meta", _vp), ("meta_bytes", _sz), ("rec", _vp), ("rec_bytes", _sz),
                ("scratch", _vp), ("scratch_bytes", _sz), ("ray_steps_out", _vp)]


class TgBatchPlan(ctypes.Structure):
    _fields_ = [("tile_j", _i32), ("blocks", _i32), ("schedule", _i32), ("pad", _i32)]


# (name, restype, argtypes) for every symbol the header declares.
SIGNATURES = [
    ("tg_last_error", ctypes.c_char_p, []),
    ("tg_version", ctypes.c_char_p, []),
    ("tg_views_bytes", _sz, [_i32, _i32, _i32, _i32]),
    ("tg_views_pack", ctypes.c_int, [_i32, _i32, _i32, _i32, _vp, _vp, _sz]),
    ("tg_conic_full_projector", ctypes.c_int, [ctypes.POINTER(TgCall), _vp]),
    ("tg_conic_image_projector", ctypes.c_int, [ctypes.POINTER(TgCall), _i32, _vp]),
    ("tg_ray_projector", ctypes.c_int, [ctypes.POINTER(TgCall), _i32, _i32, _i32, _vp]),
    ("tg_unit_vector", ctypes.c_int, [ctypes.POINTER(TgCall), _i32, _i32, _vp, _vp]),
    ("tg_intersection_parameters", ctypes.c_int,
     [ctypes.POINTER(TgCall), _i32, _i32, _vp, _vp, _vp, _vp]),
    ("tg_count_steps", ctypes.c_int, [ctypes.POINTER(TgCall), _vp, _vp, _vp]),
    ("tg_det_workspace_min_bytes", _sz, [ctypes.POINTER(TgCall)]),
    ("tg_det_plan_workspace_bytes", _sz, [ctypes.POINTER(TgCall)]),
    ("tg_scale_cube", ctypes.c_int, [ctypes.POINTER(TgCall), _vp]),
    ("tg_vox_apply", ctypes.c_int, [ctypes.POINTER(TgCall), _vp, ctypes.POINTER(_i64), _vp]),
    ("tg_regularize", ctypes.c_int, [_i32, ctypes.POINTER(_i32), _u32, _dbl, _i32, _vp, _vp, _vp,
                                     _vp, _vp]),
    ("tg_cg_dot", ctypes.c_int, [_i32, _i64, _vp, _vp, _vp, _i32, _vp]),
    ("tg_cg_update", ctypes.c_int, [_i32, _i64, _vp, _vp, _vp, _vp, _vp, _vp]),
    ("tg_cg_direction", ctypes.c_int, [_i32, _i64, _vp, _vp, _vp, _vp]),
    ("tg_smoothness", ctypes.c_int, [_i32, ctypes.POINTER(_i32), ctypes.POINTER(_dbl), _vp, _vp,
                                     _vp, _vp, _i32, _vp]),
    ("tg_bpj_plan_init", ctypes.c_int, [ctypes.POINTER(TgCall), ctypes.POINTER(TgBpjPlan)]),
    ("tg_bpj_plan_count", ctypes.c_int, [ctypes.POINTER(TgCall), ctypes.POINTER(TgBpjPlan), _vp]),
    ("tg_bpj_plan_build", ctypes.c_int, [ctypes.POINTER(TgCall), ctypes.POINTER(TgBpjPlan), _vp]),
    ("tg_bpj_plan_backproject", ctypes.c_int, [ctypes.POINTER(TgCall), ctypes.POINTER(TgBpjPlan),
                                               _vp]),
    ("tg_bpj_plan_batch_bytes", _sz, [_i32]),
    ("tg_bpj_plan_batch", ctypes.c_int, [ctypes.POINTER(TgCall), ctypes.POINTER(TgBpjPlan), _i32,
                                         _vp, _vp, _vp]),
    ("tg_cg_seg_blocks", _i64, [_i64]),
    ("tg_cg_batch", ctypes.c_int, [_i32, _i32, _vp, _i32, _i64, _vp, _vp]),
    ("tg_batch_queue_bytes", _sz, [_i32]),
    ("tg_batch_plan_table", ctypes.c_int, [ctypes.POINTER(TgCall), _i32,
                                           ctypes.POINTER(TgBatchPlan), _vp]),
    ("tg_batch_job_layout", ctypes.c_int, [_i32, ctypes.POINTER(_sz), ctypes.POINTER(_sz),
                                           ctypes.POINTER(_sz)]),
    ("tg_batch_run", ctypes.c_int, [ctypes.POINTER(TgCall), _i32, ctypes.POINTER(TgBatchPlan),
                                    _vp, _vp, _vp]),
    ("tg_conic_full_projector_static", ctypes.c_int, [ctypes.POINTER(TgCall), _vp]),
    ("tg_static_acc_capacity", _i64, []),
]


class NativeError(RuntimeError):
    """A tg_* call returned an error status."""

    def __init__(self, code, message):
        super().__init__("tomograpy_amd native error %d: %s" % (code, message))
        self.code = code


_lib = None


def lib():
    """The loaded library (loads it on first use; raises if it is missing)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(
                "tomograpy_amd: %s is missing; build it with `python -c "
                "'import __graft_entry__ as g; g.build()'` or `python tomograpy_amd/build.py`"
                % LIB_PATH)
        l = ctypes.CDLL(LIB_PATH)
        for name, res, args in SIGNATURES:
            f = getattr(l, name)
            f.restype = res
            f.argtypes = args
        _lib = l
    return _lib


def check(rc):
    """Raises NativeError when rc is not TG_OK."""
    if rc != TG_OK:
        msg = lib().tg_last_error()
        raise NativeError(rc, msg.decode() if msg else "")
    return rc

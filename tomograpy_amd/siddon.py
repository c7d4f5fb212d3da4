"""Siddon conic projector / backprojector: the reference's Python surface.

Mirrors tomograpy/siddon.py:85-294 -- ``projector``, ``backprojector``,
``projector4d``, ``backprojector4d``, ``conic_image_projector``,
``conic_image_backprojector``, ``C_full_unit_vector``,
``C_full_intersection_parameters`` -- with the same arguments, the same
in-place accumulation (the output array is added to, never cleared) and the
same errors.  Every call runs the gfx950 HIP library through its C ABI.

Arrays:
  * torch tensors on a ROCm device, with a ``header`` attribute, are used in
    place (asynchronous on the current stream);
  * numpy arrays with a ``header`` attribute (``InfoArray``) are copied to
    the device, processed, and the updated output is copied back into them.

``data`` has shape (n1, n2, nv) with the views on the last axis and
``data.header`` a list of per-view dicts; ``cube`` has shape (N0, N1, N2)
(or (N0, N1, N2, T) for the 4-D variants) and ``cube.header`` a dict.
"""
import collections
import contextlib
import ctypes

import numpy as np
import torch

from . import _native as N
from .headers import check_headers, complete_view_headers, map_record, view_records

_DTYPES = {"float32": N.TG_F32, "float64": N.TG_F64}
_TORCH = {N.TG_F32: torch.float32, N.TG_F64: torch.float64}
_OBSTACLES = {None: 0, "none": 0, "sun": N.TG_OBSTACLE_SUN}

# Packed view descriptors on the device, keyed by their content.
_VIEW_CACHE = collections.OrderedDict()
_VIEW_CACHE_MAX = 64
_STATUS = {}


def device():
    """The ROCm device the library runs on (the current torch device)."""
    if not torch.cuda.is_available():
        raise RuntimeError("tomograpy_amd needs a ROCm GPU (torch.cuda.is_available() is False)")
    return torch.device("cuda", torch.cuda.current_device())


def _dtype_code(dtype):
    name = dtype.name if isinstance(dtype, np.dtype) else str(dtype).replace("torch.", "")
    # Same failure as the reference's ctypes_inv[data.dtype.name] lookup (siddon.py:111).
    return _DTYPES[name]


def _status_tensor(dev):
    st = _STATUS.get(dev)
    if st is None:
        st = torch.zeros(1, dtype=torch.int32, device=dev)
        _STATUS[dev] = st
    return st


def check_status(dev=None, reset=True):
    """Raises if a kernel flagged a stalled (NaN) ray since the last check (synchronises)."""
    dev = dev or device()
    st = _status_tensor(dev)
    code = int(st.item())
    if reset and code:
        st.zero_()
    if code:
        raise N.NativeError(code, "a ray's traversal state was NaN (reference would hang)")


def views_tensor(records, dtype_code, n1, n2, dev):
    """Device buffer of packed view descriptors (cached by content and stream)."""
    records = np.ascontiguousarray(records, dtype=np.float64)
    # (per stream: an entry's upload is ordered before its uses on that stream)
    stream = current_stream().cuda_stream if torch.device(dev).type == "cuda" else 0
    key = (dtype_code, int(n1), int(n2), str(dev), stream, records.shape, records.tobytes())
    hit = _VIEW_CACHE.get(key)
    if hit is not None:
        _VIEW_CACHE.move_to_end(key)
        return hit
    lib = N.lib()
    nv = records.shape[0]
    nbytes = lib.tg_views_bytes(dtype_code, nv, n1, n2)
    host = np.zeros(max(nbytes, 1), dtype=np.uint8)
    N.check(lib.tg_views_pack(dtype_code, nv, n1, n2, records.ctypes.data, host.ctypes.data, nbytes))
    dev_buf = _upload(torch.from_numpy(host), dev)
    _VIEW_CACHE[key] = dev_buf
    while len(_VIEW_CACHE) > _VIEW_CACHE_MAX:
        _VIEW_CACHE.popitem(last=False)
    return dev_buf


def _upload(host, dev):
    """Host tensor to ``dev``; to a GPU from pinned memory, so the copy does not
    wait for the work already queued on the stream."""
    if torch.device(dev).type != "cuda":
        return host.to(dev)
    return host.pin_memory().to(dev, non_blocking=True)


def _ptr(t):
    return None if t is None else ctypes.c_void_p(t.data_ptr())


def make_call(kind, dtype_code, data_shape, data_strides, cube_shape, cube_strides, map_rec,
              views, cube_ptr, data_ptr, flags=0, nt=1, tile_j=0, pixmask=None, tindex=None,
              pix_weight=None, vox_weight=None, voxmask=None, status=None, workspace=None):
    """Fills a TgCall from shapes, strides (in elements) and device tensors."""
    c = N.TgCall()
    c.dtype = dtype_code
    c.kind = kind
    c.flags = flags
    c.tile_j = tile_j
    m = c.map
    for k in range(3):
        m.cdelt[k] = map_rec[k]
        m.pshape[k] = map_rec[3 + k]
        m.mmin[k] = map_rec[6 + k]
        m.mmax[k] = map_rec[9 + k]
        m.n[k] = int(cube_shape[k])
        m.stride[k] = int(cube_strides[k])
    m.nt = int(nt)
    m.stride[3] = int(cube_strides[3]) if len(cube_strides) > 3 else 0
    c.n1, c.n2, c.nv = (int(x) for x in data_shape)
    for k in range(3):
        c.data_stride[k] = int(data_strides[k])
    c.views = views.data_ptr()
    c.cube = cube_ptr
    c.data = data_ptr
    if pixmask is not None:
        c.pixmask = pixmask.data_ptr()
        for k in range(3):
            c.mask_stride[k] = int(pixmask.stride(k))
    if tindex is not None:
        c.tindex = tindex.data_ptr()
    if pix_weight is not None:
        if tuple(pix_weight.stride()) != tuple(data_strides):
            raise ValueError("pixel weights must have the data array's strides")
        c.pix_weight = pix_weight.data_ptr()
    if vox_weight is not None or voxmask is not None:
        ref = vox_weight if vox_weight is not None else voxmask
        vs = ref.stride()
        if vox_weight is not None and voxmask is not None and tuple(voxmask.stride()) != tuple(vs):
            raise ValueError("voxel mask and voxel weights must share strides")
        for k in range(3):
            c.vox_stride[k] = int(vs[k])
        if vox_weight is not None:
            c.vox_weight = vox_weight.data_ptr()
        if voxmask is not None:
            c.voxmask = voxmask.data_ptr()
    if status is not None:
        c.status = status.data_ptr()
    if workspace is not None:
        c.workspace = workspace.data_ptr()
        c.workspace_bytes = workspace.numel() * workspace.element_size()
    return c


# The stream handle and torch stream fixed by fixed_stream(), else None.
_FIXED_STREAM = None
_FIXED_TORCH_STREAM = None


def stream_handle():
    """The HIP stream the library's calls go to: torch's current stream."""
    if _FIXED_STREAM is not None:
        return _FIXED_STREAM
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def current_stream():
    """torch's current stream (the one fixed by fixed_stream() inside it)."""
    if _FIXED_TORCH_STREAM is not None:
        return _FIXED_TORCH_STREAM
    return torch.cuda.current_stream()


@contextlib.contextmanager
def fixed_stream():
    """Inside the block, stream_handle() and current_stream() return the stream
    current at entry without asking torch again (about 9 us a call), for loops
    that make many library calls on one stream.  The block must not switch
    streams, and other threads must not call the library meanwhile."""
    global _FIXED_STREAM, _FIXED_TORCH_STREAM
    prev, prev_t = _FIXED_STREAM, _FIXED_TORCH_STREAM
    _FIXED_TORCH_STREAM = torch.cuda.current_stream()
    _FIXED_STREAM = ctypes.c_void_p(_FIXED_TORCH_STREAM.cuda_stream)
    try:
        yield
    finally:
        _FIXED_STREAM, _FIXED_TORCH_STREAM = prev, prev_t


def det_workspace(call, dev, budget_bytes=None):
    """Workspace tensor for deterministic backprojection of `call`: room for the
    canonical brick plan the call builds in it (tg_det_plan_workspace_bytes, an
    upper bound) when that fits in half the free HBM, else min(half the free
    HBM, 16 GiB) for the chunked passes."""
    need = N.lib().tg_det_workspace_min_bytes(ctypes.byref(call))
    if budget_bytes is None:
        free, _ = torch.cuda.mem_get_info(dev)
        plan = int(N.lib().tg_det_plan_workspace_bytes(ctypes.byref(call)))
        budget_bytes = plan if 0 < plan <= free // 2 else min(int(free * 0.5), 16 << 30)
    nbytes = max(int(need), int(budget_bytes))
    return torch.empty(nbytes, dtype=torch.uint8, device=dev)


class _Operand:
    """A device view of an argument array, copied back on exit if it was host memory."""

    def __init__(self, x, dev, dtype=None, name="array"):
        self.host = None
        if isinstance(x, torch.Tensor):
            if x.device.type != "cuda":
                self.host = x
                x = x.to(dev)
            self.t = x
        else:
            self.host = x
            arr = np.asarray(x)
            self.t = torch.from_numpy(np.ascontiguousarray(arr)).to(dev)
        if dtype is not None and self.t.dtype != dtype:
            raise ValueError("%s has dtype %s, expected %s" % (name, self.t.dtype, dtype))

    def writeback(self):
        if self.host is None:
            return
        if isinstance(self.host, torch.Tensor):
            self.host.copy_(self.t.cpu())
        else:
            self.host[...] = self.t.cpu().numpy()


def _mask_tensor(mask, shape, dev):
    if mask is None:
        return None
    if isinstance(mask, torch.Tensor):
        m = mask.to(dev)
    else:
        m = _upload(torch.from_numpy(np.ascontiguousarray(np.asarray(mask))), dev)
    if tuple(m.shape) != tuple(shape):
        raise ValueError("mask shape %s does not match data shape %s" % (tuple(m.shape), tuple(shape)))
    # Skip a pixel iff its mask value is nonzero (C_siddon.c.template:520).
    return (m != 0).to(torch.uint8)


# A backprojection call of at least this many steps (bounded by rays x
# (n0 + n1 + n2)) goes through a brick plan built for the call -- count, fill,
# then the brick kernel -- instead of one global atomic per step.  Measured
# per call, build included, on device arrays: 256^3 seen by 128 views of 512^2
# in 20.2 against 155 ms; 128^3 by 64 of 512^2 in 6.9 against 79 ms; 128^3 by
# 64 of 256^2 in 2.0 against 10.7 ms.  Below it, a call of at least
# PLAN_CALL_DENSE_MIN_STEPS with at least one ray per voxel takes a plan too:
# there the atomics of neighbouring rays collide on the same voxels (64^3 by 32
# of 128^2: 0.55 against 0.92 ms; 64^3 by 16 of 128^2: 0.45 against 0.50 ms),
# while sparser or smaller calls keep the scatter (96^3 by 32 of 128^2: 0.63
# against 0.59 ms; 48^3 by 16 of 96^2: 0.39 against 0.22 ms).  A plan used
# once is left in trace order: sorting its passes by step count (what the
# operators do for the many backprojections of a solve) costs more than it
# saves in one (at 256^3, 128 x 512^2: 35.5 ms sorted against 29.9 ms).
PLAN_CALL_MIN_STEPS = 1 << 28
PLAN_CALL_DENSE_MIN_STEPS = 1 << 25


def _plan_backprojection(data_header, cube_header, d, c, mask, obstacle, tdtype, dev,
                         tindex=None):
    """cube += P^T data through a brick plan of this geometry built for the
    call (the operators' backward: float32 data sums in float64, independent
    of the order).  A 4-D cube (N0, N1, N2, T) takes image t into bin
    tindex[t] (default t).  Returns False, with nothing done, when the plan
    would take more than half the free device memory."""
    from .lo_wrapper import Siddon, Siddon4d
    if c.t.dim() == 4:
        hdr = dict(cube_header)
        hdr["NAXIS4"] = int(c.t.shape[3])
        ti = np.arange(len(data_header)) if tindex is None else np.asarray(tindex)
        P = Siddon4d(data_header, hdr, tindex=ti, mask=mask, obstacle=obstacle, dtype=tdtype,
                     device=dev, bpj_mode="plan", plan_order=0)
    else:
        P = Siddon(data_header, cube_header, mask=mask, obstacle=obstacle, dtype=tdtype,
                   device=dev, bpj_mode="plan", plan_order=0)
    free, _ = torch.cuda.mem_get_info(dev)
    if P.plan_reserve_bytes() > free // 2:
        return False
    P.backproject_into(d.t, c.t)
    return True


def _siddon(kind, data, cube, mask=None, obstacle=None, t=None, ray=None, deterministic=False,
            tindex=None, tile_j=0, sync=None):
    if obstacle not in _OBSTACLES:
        raise KeyError(obstacle)  # obstacles_inv[obstacle] (siddon.py:112)
    if not hasattr(data, "header") or not hasattr(cube, "header"):
        raise AttributeError("data and cube need a header attribute")
    if not isinstance(cube.header, dict):
        cube.header = dict(cube.header)
    data_header, cube_header = check_headers(data.header, cube.header)
    if str(data.dtype).replace("torch.", "") != str(cube.dtype).replace("torch.", ""):
        raise ValueError("data and cube map should have the same data-type")
    code = _dtype_code(data.dtype)
    dev = device()
    tdtype = _TORCH[code]
    d = _Operand(data, dev, tdtype, "data")
    c = _Operand(cube, dev, tdtype, "cube")
    if d.t.dim() != 3:
        raise ValueError("data must have 3 dimensions (n1, n2, n_images)")
    time = kind in (N.TG_PJT, N.TG_BPJT)
    if c.t.dim() != (4 if time else 3):
        raise ValueError("cube must have %d dimensions" % (4 if time else 3))
    n1, n2, nv = d.t.shape
    if len(data_header) != nv:
        raise ValueError("data.header has %d entries for %d images" % (len(data_header), nv))
    views = views_tensor(view_records(data_header), code, n1, n2, dev)
    pm = _mask_tensor(mask, d.t.shape, dev)
    ti = None
    nt = 1
    if time:
        nt = c.t.shape[3]
        if tindex is not None:
            ti = torch.as_tensor(np.asarray(tindex, dtype=np.int32)).to(dev)
            if ti.numel() != nv or int(ti.min()) < 0 or int(ti.max()) >= nt:
                raise ValueError("tindex must give a bin in [0, %d) for each of %d views" % (nt, nv))
        elif nv > nt:
            raise ValueError("4-D projection of %d images needs at least %d time bins" % (nv, nv))
    flags = _OBSTACLES[obstacle]
    bwd = kind in (N.TG_BPJ, N.TG_BPJT)
    if deterministic and bwd:
        flags |= N.TG_DETERMINISTIC
    ws = None
    rays, bound = n1 * n2 * nv, n1 * n2 * nv * sum(c.t.shape[:3])
    # (voxels x time bins: the rays of a 4-D call spread over its bins)
    dense = bound >= PLAN_CALL_DENSE_MIN_STEPS and rays >= int(np.prod(c.t.shape))
    planned = (bwd and not deterministic and t is None and ray is None
               and (bound >= PLAN_CALL_MIN_STEPS or dense)
               and n1 * n2 * nv < (1 << 32) and max(c.t.shape[:3]) <= 1024
               and _plan_backprojection(data_header, cube_header, d, c, mask, obstacle, tdtype,
                                        dev, tindex))
    if not planned:
        st = _status_tensor(dev)
        call = make_call(kind, code, d.t.shape, d.t.stride(), c.t.shape, c.t.stride(),
                         map_record(cube_header), views, c.t.data_ptr(), d.t.data_ptr(),
                         flags=flags, nt=nt, tile_j=tile_j, pixmask=pm, tindex=ti, status=st)
        if flags & N.TG_DETERMINISTIC:
            ws = det_workspace(call, dev)
            call.workspace = ws.data_ptr()
            call.workspace_bytes = ws.numel()
        lib = N.lib()
        s = stream_handle()
        if ray is not None:
            i, j, tt = (int(x) for x in ray)
            N.check(lib.tg_ray_projector(ctypes.byref(call), i, j, tt, s))
        elif t is not None:
            N.check(lib.tg_conic_image_projector(ctypes.byref(call), int(t), s))
        else:
            N.check(lib.tg_conic_full_projector(ctypes.byref(call), s))
    host = d.host is not None or c.host is not None
    if sync or (sync is None and host):
        torch.cuda.current_stream().synchronize()
        check_status(dev)
    if bwd:
        c.writeback()
    else:
        d.writeback()
    del ws
    return data if not bwd else cube


def projector(data, cube, mask=None, obstacle=None, nthread=0, **kw):
    """Adds the conic projection of `cube` into `data` (siddon.py:85-117).

    ``mask``: pixels whose mask value is nonzero are skipped.  ``obstacle``:
    None or "sun" (stop rays at the unit sphere).  ``nthread`` is accepted
    for compatibility and ignored (the GPU schedules the rays).
    """
    return _siddon(N.TG_PJ, data, cube, mask=mask, obstacle=obstacle, **kw)


def backprojector(data, cube, mask=None, obstacle=None, nthread=0, deterministic=False, **kw):
    """Adds the backprojection of `data` into `cube` (siddon.py:119-151).

    ``deterministic=True`` sums every voxel in the reference's single-thread
    order (views, rows, columns), bit-identical to ``nthread=1``.
    """
    return _siddon(N.TG_BPJ, data, cube, mask=mask, obstacle=obstacle,
                   deterministic=deterministic, **kw)


def projector4d(data, cube, mask=None, obstacle=None, nthread=0, tindex=None, **kw):
    """Projection of a time-dependent map (siddon.py:153-185).

    Image t reads time bin t of the 4-D cube (C_siddon.c.template:931-936), or
    bin ``tindex[t]`` when given.
    """
    return _siddon(N.TG_PJT, data, cube, mask=mask, obstacle=obstacle, tindex=tindex, **kw)


def backprojector4d(data, cube, mask=None, obstacle=None, nthread=0, tindex=None,
                    deterministic=False, **kw):
    """Backprojection into a time-dependent map (siddon.py:187-219)."""
    return _siddon(N.TG_BPJT, data, cube, mask=mask, obstacle=obstacle, tindex=tindex,
                   deterministic=deterministic, **kw)


def conic_image_projector(data, cube, t, mask=None, obstacle=None, **kw):
    """Projection of image t only (siddon.py:221-231)."""
    return _siddon(N.TG_PJ, data, cube, mask=mask, obstacle=obstacle, t=t, **kw)


def conic_image_backprojector(data, cube, t, mask=None, obstacle=None, **kw):
    """Backprojection of image t only (siddon.py:233-243).

    Returns ``data`` as the reference does (siddon.py:243); the cube is
    updated in place.
    """
    _siddon(N.TG_BPJ, data, cube, mask=mask, obstacle=obstacle, t=t, **kw)
    return data


def ray_projector(data, cube, index, mask=None, obstacle=None, backward=False, **kw):
    """Projects (or backprojects) the single ray of pixel ``index = (i, j, t)``
    (C_siddon.c.template:250-307)."""
    kind = N.TG_BPJ if backward else N.TG_PJ
    return _siddon(kind, data, cube, mask=mask, obstacle=obstacle, ray=index, **kw)


def check_projector_inputs(data, cube):
    """Completes headers and checks dtypes (siddon.py:245-270)."""
    if not isinstance(cube.header, dict):
        cube.header = dict(cube.header)
    check_headers(data.header, cube.header)
    if str(data.dtype).replace("torch.", "") != str(cube.dtype).replace("torch.", ""):
        raise ValueError("data and cube map should have the same data-type")


def _geometry_call(data, cube_header=None):
    code = _dtype_code(data.dtype)
    dev = device()
    d = _Operand(data, dev, _TORCH[code], "data")
    n1, n2, nv = d.t.shape
    views = views_tensor(view_records(data.header), code, n1, n2, dev)
    mrec = map_record(cube_header) if cube_header is not None else np.zeros(12)
    nshape = [int(cube_header["NAXIS%d" % (k + 1)]) for k in range(3)] if cube_header else [1, 1, 1]
    call = make_call(N.TG_PJ, code, d.t.shape, d.t.stride(), nshape, [0, 0, 0, 0], mrec, views,
                     None, d.t.data_ptr())
    return call, d, dev, code


def C_full_unit_vector(data):
    """Unit vector of every pixel, shape data.shape + (3,) (siddon.py:272-280).

    The output has the data's dtype (the reference allocates float64 and lets
    the float32 module write float32 into it, siddon.py:273 vs C:642-644).
    Pixels whose data value is NaN are left at zero (C_siddon.c.template:638).
    """
    complete_view_headers(data.header)
    call, d, dev, code = _geometry_call(data)
    u = torch.zeros(tuple(d.t.shape) + (3,), dtype=_TORCH[code], device=dev)
    N.check(N.lib().tg_unit_vector(ctypes.byref(call), 0, call.nv, u.data_ptr(), stream_handle()))
    return u if isinstance(data, torch.Tensor) and data.device.type == "cuda" else u.cpu().numpy()


def C_full_intersection_parameters(data, cube, u):
    """a1, an per pixel from unit vectors u (siddon.py:282-294)."""
    if not isinstance(cube.header, dict):
        cube.header = dict(cube.header)
    check_headers(data.header, cube.header)
    call, d, dev, code = _geometry_call(data, cube.header)
    ut = u if isinstance(u, torch.Tensor) else torch.from_numpy(np.ascontiguousarray(u))
    ut = ut.to(dev, _TORCH[code]).contiguous()
    a1 = torch.zeros_like(ut)
    an = torch.zeros_like(ut)
    N.check(N.lib().tg_intersection_parameters(ctypes.byref(call), 0, call.nv, ut.data_ptr(),
                                               a1.data_ptr(), an.data_ptr(), stream_handle()))
    if isinstance(data, torch.Tensor) and data.device.type == "cuda":
        return a1, an
    return a1.cpu().numpy(), an.cpu().numpy()


def count_steps(data, cube, mask=None, obstacle=None, per_ray=False):
    """Number of Siddon steps (voxel segments) a projection of `data` through
    `cube` takes: the unit of the throughput figures.  With per_ray=True also
    returns the (nv, n1, n2) per-ray counts."""
    if not isinstance(cube.header, dict):
        cube.header = dict(cube.header)
    data_header, cube_header = check_headers(data.header, cube.header)
    code = _dtype_code(data.dtype)
    dev = device()
    d = _Operand(data, dev, _TORCH[code], "data")
    n1, n2, nv = d.t.shape
    views = views_tensor(view_records(data_header), code, n1, n2, dev)
    pm = _mask_tensor(mask, d.t.shape, dev)
    shape = [int(cube_header["NAXIS%d" % (k + 1)]) for k in range(3)]
    call = make_call(N.TG_PJ, code, d.t.shape, d.t.stride(), shape, [0, 0, 0, 0],
                     map_record(cube_header), views, None, d.t.data_ptr(),
                     flags=_OBSTACLES[obstacle], pixmask=pm, status=_status_tensor(dev))
    total = torch.zeros(1, dtype=torch.int64, device=dev)
    pr = torch.zeros((nv, n1, n2), dtype=torch.int32, device=dev) if per_ray else None
    N.check(N.lib().tg_count_steps(ctypes.byref(call), _ptr(pr), total.data_ptr(), stream_handle()))
    n = int(total.item())
    return (n, pr) if per_ray else n

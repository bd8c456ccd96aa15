"""Host-side mirror of the reference's C++ API for the thinning path.

Names, argument meaning and output layout follow the reference
(2D/LibThinning2D.h:13-15, 3D/LibThinning3D.h:13-15, 2D/ExtractSegments.h:16-17,
3D/ExtractSegments.h:21-22, 2D/CenterlineExtraction.h:25-28,
3D/CenterlineExctraction.h:26-29); arrays are numpy (host) or torch tensors on
the GPU (device-resident variants).  Every call goes through the HIP C-ABI;
errors raise BGTError instead of the reference's printf/exit.
"""
import ctypes

import numpy as np

from ._lib import DTYPES, Segments, check, check_device_pair, lib


def _dtype_code(a):
    name = np.dtype(a.dtype).name
    if name not in DTYPES:
        raise TypeError(f"unsupported pixel type {a.dtype} (reference instantiations: "
                        f"{', '.join(DTYPES)})")
    return DTYPES[name]


def _ptr(a):
    return ctypes.c_void_p(a.ctypes.data)


def Thinning2D(data):
    """Thinning2D<T>(width, height, data, out): returns the thinned image (same
    dtype, 0/1) of a (height, width) array."""
    a = np.ascontiguousarray(data)
    if a.ndim != 2:
        raise ValueError("Thinning2D expects a (height, width) array")
    out = np.empty_like(a)
    check(lib().bgt_thinning2d(_dtype_code(a), a.shape[1], a.shape[0], _ptr(a), _ptr(out)))
    return out


def Thinning3D(data):
    """Thinning3D<T>(width, height, depth, data, out) of a (depth, height,
    width) array."""
    a = np.ascontiguousarray(data)
    if a.ndim != 3:
        raise ValueError("Thinning3D expects a (depth, height, width) array")
    out = np.empty_like(a)
    check(lib().bgt_thinning3d(_dtype_code(a), a.shape[2], a.shape[1], a.shape[0], _ptr(a),
                               _ptr(out)))
    return out


class SSegment:
    """One traced segment: points (n, 2) [x, y] or (n, 3) [x, y, z] int16 and
    the is_end_segment flag (2D/Structs.h:27-34)."""
    __slots__ = ("points", "is_end_segment", "pruned")

    def __init__(self, points, is_end):
        self.points = points
        self.is_end_segment = bool(is_end)
        self.pruned = False

    def __repr__(self):
        return f"SSegment(n={len(self.points)}, is_end={self.is_end_segment})"


def _take_segments(s, three_d):
    n, m = s.nseg, s.npts
    x = np.ctypeslib.as_array(s.x, (m,)).copy() if m else np.zeros(0, np.int16)
    y = np.ctypeslib.as_array(s.y, (m,)).copy() if m else np.zeros(0, np.int16)
    z = (np.ctypeslib.as_array(s.z, (m,)).copy() if m else np.zeros(0, np.int16)) if three_d else None
    st = np.ctypeslib.as_array(s.seg_start, (n + 1,)).copy()
    ie = np.ctypeslib.as_array(s.is_end, (n,)).copy() if n else np.zeros(0, np.uint8)
    lib().bgt_segments_free(ctypes.byref(s))
    return {"x": x, "y": y, "z": z, "start": st, "is_end": ie}


def extract_segments_flat(data):
    """ExtractSegmentsV as flat arrays: x, y (, z), start (nseg+1), is_end."""
    a = np.ascontiguousarray(data)
    s = Segments()
    if a.ndim == 2:
        check(lib().bgt_extract_segments2d(_dtype_code(a), _ptr(a), a.shape[1], a.shape[0],
                                           ctypes.byref(s)))
        return _take_segments(s, False)
    if a.ndim == 3:
        check(lib().bgt_extract_segments3d(_dtype_code(a), _ptr(a), a.shape[2], a.shape[1],
                                           a.shape[0], ctypes.byref(s)))
        return _take_segments(s, True)
    raise ValueError("ExtractSegmentsV expects a 2D or 3D array")


def ExtractSegmentsV(data):
    """ExtractSegmentsV<T>(data, width, height[, depth]) -> list of SSegment."""
    f = extract_segments_flat(data)
    cols = [f["x"], f["y"]] + ([f["z"]] if f["z"] is not None else [])
    pts = np.stack(cols, axis=1) if len(f["x"]) else np.zeros((0, len(cols)), np.int16)
    st = f["start"]
    return [SSegment(pts[st[i]:st[i + 1]], f["is_end"][i]) for i in range(len(st) - 1)]


def CenterlineExtraction(data, max_segment_length, smoothing_span):
    """CenterlineExtraction<T>: returns (out_x, out_y, out_idx) for a 2D image
    or (out_x, out_y, out_z, out_idx) for a volume; out_idx has nsegs + 1
    entries with out_idx[nsegs] = number of points."""
    a = np.ascontiguousarray(data)
    px = ctypes.POINTER(ctypes.c_float)()
    py = ctypes.POINTER(ctypes.c_float)()
    pz = ctypes.POINTER(ctypes.c_float)()
    pi = ctypes.POINTER(ctypes.c_uint32)()
    ns = ctypes.c_uint32(0)
    if a.ndim == 2:
        check(lib().bgt_centerline2d(_dtype_code(a), _ptr(a), a.shape[1], a.shape[0],
                                     int(max_segment_length), int(smoothing_span),
                                     ctypes.byref(px), ctypes.byref(py), ctypes.byref(pi),
                                     ctypes.byref(ns)))
    elif a.ndim == 3:
        check(lib().bgt_centerline3d(_dtype_code(a), _ptr(a), a.shape[2], a.shape[1], a.shape[0],
                                     int(max_segment_length), int(smoothing_span),
                                     ctypes.byref(px), ctypes.byref(py), ctypes.byref(pz),
                                     ctypes.byref(pi), ctypes.byref(ns)))
    else:
        raise ValueError("CenterlineExtraction expects a 2D or 3D array")
    n = ns.value
    idx = np.ctypeslib.as_array(pi, (n + 1,)).copy()
    m = int(idx[n])
    take = lambda p: np.ctypeslib.as_array(p, (m,)).copy() if m else np.zeros(0, np.float32)
    x, y = take(px), take(py)
    z = take(pz) if a.ndim == 3 else None
    for p in (px, py, pi) + ((pz,) if a.ndim == 3 else ()):
        lib().bgt_free(ctypes.cast(p, ctypes.c_void_p))
    return (x, y, idx) if a.ndim == 2 else (x, y, z, idx)


# ---- device-resident variants (torch tensors on the GPU) --------------------

def _stream_handle(stream):
    import torch
    s = stream if stream is not None else torch.cuda.current_stream()
    return ctypes.c_void_p(s.cuda_stream)


def thin2d_device(img, out=None, stream=None):
    """Thin a (h, w) uint8 CUDA tensor; returns (out, iterations)."""
    import torch
    if not isinstance(img, torch.Tensor) or img.dtype != torch.uint8 or img.dim() != 2 or not img.is_cuda:
        raise TypeError("thin2d_device expects a 2D uint8 GPU tensor")
    h, w = img.shape
    if out is None:
        out = torch.empty_like(img)
    check_device_pair(img, out, 2, "thin2d_device")
    it = ctypes.c_int(0)
    check(lib().bgt_thin2d_device(ctypes.c_void_p(img.data_ptr()), img.stride(0), w, h,
                                  ctypes.c_void_p(out.data_ptr()), out.stride(0),
                                  _stream_handle(stream), ctypes.byref(it)))
    return out, it.value


def thin3d_device(vol, out=None, stream=None):
    """Thin a contiguous (d, h, w) uint8 CUDA tensor; returns (out, iterations)."""
    import torch
    if not isinstance(vol, torch.Tensor) or vol.dtype != torch.uint8 or vol.dim() != 3 or not vol.is_cuda:
        raise TypeError("thin3d_device expects a contiguous 3D uint8 GPU tensor")
    d, h, w = vol.shape
    if out is None:
        out = torch.empty_like(vol)
    check_device_pair(vol, out, 3, "thin3d_device")
    it = ctypes.c_int(0)
    check(lib().bgt_thin3d_device(ctypes.c_void_p(vol.data_ptr()), w, h, d,
                                  ctypes.c_void_p(out.data_ptr()), _stream_handle(stream),
                                  ctypes.byref(it)))
    return out, it.value


def set_config2d(waves=4, rows=16, iters_per_pass=3, blocks_per_cu=-1):
    """Launch shape of the 2D pass kernel for later calls: waves per
    workgroup, register rows per lane, iterations per HBM pass and batch
    workgroups per CU (0 = occupancy limit, -1 = automatic: 2 in a batch
    that also holds volumes, else the occupancy limit)."""
    check(lib().bgt_set_config2d(waves, rows, iters_per_pass, blocks_per_cu))


def set_config3d(blocks_per_cu=2, max_chunk=16, tail_words=16384):
    """Launch shape of the 3D subiteration kernel for later calls: workgroups
    per CU (0 = occupancy limit), iterations queued per host poll, and the
    average active words per volume at which the remaining iterations run in
    one launch, one workgroup per volume (0 = never)."""
    check(lib().bgt_set_config3d(blocks_per_cu, max_chunk, tail_words))


def set_engine3d(engine=1):
    """3D engine for later calls (results do not depend on it): 0 = word
    lists with a one-launch tail, 1 = word lists then tiles from tail_words
    on, 2 = tiles throughout."""
    check(lib().bgt_set_engine3d(engine))


def set_lane_words2d(words=0):
    """64-bit words per lane of the 2D pass kernel for later calls (results do
    not depend on it): 1, 2 or 4, 0 = automatic (the fewest that hold a row
    in one wave64; widths up to 16384)."""
    check(lib().bgt_set_lane_words2d(words))


def set_fused_io2d(on=False):
    """Batches of 2D images up to 4096 wide: the first band pass reads the
    input bytes itself (True) or a separate pack launch does (False, the
    default).  Results do not depend on it."""
    check(lib().bgt_set_fused_io2d(int(bool(on))))


def set_persistent2d(on=False):
    """Batches of 2D images: the whole band-pass chain in one persistent
    launch (True) or a launch per pass driven by the host (False, the
    default).  Results do not depend on it."""
    check(lib().bgt_set_persistent2d(int(bool(on))))


def debug_fail_device_allocs(n):
    """Test hook: the library's next n device allocations fail (its calls
    then raise BGTError with BGT_ERR_NOMEM); 0 clears it."""
    check(lib().bgt_debug_fail_device_allocs(int(n)))


def extract_segments_device(skel, padded=None, stream=None):
    """ExtractSegmentsV over a contiguous uint8 torch tensor on the GPU (h x w
    or d x h x w), the top-left corner of a zero-padded frame of shape
    `padded` (default: the tensor's own): the skeleton points are classified
    and compacted on the GPU, and the host follows only them.  Returns the
    same flat arrays as extract_segments_flat on the padded frame."""
    import torch
    if skel.dtype != torch.uint8 or not skel.is_cuda or not skel.is_contiguous():
        raise ValueError("expects a contiguous uint8 CUDA tensor")
    shape = tuple(skel.shape)
    pad = tuple(padded) if padded is not None else shape
    s = Segments()
    if skel.dim() == 2:
        check(lib().bgt_extract_segments2d_device(ctypes.c_void_p(skel.data_ptr()), shape[1], shape[0],
                                                  pad[1], pad[0], _stream_handle(stream),
                                                  ctypes.byref(s)))
        return _take_segments(s, False)
    if skel.dim() == 3:
        check(lib().bgt_extract_segments3d_device(ctypes.c_void_p(skel.data_ptr()), shape[2], shape[1],
                                                  shape[0], pad[2], pad[1], pad[0],
                                                  _stream_handle(stream), ctypes.byref(s)))
        return _take_segments(s, True)
    raise ValueError("expects a 2D or 3D tensor")


def mex_columns(centerline):
    """The MEX gateways' output (2D/Matlab2DThinning.cpp:163-175,
    3D/Matlab3DThinning.cpp:157-170) from CenterlineExtraction's arrays:
    one (n, 2) [y x] or (n, 3) [y x z] float64 matrix per segment, laid out
    column-major as MATLAB holds it."""
    if len(centerline) == 3:
        x, y, idx = centerline
        z = None
    else:
        x, y, z, idx = centerline
    x = np.ascontiguousarray(x, np.float32)
    y = np.ascontiguousarray(y, np.float32)
    z = None if z is None else np.ascontiguousarray(z, np.float32)
    idx = np.ascontiguousarray(idx, np.uint32)
    nc = 2 if z is None else 3
    nsegs = len(idx) - 1
    out = np.empty(max(1, int(idx[-1]) * nc), np.float64)
    check(lib().bgt_mex_columns(_ptr(x), _ptr(y), _ptr(z) if z is not None else None, _ptr(idx),
                                nsegs, _ptr(out)))
    return [out[idx[i] * nc:idx[i + 1] * nc].reshape((nc, idx[i + 1] - idx[i])).T
            for i in range(nsegs)]

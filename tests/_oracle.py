"""ctypes bindings to the oracle (oracle/liboracle.so) and, when built, to the
reference's own host-compiled sources (oracle/_ref/*.so).

Test infrastructure only: the oracle is the checker the HIP path is compared
against; nothing in begputhinning_amd imports this module.
"""
import ctypes
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(ROOT, "oracle")
REF_DIR = os.path.join(ORACLE_DIR, "_ref")

_u8p = ctypes.POINTER(ctypes.c_uint8)
_i16p = ctypes.POINTER(ctypes.c_int16)
_u32p = ctypes.POINTER(ctypes.c_uint32)
_u64p = ctypes.POINTER(ctypes.c_uint64)
_f32p = ctypes.POINTER(ctypes.c_float)


class _Segs(ctypes.Structure):
    _fields_ = [("nseg", ctypes.c_int), ("npts", ctypes.c_int),
                ("x", _i16p), ("y", _i16p), ("z", _i16p),
                ("seg_start", _u32p), ("is_end", _u8p)]


class _CL(ctypes.Structure):
    _fields_ = [("nseg", ctypes.c_int), ("npts", ctypes.c_int),
                ("x", _f32p), ("y", _f32p), ("z", _f32p),
                ("idx", _u32p), ("iterations", ctypes.c_int)]


def _load(path):
    if not os.path.exists(path):
        return None
    return ctypes.CDLL(path)


_lib = None


def lib():
    global _lib
    if _lib is None:
        path = os.path.join(ORACLE_DIR, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle/liboracle.so is not built (run __graft_entry__.build())")
        L = ctypes.CDLL(path)
        L.or2d_match.restype = ctypes.c_uint64
        L.or2d_match.argtypes = [ctypes.c_int] + [ctypes.c_uint64] * 9
        L.or2d_table.argtypes = [ctypes.c_int, _u8p]
        L.or2d_thin.restype = ctypes.c_int
        L.or2d_thin.argtypes = [_u8p, ctypes.c_int, ctypes.c_int, ctypes.c_int]
        L.or2d_iterate.restype = ctypes.c_int
        L.or2d_iterate.argtypes = [_u8p, ctypes.c_int, ctypes.c_int, ctypes.c_int]
        L.or3d_comb.restype = ctypes.c_uint64
        L.or3d_comb.argtypes = [ctypes.c_int, _u64p]
        L.or3d_table.argtypes = [ctypes.c_int, _u64p]
        L.or3d_thin.restype = ctypes.c_int
        L.or3d_thin.argtypes = [_u8p] + [ctypes.c_int] * 4
        L.or3d_iterate.restype = ctypes.c_int
        L.or3d_iterate.argtypes = [_u8p] + [ctypes.c_int] * 4
        L.or2d_extract.restype = ctypes.c_int
        L.or2d_extract.argtypes = [_u8p, ctypes.c_int, ctypes.c_int, ctypes.POINTER(_Segs)]
        L.or3d_extract.restype = ctypes.c_int
        L.or3d_extract.argtypes = [_u8p, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                   ctypes.POINTER(_Segs)]
        L.or_segments_free.argtypes = [ctypes.POINTER(_Segs)]
        L.or2d_centerline.restype = ctypes.c_int
        L.or2d_centerline.argtypes = [_u8p, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                      ctypes.c_int, ctypes.POINTER(_CL)]
        L.or3d_centerline.restype = ctypes.c_int
        L.or3d_centerline.argtypes = [_u8p, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                      ctypes.c_int, ctypes.c_int, ctypes.POINTER(_CL)]
        L.or_centerline_free.argtypes = [ctypes.POINTER(_CL)]
        L.or_smooth.argtypes = [_i16p, _i16p, _i16p, ctypes.c_int, ctypes.c_int,
                                _f32p, _f32p, _f32p]
        _lib = L
    return _lib


def _ptr(a, t):
    return a.ctypes.data_as(t)


def table2d(direction):
    out = np.zeros(512, np.uint8)
    lib().or2d_table(direction, _ptr(out, _u8p))
    return out


def table3d(direction):
    """Bit-packed 2^26-entry table (uint64[1<<20]) of C_d."""
    out = np.zeros(1 << 20, np.uint64)
    lib().or3d_table(direction, _ptr(out, _u64p))
    return out


def comb3d(direction, words):
    n = np.ascontiguousarray(words, dtype=np.uint64)
    assert n.shape == (26,)
    return int(lib().or3d_comb(direction, _ptr(n, _u64p)))


def thin2d(img, max_iter=0):
    a = np.ascontiguousarray(img, dtype=np.uint8).copy()
    h, w = a.shape
    it = lib().or2d_thin(_ptr(a, _u8p), w, h, max_iter)
    return a, it


def iterate2d(img, n_iter):
    a = np.ascontiguousarray(img, dtype=np.uint8).copy()
    h, w = a.shape
    ch = lib().or2d_iterate(_ptr(a, _u8p), w, h, n_iter)
    return a, ch


def thin3d(vol, max_iter=0):
    a = np.ascontiguousarray(vol, dtype=np.uint8).copy()
    d, h, w = a.shape
    it = lib().or3d_thin(_ptr(a, _u8p), w, h, d, max_iter)
    return a, it


def iterate3d(vol, n_iter):
    a = np.ascontiguousarray(vol, dtype=np.uint8).copy()
    d, h, w = a.shape
    ch = lib().or3d_iterate(_ptr(a, _u8p), w, h, d, n_iter)
    return a, ch


def _segs_to_py(s, three_d):
    n, m = s.nseg, s.npts
    xs = np.ctypeslib.as_array(s.x, (m,)).copy() if m else np.zeros(0, np.int16)
    ys = np.ctypeslib.as_array(s.y, (m,)).copy() if m else np.zeros(0, np.int16)
    zs = (np.ctypeslib.as_array(s.z, (m,)).copy() if m else np.zeros(0, np.int16)) if three_d else None
    st = np.ctypeslib.as_array(s.seg_start, (n + 1,)).copy()
    ie = np.ctypeslib.as_array(s.is_end, (n,)).copy() if n else np.zeros(0, np.uint8)
    return {"x": xs, "y": ys, "z": zs, "start": st, "is_end": ie}


def extract2d(img):
    a = np.ascontiguousarray(img, dtype=np.uint8)
    h, w = a.shape
    s = _Segs()
    lib().or2d_extract(_ptr(a, _u8p), w, h, ctypes.byref(s))
    r = _segs_to_py(s, False)
    lib().or_segments_free(ctypes.byref(s))
    return r


def extract3d(vol):
    a = np.ascontiguousarray(vol, dtype=np.uint8)
    d, h, w = a.shape
    s = _Segs()
    lib().or3d_extract(_ptr(a, _u8p), w, h, d, ctypes.byref(s))
    r = _segs_to_py(s, True)
    lib().or_segments_free(ctypes.byref(s))
    return r


def _cl_to_py(c, three_d):
    n, m = c.nseg, c.npts
    g = lambda p: np.ctypeslib.as_array(p, (m,)).copy() if m else np.zeros(0, np.float32)
    return {"x": g(c.x), "y": g(c.y), "z": g(c.z) if three_d else None,
            "idx": np.ctypeslib.as_array(c.idx, (n + 1,)).copy(), "iterations": c.iterations}


def centerline2d(img, max_segment_length, smoothing_span):
    a = np.ascontiguousarray(img, dtype=np.uint8)
    h, w = a.shape
    c = _CL()
    lib().or2d_centerline(_ptr(a, _u8p), w, h, max_segment_length, smoothing_span, ctypes.byref(c))
    r = _cl_to_py(c, False)
    lib().or_centerline_free(ctypes.byref(c))
    return r


def centerline3d(vol, max_segment_length, smoothing_span):
    a = np.ascontiguousarray(vol, dtype=np.uint8)
    d, h, w = a.shape
    c = _CL()
    lib().or3d_centerline(_ptr(a, _u8p), w, h, d, max_segment_length, smoothing_span,
                          ctypes.byref(c))
    r = _cl_to_py(c, True)
    lib().or_centerline_free(ctypes.byref(c))
    return r


def smooth(px, py, pz, span):
    px = np.ascontiguousarray(px, np.int16)
    py = np.ascontiguousarray(py, np.int16)
    n = len(px)
    ox = np.zeros(n, np.float32); oy = np.zeros(n, np.float32)
    if pz is not None:
        pz = np.ascontiguousarray(pz, np.int16)
        oz = np.zeros(n, np.float32)
        lib().or_smooth(_ptr(px, _i16p), _ptr(py, _i16p), _ptr(pz, _i16p), n, span,
                        _ptr(ox, _f32p), _ptr(oy, _f32p), _ptr(oz, _f32p))
        return ox, oy, oz
    lib().or_smooth(_ptr(px, _i16p), _ptr(py, _i16p), None, n, span,
                    _ptr(ox, _f32p), _ptr(oy, _f32p), None)
    return ox, oy, None


# --- reference builds (present only where oracle/_ref was built) -----------

def ref_available(name):
    return os.path.exists(os.path.join(REF_DIR, name))


def ref_table3d(direction):
    L = ctypes.CDLL(os.path.join(REF_DIR, "libref_pred3d.so"))
    L.ref3d_table.argtypes = [ctypes.c_int, _u32p]
    out = np.zeros(1 << 21, np.uint32)
    L.ref3d_table(direction, _ptr(out, _u32p))
    return out


def ref_extract2d(img):
    L = ctypes.CDLL(os.path.join(REF_DIR, "libref2d.so"))
    L.ref2d_extract.restype = ctypes.c_int
    L.ref2d_extract.argtypes = [_u8p, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                _i16p, _i16p, _u32p, _u8p, ctypes.POINTER(ctypes.c_int)]
    a = np.ascontiguousarray(img, dtype=np.uint8)
    h, w = a.shape
    cnt = int(a.astype(bool).sum())
    cap_p, cap_s = cnt * 20 + 16, cnt * 10 + 16
    xs = np.zeros(cap_p, np.int16); ys = np.zeros(cap_p, np.int16)
    st = np.zeros(cap_s + 1, np.uint32); ie = np.zeros(cap_s, np.uint8)
    npts = ctypes.c_int(0)
    ns = L.ref2d_extract(_ptr(a, _u8p), w, h, cap_p, cap_s, _ptr(xs, _i16p), _ptr(ys, _i16p),
                         _ptr(st, _u32p), _ptr(ie, _u8p), ctypes.byref(npts))
    assert ns >= 0, "capacity too small"
    m = npts.value
    return {"x": xs[:m], "y": ys[:m], "z": None, "start": st[:ns + 1], "is_end": ie[:ns]}


def ref_extract3d(vol):
    L = ctypes.CDLL(os.path.join(REF_DIR, "libref3d.so"))
    L.ref3d_extract.restype = ctypes.c_int
    L.ref3d_extract.argtypes = [_u8p, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                ctypes.c_int, _i16p, _i16p, _i16p, _u32p, _u8p,
                                ctypes.POINTER(ctypes.c_int)]
    a = np.ascontiguousarray(vol, dtype=np.uint8)
    d, h, w = a.shape
    cnt = int(a.astype(bool).sum())
    cap_p, cap_s = cnt * 60 + 16, cnt * 30 + 16
    xs = np.zeros(cap_p, np.int16); ys = np.zeros(cap_p, np.int16); zs = np.zeros(cap_p, np.int16)
    st = np.zeros(cap_s + 1, np.uint32); ie = np.zeros(cap_s, np.uint8)
    npts = ctypes.c_int(0)
    ns = L.ref3d_extract(_ptr(a, _u8p), w, h, d, cap_p, cap_s, _ptr(xs, _i16p), _ptr(ys, _i16p),
                         _ptr(zs, _i16p), _ptr(st, _u32p), _ptr(ie, _u8p), ctypes.byref(npts))
    assert ns >= 0, "capacity too small"
    m = npts.value
    return {"x": xs[:m], "y": ys[:m], "z": zs[:m], "start": st[:ns + 1], "is_end": ie[:ns]}


def ref_extract3d_plain(vol):
    """The reference's ExtractSegments (point lists, no end flags)."""
    L = ctypes.CDLL(os.path.join(REF_DIR, "libref3d.so"))
    L.ref3d_extract_plain.restype = ctypes.c_int
    L.ref3d_extract_plain.argtypes = [_u8p, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                      ctypes.c_int, _i16p, _i16p, _i16p, _u32p,
                                      ctypes.POINTER(ctypes.c_int)]
    a = np.ascontiguousarray(vol, dtype=np.uint8)
    d, h, w = a.shape
    cnt = int(a.astype(bool).sum())
    cap_p, cap_s = cnt * 60 + 16, cnt * 30 + 16
    xs = np.zeros(cap_p, np.int16); ys = np.zeros(cap_p, np.int16); zs = np.zeros(cap_p, np.int16)
    st = np.zeros(cap_s + 1, np.uint32)
    npts = ctypes.c_int(0)
    ns = L.ref3d_extract_plain(_ptr(a, _u8p), w, h, d, cap_p, cap_s, _ptr(xs, _i16p),
                               _ptr(ys, _i16p), _ptr(zs, _i16p), _ptr(st, _u32p),
                               ctypes.byref(npts))
    assert ns >= 0, "capacity too small"
    m = npts.value
    return {"x": xs[:m], "y": ys[:m], "z": zs[:m], "start": st[:ns + 1]}

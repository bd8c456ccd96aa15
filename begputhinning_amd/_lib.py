"""ctypes binding of the C-ABI in include/bgt.h (libbgt.so, built in-tree).

Loading fails loudly when the HIP library has not been built: there is no CPU
fallback anywhere on the product path.
"""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
# BGT_LIBRARY names another build of the same library (a sanitizer build, say)
LIB_PATH = os.environ.get("BGT_LIBRARY") or os.path.join(HERE, "libbgt.so")

BGT_OK = 0
BGT_ERR_INVALID = 1
BGT_ERR_HIP = 2
BGT_ERR_NOMEM = 3
BGT_ERR_UNSUPPORTED = 4
BGT_ERR_NOT_CONVERGED = 5
BGT_ERR_ADMISSION = 6

DTYPES = {"uint8": 0, "int8": 1, "uint16": 2, "int16": 3, "uint32": 4, "int32": 5,
          "float32": 6, "float64": 7, "bool": 8}


class BGTError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"bgt error {code}: {msg}")
        self.code = code


class Segments(ctypes.Structure):
    _fields_ = [("nseg", ctypes.c_int), ("npts", ctypes.c_int),
                ("x", ctypes.POINTER(ctypes.c_int16)), ("y", ctypes.POINTER(ctypes.c_int16)),
                ("z", ctypes.POINTER(ctypes.c_int16)),
                ("seg_start", ctypes.POINTER(ctypes.c_uint32)),
                ("is_end", ctypes.POINTER(ctypes.c_uint8))]


_lib = None

_vp = ctypes.c_void_p
_i = ctypes.c_int
_sz = ctypes.c_size_t
_ip = ctypes.POINTER(ctypes.c_int)
_fpp = ctypes.POINTER(ctypes.POINTER(ctypes.c_float))
_upp = ctypes.POINTER(ctypes.POINTER(ctypes.c_uint32))
_up = ctypes.POINTER(ctypes.c_uint32)

SIGNATURES = {
    "bgt_last_error": (ctypes.c_char_p, []),
    "bgt_version": (_i, []),
    "bgt_thinning2d": (_i, [_i, _i, _i, _vp, _vp]),
    "bgt_thinning3d": (_i, [_i, _i, _i, _i, _vp, _vp]),
    "bgt_thin2d_device": (_i, [_vp, _sz, _i, _i, _vp, _sz, _vp, _ip]),
    "bgt_thin3d_device": (_i, [_vp, _i, _i, _i, _vp, _vp, _ip]),
    "bgt_thin2d_device_noalloc": (_i, [_vp, _sz, _i, _i, _vp, _sz, _vp, _ip]),
    "bgt_thin2d_noalloc_max_bytes": (_sz, []),
    "bgt_extract_segments2d": (_i, [_i, _vp, _i, _i, ctypes.POINTER(Segments)]),
    "bgt_extract_segments3d": (_i, [_i, _vp, _i, _i, _i, ctypes.POINTER(Segments)]),
    "bgt_extract_segments2d_device": (_i, [_vp, _i, _i, _i, _i, _vp, ctypes.POINTER(Segments)]),
    "bgt_extract_segments3d_device": (_i, [_vp, _i, _i, _i, _i, _i, _i, _vp,
                                           ctypes.POINTER(Segments)]),
    "bgt_segments_free": (None, [ctypes.POINTER(Segments)]),
    "bgt_centerline2d": (_i, [_i, _vp, _i, _i, _i, _i, _fpp, _fpp, _upp, _up]),
    "bgt_centerline3d": (_i, [_i, _vp, _i, _i, _i, _i, _i, _fpp, _fpp, _fpp, _upp, _up]),
    "bgt_free": (None, [_vp]),
    "bgt_mex_columns": (_i, [_vp, _vp, _vp, _vp, ctypes.c_uint32, _vp]),
    "bgt_synth_item_seed": (ctypes.c_uint64, [_i, _i]),
    "bgt_synth_polylines": (_i, [_vp, _i, _i, ctypes.c_uint64, _i, ctypes.c_double]),
    "bgt_synth_tree2d": (_i, [_vp, _i, _i, ctypes.c_uint64, _i, ctypes.c_double, ctypes.c_double]),
    "bgt_synth_tree3d": (_i, [_vp, _i, _i, _i, ctypes.c_uint64, _i, _i, ctypes.c_double]),
    "bgt_set_config2d": (_i, [_i, _i, _i, _i]),
    "bgt_set_config3d": (_i, [_i, _i, _i]),
    "bgt_set_engine3d": (_i, [_i]),
    "bgt_set_lane_words2d": (_i, [_i]),
    "bgt_set_fused_io2d": (_i, [_i]),
    "bgt_set_persistent2d": (_i, [_i]),
    "bgt_debug_fail_device_allocs": (_i, [_i]),
    "bgt_batch_create": (_i, [ctypes.POINTER(ctypes.c_void_p)]),
    "bgt_batch_use_current_config": (_i, [_vp]),
    "bgt_batch_add2d": (_i, [_vp, _vp, _sz, _i, _i, _vp, _sz]),
    "bgt_batch_add3d": (_i, [_vp, _vp, _i, _i, _i, _vp]),
    "bgt_batch_run": (_i, [_vp, _vp, _i]),
    "bgt_batch_iterations": (_i, [_vp, _ip]),
    "bgt_batch_device_bytes": (_i, [_vp, ctypes.POINTER(ctypes.c_size_t)]),
    "bgt_batch_destroy": (_i, [_vp]),
    "bgt_batch_autotune": (_i, [_vp, _vp, _i, _ip, _ip, _ip, _ip, _ip,
                                ctypes.POINTER(ctypes.c_double)]),
    "bgt_batch_set_timing": (_i, [_vp, _i]),
    "bgt_batch_kernel_time": (_i, [_vp, ctypes.POINTER(ctypes.c_double), _ip,
                                   ctypes.POINTER(ctypes.c_double), _ip]),
    "bgt_server_create": (_i, [ctypes.POINTER(ctypes.c_void_p), _sz, _i, _i, _i]),
    "bgt_server_submit2d": (_i, [_vp, _vp, _sz, _i, _i, _vp, _sz, ctypes.POINTER(ctypes.c_uint64)]),
    "bgt_server_submit3d": (_i, [_vp, _vp, _i, _i, _i, _vp, ctypes.POINTER(ctypes.c_uint64)]),
    "bgt_server_wait": (_i, [_vp, ctypes.c_uint64, _ip, ctypes.POINTER(ctypes.c_double)]),
    "bgt_server_enable_segments": (_i, [_vp, _i]),
    "bgt_server_wait_segments": (_i, [_vp, ctypes.c_uint64, _ip, ctypes.POINTER(ctypes.c_double),
                                      ctypes.POINTER(ctypes.c_double), ctypes.POINTER(Segments)]),
    "bgt_server_poll": (_i, [_vp, ctypes.c_uint64, _ip]),
    "bgt_server_stats": (_i, [_vp, ctypes.POINTER(ctypes.c_uint64), ctypes.POINTER(ctypes.c_uint64),
                              ctypes.POINTER(ctypes.c_size_t)]),
    "bgt_server_destroy": (_i, [_vp]),
    "bgt_synth_batch_dims": (_i, [ctypes.c_uint64, _i, _i, _i, _ip, _i, _i, _i, _ip]),
    "bgt_synth_arrivals": (_i, [ctypes.c_uint64, _i, ctypes.c_double,
                                ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double)]),
}

# entry points added by later subsystems register their signatures here
EXTRA_SIGNATURES = {}


def lib():
    """The loaded HIP library; raises if it is not built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"{LIB_PATH} is missing: the HIP library must be built first "
                "(python -m begputhinning_amd.build); there is no CPU fallback")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in list(SIGNATURES.items()) + list(EXTRA_SIGNATURES.items()):
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def check(code):
    if code != BGT_OK:
        msg = lib().bgt_last_error()
        raise BGTError(code, msg.decode() if msg else "")
    return code


def check_device_pair(inp, out, ndim, what):
    """Validates a device-resident (input, output) pair before any kernel
    runs: both uint8 tensors on the current GPU, `ndim` dimensions, the same
    shape, rows contiguous (2D) or fully contiguous (3D).  The library writes
    exactly w*h(*d) bytes through `out` at the given pitch, so a smaller,
    differently typed or host `out` would be written past its end."""
    import torch
    for name, t in (("input", inp), ("out", out)):
        if not isinstance(t, torch.Tensor):
            raise TypeError(f"{what}: {name} must be a torch tensor")
        if not t.is_cuda:
            raise TypeError(f"{what}: {name} must be a GPU tensor (got device {t.device})")
        if t.dtype != torch.uint8:
            raise TypeError(f"{what}: {name} must be uint8 (got {t.dtype})")
        if t.dim() != ndim:
            raise ValueError(f"{what}: {name} must have {ndim} dimensions (got {t.dim()})")
    if out.device != inp.device or inp.device.index != torch.cuda.current_device():
        raise ValueError(f"{what}: input and out must be on the current GPU "
                         f"(cuda:{torch.cuda.current_device()}), got {inp.device} and {out.device}")
    if tuple(out.shape) != tuple(inp.shape):
        raise ValueError(f"{what}: out shape {tuple(out.shape)} != input shape {tuple(inp.shape)}")
    if ndim == 2:
        for name, t in (("input", inp), ("out", out)):
            if t.stride(1) != 1 or (t.shape[0] > 1 and t.stride(0) < t.shape[1]):
                raise ValueError(f"{what}: {name} rows must be contiguous")
    else:
        for name, t in (("input", inp), ("out", out)):
            if not t.is_contiguous():
                raise ValueError(f"{what}: {name} must be contiguous")

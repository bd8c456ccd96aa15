"""Batched thinning of device-resident images and volumes of uneven size
(bgt_batch_* in include/bgt.h)."""
import ctypes

import numpy as np

from ._lib import check, check_device_pair, lib


class ThinningBatch:
    """Collects (input, output) uint8 GPU tensors and thins them together.

    add2d(img, out) takes (h, w) tensors with contiguous rows; add3d(vol, out)
    contiguous (d, h, w) tensors.  run() thins everything; iterations() gives
    the per-item iteration counts in the order added."""

    def __init__(self):
        h = ctypes.c_void_p()
        check(lib().bgt_batch_create(ctypes.byref(h)))
        self._h = h
        self._keep = []
        self.n = 0

    def use_current_config(self):
        check(lib().bgt_batch_use_current_config(self._h))

    def add2d(self, img, out):
        check_device_pair(img, out, 2, "ThinningBatch.add2d")
        h, w = img.shape
        check(lib().bgt_batch_add2d(self._h, ctypes.c_void_p(img.data_ptr()), img.stride(0), w, h,
                                    ctypes.c_void_p(out.data_ptr()), out.stride(0)))
        self._keep.append((img, out))
        self.n += 1

    def add3d(self, vol, out):
        check_device_pair(vol, out, 3, "ThinningBatch.add3d")
        d, h, w = vol.shape
        check(lib().bgt_batch_add3d(self._h, ctypes.c_void_p(vol.data_ptr()), w, h, d,
                                    ctypes.c_void_p(out.data_ptr())))
        self._keep.append((vol, out))
        self.n += 1

    def run(self, stream=None, use_graph=True):
        import torch
        s = stream if stream is not None else torch.cuda.current_stream()
        check(lib().bgt_batch_run(self._h, ctypes.c_void_p(s.cuda_stream), 1 if use_graph else 0))

    def autotune(self, stream=None, reps=3):
        """Pick the fastest launch shapes for this batch (the outputs hold a
        complete result afterwards); returns the chosen shapes and time."""
        import torch
        st = stream if stream is not None else torch.cuda.current_stream()
        v = [ctypes.c_int(0) for _ in range(5)]
        ms = ctypes.c_double(0.0)
        check(lib().bgt_batch_autotune(self._h, ctypes.c_void_p(st.cuda_stream), reps,
                                       *[ctypes.byref(x) for x in v], ctypes.byref(ms)))
        return {"waves": v[0].value, "rows": v[1].value, "iters_per_pass": v[2].value,
                "blocks_per_cu_2d": v[3].value, "blocks_per_cu_3d": v[4].value, "ms": ms.value}

    def iterations(self):
        out = np.zeros(self.n, np.int32)
        check(lib().bgt_batch_iterations(self._h, out.ctypes.data_as(ctypes.POINTER(ctypes.c_int))))
        return out

    def device_bytes(self):
        v = ctypes.c_size_t(0)
        check(lib().bgt_batch_device_bytes(self._h, ctypes.byref(v)))
        return v.value

    def set_timing(self, on=True):
        check(lib().bgt_batch_set_timing(self._h, 1 if on else 0))

    def kernel_time(self):
        """(ms_2d, launches_2d, ms_3d, launches_3d) of the last timed run."""
        a, c = ctypes.c_double(0), ctypes.c_double(0)
        b, d = ctypes.c_int(0), ctypes.c_int(0)
        check(lib().bgt_batch_kernel_time(self._h, ctypes.byref(a), ctypes.byref(b),
                                          ctypes.byref(c), ctypes.byref(d)))
        return a.value, b.value, c.value, d.value

    def close(self):
        if self._h:
            lib().bgt_batch_destroy(self._h)
            self._h = None
            self._keep = []

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

"""Serving: device-resident requests batched as they arrive (bgt_server_* in
include/bgt.h).

    srv = ThinningServer(budget_bytes=8 << 30)
    t = srv.submit(img_gpu, out_gpu)        # (h, w) or (d, h, w) uint8 CUDA tensors
    iterations, latency_ms = srv.wait(t)

Inputs must be complete on the device when submitted (synchronize the stream
that produced them); outputs are complete when wait() returns.
"""
import ctypes

from ._lib import check, check_device_pair, lib


class ThinningServer:
    def __init__(self, budget_bytes=8 << 30, workers2d=2, workers3d=1, max_items=64):
        h = ctypes.c_void_p()
        check(lib().bgt_server_create(ctypes.byref(h), int(budget_bytes), workers2d, workers3d,
                                      max_items))
        self._h = h
        self._keep = {}

    def submit(self, inp, out):
        """Queue one image ((h, w), contiguous rows) or volume ((d, h, w),
        contiguous); returns a ticket."""
        t = ctypes.c_uint64(0)
        ndim = inp.dim() if hasattr(inp, "dim") else -1
        if ndim not in (2, 3):
            raise ValueError("requests are 2D or 3D uint8 GPU tensors")
        check_device_pair(inp, out, ndim, "ThinningServer.submit")
        if ndim == 2:
            hh, w = inp.shape
            check(lib().bgt_server_submit2d(self._h, ctypes.c_void_p(inp.data_ptr()), inp.stride(0),
                                            w, hh, ctypes.c_void_p(out.data_ptr()), out.stride(0),
                                            ctypes.byref(t)))
        else:
            d, hh, w = inp.shape
            check(lib().bgt_server_submit3d(self._h, ctypes.c_void_p(inp.data_ptr()), w, hh, d,
                                            ctypes.c_void_p(out.data_ptr()), ctypes.byref(t)))
        self._keep[t.value] = (inp, out)
        return t.value

    def wait(self, ticket):
        """Block until the request is done: (iterations, submit-to-done ms)."""
        it = ctypes.c_int(0)
        ms = ctypes.c_double(0.0)
        try:
            check(lib().bgt_server_wait(self._h, ctypes.c_uint64(ticket), ctypes.byref(it),
                                        ctypes.byref(ms)))
        finally:
            self._keep.pop(ticket, None)
        return it.value, ms.value

    def enable_segments(self, threads=4):
        """Have the server trace every served skeleton into its segment list
        (`threads` tracer threads; before the first submit).  wait() then
        also waits for the trace."""
        check(lib().bgt_server_enable_segments(self._h, int(threads)))

    def wait_segments(self, ticket):
        """Block until the request's segment list is ready: (iterations,
        submit-to-skeleton ms, submit-to-segments ms, segments) with the
        segments as extract_segments_flat gives them."""
        from ._lib import Segments
        from .thinning import _take_segments
        it = ctypes.c_int(0)
        ms = ctypes.c_double(0.0)
        sms = ctypes.c_double(0.0)
        seg = Segments()
        inp = self._keep.get(ticket, (None,))[0]
        try:
            check(lib().bgt_server_wait_segments(self._h, ctypes.c_uint64(ticket), ctypes.byref(it),
                                                 ctypes.byref(ms), ctypes.byref(sms),
                                                 ctypes.byref(seg)))
        finally:
            self._keep.pop(ticket, None)
        three_d = inp is not None and inp.dim() == 3
        return it.value, ms.value, sms.value, _take_segments(seg, three_d)

    def done(self, ticket):
        d = ctypes.c_int(0)
        check(lib().bgt_server_poll(self._h, ctypes.c_uint64(ticket), ctypes.byref(d)))
        return bool(d.value)

    def stats(self):
        b, c, p = ctypes.c_uint64(0), ctypes.c_uint64(0), ctypes.c_size_t(0)
        check(lib().bgt_server_stats(self._h, ctypes.byref(b), ctypes.byref(c), ctypes.byref(p)))
        return {"batches": b.value, "completed": c.value, "peak_bytes": p.value}

    def close(self):
        if self._h:
            check(lib().bgt_server_destroy(self._h))
            self._h = None
            self._keep.clear()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

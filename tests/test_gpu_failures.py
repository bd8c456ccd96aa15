"""Failure handling on the GPU, forced from the host: device allocations made
to fail through the library's hook (bgt_debug_fail_device_allocs) return
BGT_ERR_NOMEM from the call that needed them, leave no half-made state (the
next call succeeds and equals the oracle), and do not touch the no-allocation
entry point; a poisoned item of a batch (too wide, or of zero size) is
rejected when added while the batch's other items still thin correctly; a
served request whose memory cannot be had fails alone while the others
complete."""
import threading

import numpy as np
import pytest

import _oracle as O

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


@pytest.fixture(scope="module", autouse=True)
def _gpu():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import begputhinning_amd as B
    yield
    B.debug_fail_device_allocs(0)


def _in_thread(fn):
    """Run fn in a fresh thread (its per-thread device workspace starts
    empty, so the call has to allocate); re-raise what it raised."""
    box = {}

    def run():
        try:
            box["v"] = fn()
        except BaseException as e:  # noqa: BLE001 -- handed to the caller
            box["e"] = e
    t = threading.Thread(target=run)
    t.start()
    t.join()
    if "e" in box:
        raise box["e"]
    return box.get("v")


def test_per_image_allocation_failure():
    import begputhinning_amd as B
    from begputhinning_amd import synth
    from begputhinning_amd._lib import BGT_ERR_NOMEM
    img = synth.polylines(300, 200, 17)
    t = torch.from_numpy(img).cuda()
    B.debug_fail_device_allocs(1)
    with pytest.raises(B.BGTError) as e:
        _in_thread(lambda: B.thin2d_device(t))
    assert e.value.code == BGT_ERR_NOMEM
    B.debug_fail_device_allocs(0)
    out, it = _in_thread(lambda: B.thin2d_device(t))
    torch.cuda.synchronize()
    ref, rit = O.thin2d(img)
    assert np.array_equal(out.cpu().numpy(), ref) and it == rit


def test_batch_allocation_failure_then_success():
    import begputhinning_amd as B
    from begputhinning_amd import synth
    from begputhinning_amd._lib import BGT_ERR_NOMEM
    from begputhinning_amd.batch import ThinningBatch
    items = [synth.polylines(200 + 37 * k, 150 + 11 * k, 50 + k) for k in range(5)]
    items.append(synth.tree3d(40, 30, 20, 8))
    ins = [torch.from_numpy(a).cuda() for a in items]
    outs = [torch.full_like(x, 9) for x in ins]
    b = ThinningBatch()
    for x, o in zip(ins, outs):
        (b.add2d if x.dim() == 2 else b.add3d)(x, o)
    B.debug_fail_device_allocs(1)
    with pytest.raises(B.BGTError) as e:
        b.run(use_graph=False)
    assert e.value.code == BGT_ERR_NOMEM
    B.debug_fail_device_allocs(0)
    b.run(use_graph=False)
    torch.cuda.synchronize()
    its = b.iterations()
    for k, a in enumerate(items):
        ref, rit = O.thin2d(a) if a.ndim == 2 else O.thin3d(a)
        assert np.array_equal(outs[k].cpu().numpy(), ref) and its[k] == rit, k
    b.close()


def test_noalloc_entry_needs_no_allocation():
    import ctypes
    import begputhinning_amd as B
    from begputhinning_amd import synth
    from begputhinning_amd._lib import check, lib
    img = synth.polylines(256, 120, 23)
    t = torch.from_numpy(img).cuda()
    o = torch.empty_like(t)
    torch.cuda.synchronize()
    B.debug_fail_device_allocs(1000)
    try:
        it = ctypes.c_int(0)
        s = torch.cuda.current_stream().cuda_stream
        check(lib().bgt_thin2d_device_noalloc(ctypes.c_void_p(t.data_ptr()), t.stride(0), 256, 120,
                                              ctypes.c_void_p(o.data_ptr()), o.stride(0),
                                              ctypes.c_void_p(s), ctypes.byref(it)))
        torch.cuda.synchronize()
    finally:
        B.debug_fail_device_allocs(0)
    ref, rit = O.thin2d(img)
    assert np.array_equal(o.cpu().numpy(), ref) and it.value == rit


def test_poisoned_items_leave_the_batch_intact():
    import begputhinning_amd as B
    from begputhinning_amd import synth
    from begputhinning_amd.batch import ThinningBatch
    good = [synth.polylines(310, 170, 61), synth.tree3d(30, 40, 25, 62), synth.polylines(90, 500, 63)]
    ins = [torch.from_numpy(a).cuda() for a in good]
    outs = [torch.full_like(x, 9) for x in ins]
    b = ThinningBatch()
    b.add2d(ins[0], outs[0])
    wide = torch.zeros((3, 16385), dtype=torch.uint8, device="cuda")
    with pytest.raises(B.BGTError):
        b.add2d(wide, torch.empty_like(wide))  # wider than any 2D image
    b.add3d(ins[1], outs[1])
    empty = torch.zeros((0, 7), dtype=torch.uint8, device="cuda")
    with pytest.raises((B.BGTError, ValueError)):
        b.add2d(empty, torch.empty_like(empty))
    b.add2d(ins[2], outs[2])
    b.run(use_graph=False)
    torch.cuda.synchronize()
    its = b.iterations()
    assert len(its) == 3
    for k, a in enumerate(good):
        ref, rit = O.thin2d(a) if a.ndim == 2 else O.thin3d(a)
        assert np.array_equal(outs[k].cpu().numpy(), ref) and its[k] == rit, k
    b.close()


def test_served_request_fails_alone():
    import begputhinning_amd as B
    from begputhinning_amd import synth
    from begputhinning_amd._lib import BGT_ERR_NOMEM
    from begputhinning_amd.serve import ThinningServer
    frames = [synth.polylines(200 + 20 * k, 180, 70 + k) for k in range(4)]
    vol = synth.tree3d(48, 40, 36, 71)
    fins = [torch.from_numpy(a).cuda() for a in frames]
    fouts = [torch.full_like(x, 9) for x in fins]
    vin = torch.from_numpy(vol).cuda()
    vout = torch.full_like(vin, 9)
    torch.cuda.synchronize()
    srv = ThinningServer(budget_bytes=2 << 30, workers2d=1, workers3d=1, max_items=8)
    try:
        # the server is up (its frame queue allocated); the volume's worker
        # has not allocated yet, so the volume's first allocation fails
        B.debug_fail_device_allocs(1)
        tv = srv.submit(vin, vout)
        with pytest.raises(B.BGTError) as e:
            srv.wait(tv)
        assert e.value.code == BGT_ERR_NOMEM
        B.debug_fail_device_allocs(0)
        tks = [srv.submit(x, o) for x, o in zip(fins, fouts)]
        tv2 = srv.submit(vin, vout)
        its = [srv.wait(t)[0] for t in tks]
        it3 = srv.wait(tv2)[0]
    finally:
        B.debug_fail_device_allocs(0)
        srv.close()
    for k, a in enumerate(frames):
        ref, rit = O.thin2d(a)
        assert np.array_equal(fouts[k].cpu().numpy(), ref) and its[k] == rit, k
    ref3, rit3 = O.thin3d(vol)
    assert np.array_equal(vout.cpu().numpy(), ref3) and it3 == rit3

"""Segment lists traced by the server itself (bgt_server_enable_segments):
each served skeleton's list equals the oracle's tracer on the oracle's
skeleton of the same input (frames in the running 2D queue and on the
workers, volumes), segments-ready comes no earlier than skeleton-ready, a
plain wait on a tracing server still works, and a pitched 2D output (which
the device tracer cannot read) fails alone with BGT_ERR_UNSUPPORTED."""
import numpy as np
import pytest

import _oracle as O

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


@pytest.fixture(scope="module", autouse=True)
def _gpu():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")


def _same(got, ref, three_d):
    keys = ("x", "y", "z", "start", "is_end") if three_d else ("x", "y", "start", "is_end")
    for k in keys:
        assert np.array_equal(np.asarray(got[k]), np.asarray(ref[k])), k


def test_served_segments_equal_the_oracle():
    from begputhinning_amd import synth
    from begputhinning_amd.serve import ThinningServer
    rng = np.random.default_rng(21)
    items = []
    for k in range(10):
        w, h = (int(v) for v in rng.integers(64, 700, 2))
        items.append(synth.tree2d(w, h, 300 + k) if k % 2 else synth.polylines(w, h, 300 + k))
    items.append(synth.polylines(2300, 90, 399))  # wider than the running queue's slots: a worker
    items += [synth.tree3d(50, 40, 30, 310), synth.tree3d(70, 60, 45, 311)]
    ins = [torch.from_numpy(a).cuda() for a in items]
    outs = [torch.full_like(x, 9) for x in ins]
    torch.cuda.synchronize()
    srv = ThinningServer(budget_bytes=4 << 30, workers2d=2, workers3d=1, max_items=8)
    try:
        srv.enable_segments(3)
        tks = [srv.submit(x, o) for x, o in zip(ins, outs)]
        res = [srv.wait_segments(t) for t in tks]
        extra = srv.submit(ins[0], outs[0])  # a plain wait on a tracing server
        it0, _ = srv.wait(extra)
    finally:
        srv.close()
    for k, a in enumerate(items):
        three_d = a.ndim == 3
        ref, rit = (O.thin3d(a) if three_d else O.thin2d(a))
        its, lat, seg_ms, segs = res[k]
        assert np.array_equal(outs[k].cpu().numpy(), ref) and its == rit, k
        assert seg_ms >= lat >= 0, (k, lat, seg_ms)
        _same(segs, O.extract3d(ref) if three_d else O.extract2d(ref), three_d)
    assert it0 == O.thin2d(items[0])[1]


def test_pitched_output_trace_fails_alone():
    from begputhinning_amd import synth
    from begputhinning_amd._lib import BGT_ERR_UNSUPPORTED
    import begputhinning_amd as B
    from begputhinning_amd.serve import ThinningServer
    a = synth.polylines(200, 150, 401)
    b = synth.polylines(180, 120, 402)
    ta, tb = torch.from_numpy(a).cuda(), torch.from_numpy(b).cuda()
    big = torch.zeros((150, 256), dtype=torch.uint8, device="cuda")
    oa = big[:, :200]  # pitched view
    ob = torch.empty_like(tb)
    torch.cuda.synchronize()
    srv = ThinningServer(budget_bytes=2 << 30, workers2d=1, workers3d=0, max_items=4)
    try:
        srv.enable_segments(1)
        t1, t2 = srv.submit(ta, oa), srv.submit(tb, ob)
        with pytest.raises(B.BGTError) as e:
            srv.wait_segments(t1)
        assert e.value.code == BGT_ERR_UNSUPPORTED
        it, _, _, segs = srv.wait_segments(t2)
    finally:
        srv.close()
    ref, rit = O.thin2d(b)
    assert it == rit and np.array_equal(ob.cpu().numpy(), ref)
    _same(segs, O.extract2d(ref), False)
    assert np.array_equal(oa.cpu().numpy(), O.thin2d(a)[0])  # the skeleton itself was served

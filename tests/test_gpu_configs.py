"""The benchmark configurations at their full sizes on the MI355X, bit-exact
against the oracle:

- the mixed batch (512 images with sides U{128..2048}, 16 volumes with sides
  U{64..384}) through one ThinningBatch, every item's skeleton hash and
  iteration count equal to the oracle-written tests/golden/workload4.json;
- 384^3 and 512^3 vessel-tree volumes through bgt_thin3d_device;
- the 256^3 volume's CenterlineExtraction (float32 outputs compared as bits);
- serving-sized requests (1024x1024 frames of both contents, a 384^3 volume,
  frames and volumes of the serving ranges) submitted together to the server.
"""
import hashlib
import json
import os

import numpy as np
import pytest

import _oracle as O

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "workload4.json")


@pytest.fixture(scope="module", autouse=True)
def _gpu():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()[:16]


def test_mixed_batch_full_workload_matches_golden():
    from concurrent.futures import ThreadPoolExecutor
    from begputhinning_amd import synth
    from begputhinning_amd.batch import ThinningBatch
    gold = json.load(open(GOLD))
    d2, d3 = synth.batch_dims()
    n = len(d2) + len(d3)
    assert n == len(gold["skeleton_sha"]) == 528
    with ThreadPoolExecutor(8) as ex:
        items = list(ex.map(lambda i: synth.workload4_item(i, d2, d3), range(n)))
    assert [sha(a) for a in items] == gold["input_sha"], "generator drift"
    print(f"[configs] mixed batch: {n} items, {sum(a.size for a in items) / 1e6:.1f} Mvoxel",
          flush=True)
    ins = [torch.from_numpy(a).cuda() for a in items]
    outs = [torch.full_like(t, 7) for t in ins]
    b = ThinningBatch()
    for t, o in zip(ins, outs):
        (b.add2d if t.dim() == 2 else b.add3d)(t, o)
    for rep in range(2):  # a second run of the same batch re-reads the inputs
        for o in outs:
            o.fill_(7)
        b.run(use_graph=bool(rep))
        torch.cuda.synchronize()
        its = b.iterations()
        bad = [k for k in range(n) if sha(outs[k].cpu().numpy()) != gold["skeleton_sha"][k]
               or int(its[k]) != gold["iterations"][k]]
        assert not bad, f"run {rep}: items {bad[:10]} differ from the oracle's hashes"
    b.close()


@pytest.mark.parametrize("side", [384, 512])
def test_large_volume_vs_oracle(side):
    import begputhinning_amd as B
    from begputhinning_amd import synth
    vol = synth.tree3d(side, side, side, synth.item_seed(3, side))
    print(f"[configs] {side}^3 volume", flush=True)
    t = torch.from_numpy(vol).cuda()
    out, it = B.thin3d_device(t)
    torch.cuda.synchronize()
    ref, rit = O.thin3d(vol)
    assert it == rit
    assert np.array_equal(out.cpu().numpy(), ref)


def test_centerline3d_256_vs_oracle():
    import begputhinning_amd as B
    from begputhinning_amd import synth
    vol = synth.workload3()
    print("[configs] CenterlineExtraction 256^3", flush=True)
    ref = O.centerline3d(vol, 10, 3)
    x, y, z, idx = B.CenterlineExtraction(vol, 10, 3)
    assert np.array_equal(idx, ref["idx"])
    for a, r in ((x, ref["x"]), (y, ref["y"]), (z, ref["z"])):
        assert a.dtype == np.float32 and np.array_equal(a.view(np.uint32), r.view(np.uint32))


def test_server_serving_sized_requests_vs_oracle():
    from begputhinning_amd import synth
    from begputhinning_amd.serve import ThinningServer
    rng = np.random.default_rng(0xC5)
    items = [synth.tree2d(1024, 1024, synth.item_seed(5, 0)),
             synth.polylines(1024, 1024, synth.item_seed(5, 1)),
             synth.tree3d(384, 384, 384, synth.item_seed(5, 2))]
    for i in range(8):
        w, h = (int(v) for v in rng.integers(256, 1025, 2))
        items.append((synth.polylines if i % 2 else synth.tree2d)(w, h, synth.item_seed(5, 10 + i)))
    for i in range(2):
        w, h, d = (int(v) for v in rng.integers(64, 257, 3))
        items.append(synth.tree3d(w, h, d, synth.item_seed(5, 20 + i)))
    print(f"[configs] server: {len(items)} requests", flush=True)
    ins = [torch.from_numpy(a).cuda() for a in items]
    outs = [torch.full_like(t, 3) for t in ins]
    torch.cuda.synchronize()
    srv = ThinningServer(budget_bytes=16 << 30, workers2d=2, workers3d=2, max_items=64)
    try:
        tickets = [srv.submit(t, o) for t, o in zip(ins, outs)]
        its = [srv.wait(tk)[0] for tk in tickets]
    finally:
        srv.close()
    for k, a in enumerate(items):
        ref, rit = O.thin2d(a) if a.ndim == 2 else O.thin3d(a)
        assert its[k] == rit, k
        assert np.array_equal(outs[k].cpu().numpy(), ref), k


def test_centerline2d_large_parallel_smoothing_vs_oracle():
    """A 2048x2048 polyline frame: enough points that the segments are
    smoothed on the host pool; the float32 outputs equal the oracle's bits."""
    import begputhinning_amd as B
    from begputhinning_amd import synth
    img = synth.polylines(2048, 2048, synth.item_seed(1, 2048))
    print("[configs] CenterlineExtraction 2048^2", flush=True)
    ref = O.centerline2d(img, 10, 3)
    assert ref["idx"][-1] >= 1 << 14  # above the parallel threshold
    x, y, idx = B.CenterlineExtraction(img, 10, 3)
    assert np.array_equal(idx, ref["idx"])
    for a, r in ((x, ref["x"]), (y, ref["y"])):
        assert a.dtype == np.float32 and np.array_equal(a.view(np.uint32), r.view(np.uint32))

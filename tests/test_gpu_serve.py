"""Serving on the MI355X: requests batched as they arrive give, for every
request, the same skeleton and iteration count as the request served alone
and as the oracle; admission by device memory rejects what can never fit and
queues what does not fit yet; concurrent submitters are safe."""
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


def requests(seed, n2=14, n3=4):
    from begputhinning_amd import synth
    rng = np.random.default_rng(seed)
    items = []
    for i in range(n2):
        w, h = (int(v) for v in rng.integers(40, 700, 2))
        gen = synth.polylines if i % 2 else synth.tree2d
        items.append(gen(w, h, 100 * seed + i))
    for i in range(n3):
        w, h, d = (int(v) for v in rng.integers(16, 90, 3))
        items.append(synth.tree3d(w, h, d, 300 * seed + i, target_fraction=0.03))
    return items


def test_alone_equals_mix_equals_oracle():
    from begputhinning_amd.serve import ThinningServer
    items = requests(1)
    print(f"[serve] {len(items)} requests", flush=True)
    ins = [torch.from_numpy(a).cuda() for a in items]
    torch.cuda.synchronize()
    srv = ThinningServer(budget_bytes=2 << 30)
    try:
        alone, alone_it = [], []
        for t in ins:
            o = torch.full_like(t, 5)
            it, _ = srv.wait(srv.submit(t, o))
            alone.append(o.cpu().numpy())
            alone_it.append(it)
        outs = [torch.full_like(t, 5) for t in ins]
        tickets = [srv.submit(t, o) for t, o in zip(ins, outs)]
        mix_it = [srv.wait(tk)[0] for tk in tickets]
        st = srv.stats()
    finally:
        srv.close()
    for k, (a, o) in enumerate(zip(items, outs)):
        ref, rit = O.thin2d(a) if a.ndim == 2 else O.thin3d(a)
        got = o.cpu().numpy()
        assert np.array_equal(got, alone[k]) and mix_it[k] == alone_it[k], k
        assert np.array_equal(got, ref) and mix_it[k] == rit, k
    assert st["completed"] == 2 * len(items)
    assert st["batches"] < 2 * len(items), "arrivals were never batched together"


def test_admission_by_device_memory():
    from begputhinning_amd import BGTError, synth
    from begputhinning_amd.serve import ThinningServer
    big = torch.from_numpy(synth.polylines(4000, 3000, 7)).cuda()  # ~3 MB of engine state
    small = [torch.from_numpy(synth.polylines(300, 200, 8 + i)).cuda() for i in range(12)]
    torch.cuda.synchronize()
    srv = ThinningServer(budget_bytes=1 << 20, workers2d=2, workers3d=0)
    try:
        with pytest.raises(BGTError) as e:
            srv.submit(big, torch.empty_like(big))
        assert e.value.code == 6  # BGT_ERR_ADMISSION
        outs = [torch.empty_like(t) for t in small]
        tickets = [srv.submit(t, o) for t, o in zip(small, outs)]
        for tk in tickets:
            srv.wait(tk)
        st = srv.stats()
    finally:
        srv.close()
    assert st["peak_bytes"] <= 1 << 20
    for t, o in zip(small, outs):
        ref, _ = O.thin2d(t.cpu().numpy())
        assert np.array_equal(o.cpu().numpy(), ref)


def test_concurrent_submitters():
    from begputhinning_amd.serve import ThinningServer
    items = requests(2, n2=16, n3=4)
    ins = [torch.from_numpy(a).cuda() for a in items]
    outs = [torch.empty_like(t) for t in ins]
    torch.cuda.synchronize()
    srv = ThinningServer(budget_bytes=1 << 30)
    errors = []

    def client(ks):
        try:
            for k in ks:
                srv.wait(srv.submit(ins[k], outs[k]))
        except Exception as exc:  # reported below
            errors.append(exc)

    threads = [threading.Thread(target=client, args=(list(range(j, len(items), 4)),))
               for j in range(4)]
    try:
        for t in threads:
            t.start()
        for t in threads:
            t.join(120)
    finally:
        srv.close()
    assert not errors, errors
    for a, o in zip(items, outs):
        ref, _ = O.thin2d(a) if a.ndim == 2 else O.thin3d(a)
        assert np.array_equal(o.cpu().numpy(), ref)


def test_server_memory_returned_on_destroy():
    """Workers release their arenas, per-thread workspaces and pinned buffers
    when the server is destroyed: device free memory is the same after each
    create/submit/destroy cycle (single requests take the per-item path, so
    the workers' thread workspaces are exercised too)."""
    from begputhinning_amd import synth
    from begputhinning_amd.serve import ThinningServer
    items = [synth.tree3d(160, 150, 140, 41), synth.polylines(900, 700, 42),
             synth.tree3d(60, 50, 40, 43), synth.polylines(300, 200, 44)]
    ins = [torch.from_numpy(a).cuda() for a in items]
    outs = [torch.empty_like(t) for t in ins]
    torch.cuda.synchronize()

    def cycle():
        srv = ThinningServer(budget_bytes=4 << 30, workers2d=2, workers3d=2)
        try:
            for t, o in zip(ins, outs):  # one at a time: the per-item path
                srv.wait(srv.submit(t, o))
            tks = [srv.submit(t, o) for t, o in zip(ins, outs)]  # together: a batch
            for tk in tks:
                srv.wait(tk)
        finally:
            srv.close()
        torch.cuda.synchronize()
        return torch.cuda.mem_get_info()[0]

    free = [cycle() for _ in range(4)]
    print(f"[serve] free bytes after each cycle: {free}", flush=True)
    # the first cycle may leave the runtime's own pools grown; later cycles
    # must not lose memory
    assert max(free[1:]) - min(free[1:]) < 64 << 20, free
    for a, o in zip(items, outs):
        ref, _ = O.thin2d(a) if a.ndim == 2 else O.thin3d(a)
        assert np.array_equal(o.cpu().numpy(), ref)


def test_frames_join_the_running_queue():
    """Frames submitted while others are mid-run (staggered, from several
    threads; deep small frames beside large shallow ones) join the running
    2D queue at a pass boundary; every result and iteration count equals
    the oracle's and the single-image path's."""
    import time
    import begputhinning_amd as B
    from begputhinning_amd import synth
    from begputhinning_amd.serve import ThinningServer
    rng = np.random.default_rng(123)
    items = []
    for i in range(40):
        w, h = (int(v) for v in rng.integers(64, 1400, 2))
        items.append((synth.polylines if i % 3 else synth.tree2d)(w, h, 900 + i))
    items += [np.ones((150, 130), np.uint8), np.zeros((9, 700), np.uint8), np.ones((1, 1), np.uint8)]
    ins = [torch.from_numpy(a).cuda() for a in items]
    outs = [torch.full_like(t, 4) for t in ins]
    torch.cuda.synchronize()
    srv = ThinningServer(budget_bytes=4 << 30, workers2d=2, workers3d=1, max_items=16)
    res = [None] * len(items)
    errors = []

    def client(ks, seed):
        r = np.random.default_rng(seed)
        try:
            tks = []
            for k in ks:
                time.sleep(float(r.uniform(0, 0.002)))
                tks.append((k, srv.submit(ins[k], outs[k])))
            for k, tk in tks:
                res[k] = srv.wait(tk)[0]
        except Exception as exc:  # reported below
            errors.append(exc)

    threads = [threading.Thread(target=client, args=(list(range(j, len(items), 3)), j)) for j in range(3)]
    try:
        for t in threads:
            t.start()
        for t in threads:
            t.join(120)
    finally:
        srv.close()
    assert not errors, errors
    for k, a in enumerate(items):
        ref, rit = O.thin2d(a)
        assert res[k] == rit, (k, a.shape, res[k], rit)
        assert np.array_equal(outs[k].cpu().numpy(), ref), (k, a.shape)
        single, sit = B.thin2d_device(ins[k])
        assert sit == rit and torch.equal(single, outs[k]), k

"""2D batches whose first band pass reads the caller's bytes itself
(bgt_set_fused_io2d(1)) against a separate pack launch (0, the default):
every item equals the oracle in both modes, iteration counts included, on
pitched views at odd offsets with the bytes around each output row left
untouched, on images whose bands are mostly empty (the first pass skips
empty windows and queues every neighbour of a changed band), on images that
converge many passes apart, and with the pass chunks replayed from a
captured graph (the first chunk then runs as direct launches)."""
import numpy as np
import pytest

import _oracle as O

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


@pytest.fixture(scope="module", autouse=True)
def _gpu():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")


def _items(seed):
    from begputhinning_amd import synth
    rng = np.random.default_rng(seed)
    items = []
    for i in range(14):
        w, h = (int(v) for v in rng.integers(30, 1300, 2))
        items.append(synth.polylines(w, h, 300 * seed + i, n_lines=int(rng.integers(1, 30))))
    # bands far apart: a dot at the top, a blob at the bottom, empty bands between
    a = np.zeros((1500, 300), np.uint8)
    a[3, 7] = 1
    a[1300:1420, 40:220] = 1
    items.append(a)
    # a thick disc (many iterations) beside thin lines (few): bands settle in
    # different passes
    yy, xx = np.mgrid[0:900, 0:700]
    items.append(np.maximum(((xx - 350) ** 2 + (yy - 600) ** 2 < 200 ** 2).astype(np.uint8),
                            synth.polylines(700, 900, 77, n_lines=8)))
    items.append(np.zeros((700, 64), np.uint8))
    items.append(np.ones((2, 2), np.uint8))
    # non-binary foreground values (any nonzero byte is foreground)
    items.append((rng.integers(0, 4, (333, 517)) * 60).astype(np.uint8))
    return items


def _run(items, mode, pitched, with_volume):
    import begputhinning_amd as B
    from begputhinning_amd.batch import ThinningBatch
    B.set_fused_io2d(mode)
    try:
        b = ThinningBatch()
        ins, outs, bigs = [], [], []
        for k, a in enumerate(items):
            if pitched:
                h, w = a.shape
                ox, oy, pad = 1 + k % 13, 2, 5 + (k * 7) % 19
                src = torch.full((h + 4, w + ox + pad), 7, dtype=torch.uint8, device="cuda")
                src[oy:oy + h, ox:ox + w] = torch.from_numpy(a).cuda()
                big = torch.full((h + 4, w + ox + pad), 9, dtype=torch.uint8, device="cuda")
                ins.append(src[oy:oy + h, ox:ox + w])
                outs.append(big[oy:oy + h, ox:ox + w])
                bigs.append((big, oy, ox))
            else:
                ins.append(torch.from_numpy(a).cuda())
                outs.append(torch.full_like(ins[-1], 9))
        for t, o in zip(ins, outs):
            b.add2d(t, o)
        vol = out3 = None
        if with_volume:
            from begputhinning_amd import synth
            vol = synth.tree3d(60, 50, 40, 5)
            t3 = torch.from_numpy(vol).cuda()
            out3 = torch.full_like(t3, 9)
            b.add3d(t3, out3)
        for rep in range(2):
            for o in outs:
                o.fill_(9)
            b.run(use_graph=rep == 1)  # the second run replays captured pass chunks
            torch.cuda.synchronize()
            its = b.iterations()
            for k, a in enumerate(items):
                ref, rit = O.thin2d(a)
                assert np.array_equal(outs[k].cpu().numpy(), ref), (mode, pitched, rep, k, a.shape)
                assert its[k] == rit, (mode, pitched, rep, k, its[k], rit)
                if pitched:
                    big, oy, ox = bigs[k]
                    g = big.cpu().numpy()
                    h, w = a.shape
                    mask = np.ones(g.shape, bool)
                    mask[oy:oy + h, ox:ox + w] = False
                    assert (g[mask] == 9).all(), (mode, rep, k, "bytes outside the view written")
            if with_volume:
                ref3, rit3 = O.thin3d(vol)
                assert np.array_equal(out3.cpu().numpy(), ref3)
                assert its[len(items)] == rit3
        b.close()
    finally:
        B.set_fused_io2d()


@pytest.mark.parametrize("mode", [0, 1])
@pytest.mark.parametrize("pitched", [False, True])
def test_fused_io_batch(mode, pitched):
    print(f"[fused_io] mode={mode} pitched={pitched}", flush=True)
    _run(_items(1), mode, pitched, with_volume=not pitched)


@pytest.mark.parametrize("cfg", [(4, 16, 3), (4, 8, 2), (1, 16, 1), (8, 16, 6)])
def test_fused_io_launch_shapes(cfg):
    """Other window and K choices (the first pass split into parts when few
    bands are queued)."""
    import begputhinning_amd as B
    print(f"[fused_io] shape={cfg}", flush=True)
    B.set_config2d(cfg[0], cfg[1], cfg[2], -1)
    try:
        _run(_items(2), 1, False, with_volume=False)
    finally:
        B.set_config2d()

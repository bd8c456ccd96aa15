"""2D bands with 1, 2 or 4 64-bit words per lane on the MI355X: every choice
gives the oracle's skeletons and iteration counts (per-image path, batch,
serving queue), and images wider than 4096 pixels (up to 16384, which need
two or four words a lane) thin like any other."""
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
    B.set_config2d(4, 16, 3)
    B.set_lane_words2d(0)
    yield
    B.set_config2d(4, 16, 3)
    B.set_lane_words2d(0)


def check(img, label):
    import begputhinning_amd as B
    print(f"[lanes] {label} shape={img.shape}", flush=True)
    ref, rit = O.thin2d(img)
    t = torch.from_numpy(np.ascontiguousarray(img, np.uint8)).cuda()
    got, it = B.thin2d_device(t)
    torch.cuda.synchronize()
    got = got.cpu().numpy()
    assert np.array_equal(got, ref), f"{label}: {np.argwhere(got != ref)[:5]}"
    assert it == rit, f"{label}: iterations {it} != {rit}"


def random_image(w, h, p=0.6):
    rng = np.random.default_rng(w * 7919 + h)
    return (rng.random((h, w)) < p).astype(np.uint8)


@pytest.mark.parametrize("w,h", [(4097, 30), (5000, 21), (8192, 17), (8193, 12), (12000, 11),
                                 (16320, 9), (16384, 13)])
def test_wide_random(w, h):
    check(random_image(w, h), f"random {w}x{h}")


def test_wide_structures():
    from begputhinning_amd import synth
    check(synth.polylines(9000, 700, 31, n_lines=60), "polylines 9000x700")
    check(synth.tree2d(16384, 300, 32), "tree 16384x300")
    yy, xx = np.mgrid[0:400, 0:6000]
    check(((xx - 3000) ** 2 / 9 + (yy - 200) ** 2 < 180 ** 2).astype(np.uint8), "ellipse 6000x400")


def test_too_wide_rejected():
    import begputhinning_amd as B
    t = torch.zeros((4, 16385), dtype=torch.uint8, device="cuda")
    with pytest.raises(B.BGTError):
        B.thin2d_device(t)


@pytest.mark.parametrize("words", [1, 2, 4])
@pytest.mark.parametrize("cfg", [(4, 16, 3), (4, 8, 2), (1, 16, 1), (8, 16, 6), (2, 16, 2)])
def test_forced_lane_words(words, cfg):
    """Rows of one word and of many, single- and multi-band, with each
    forced count of words per lane (rows narrower than a lane leave words
    that read as zero and are never stored)."""
    import begputhinning_amd as B
    from begputhinning_amd import synth
    B.set_config2d(*cfg)
    B.set_lane_words2d(words)
    try:
        for w, h in [(1, 5), (40, 300), (64, 64), (65, 90), (130, 33), (200, 700), (257, 40),
                     (700, 900), (2049, 60), (4096, 25)]:
            check(random_image(w, h), f"V={words} cfg={cfg} random {w}x{h}")
        check(synth.polylines(700, 900, 12345), f"V={words} cfg={cfg} polylines")
        check(synth.tree2d(1024, 1024, 8), f"V={words} cfg={cfg} tree")
    finally:
        B.set_config2d(4, 16, 3)
        B.set_lane_words2d(0)


@pytest.mark.parametrize("words", [0, 2, 4])
def test_batch_lane_words(words):
    """A batch plans every image with the words per lane of its widest one: a
    batch holding an image wider than 4096 (or a forced count) still gives
    each item the oracle's skeleton and iteration count."""
    import begputhinning_amd as B
    from begputhinning_amd import synth
    from begputhinning_amd.batch import ThinningBatch
    rng = np.random.default_rng(77 + words)
    items = []
    for i in range(10):
        w, h = (int(v) for v in rng.integers(20, 1500, 2))
        items.append(synth.polylines(w, h, 500 + i, n_lines=int(rng.integers(1, 30))))
    items += [random_image(6000, 40), synth.tree2d(10000, 200, 3), np.ones((3, 9000), np.uint8),
              np.zeros((5, 7), np.uint8), synth.tree3d(60, 50, 40, 9)]
    ins = [torch.from_numpy(a).cuda() for a in items]
    outs = [torch.full_like(t, 9) for t in ins]
    B.set_lane_words2d(words)
    try:
        b = ThinningBatch()
        for t, o in zip(ins, outs):
            (b.add2d if t.dim() == 2 else b.add3d)(t, o)
        for rep in range(2):
            b.run(use_graph=rep == 1)
            torch.cuda.synchronize()
            its = b.iterations()
            for k, (a, o) in enumerate(zip(items, outs)):
                ref, rit = O.thin2d(a) if a.ndim == 2 else O.thin3d(a)
                assert np.array_equal(o.cpu().numpy(), ref), (words, rep, k, a.shape)
                assert its[k] == rit, (words, rep, k, its[k], rit)
        b.close()
    finally:
        B.set_lane_words2d(0)


def test_serving_queue_two_lane_words():
    """The serving queue's slots take the words per lane in force when the
    server starts; with two, every frame still equals the oracle."""
    import begputhinning_amd as B
    from begputhinning_amd import synth
    from begputhinning_amd.serve import ThinningServer
    rng = np.random.default_rng(5)
    items = [synth.polylines(int(w), int(h), 40 + i)
             for i, (w, h) in enumerate(rng.integers(64, 1800, (12, 2)))]
    ins = [torch.from_numpy(a).cuda() for a in items]
    outs = [torch.full_like(t, 4) for t in ins]
    torch.cuda.synchronize()
    B.set_lane_words2d(2)
    try:
        srv = ThinningServer(budget_bytes=2 << 30, workers2d=2, workers3d=0, max_items=8)
        try:
            tks = [srv.submit(i, o) for i, o in zip(ins, outs)]
            its = [srv.wait(tk)[0] for tk in tks]
        finally:
            srv.close()
    finally:
        B.set_lane_words2d(0)
    for k, a in enumerate(items):
        ref, rit = O.thin2d(a)
        assert its[k] == rit, (k, a.shape)
        assert np.array_equal(outs[k].cpu().numpy(), ref), (k, a.shape)

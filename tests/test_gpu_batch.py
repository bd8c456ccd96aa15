"""Batched engine on the MI355X: every item of a mixed batch equals the oracle
(bit-exact, with its own iteration count) and equals the single-item path;
graph replay and direct launches agree; items of very different depth converge
independently."""
import numpy as np
import pytest

import _oracle as O

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


@pytest.fixture(scope="module", autouse=True)
def _gpu():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")


def make_items(seed, n2=24, n3=5):
    from begputhinning_amd import synth
    rng = np.random.default_rng(seed)
    items = []
    for i in range(n2):
        w, h = (int(v) for v in rng.integers(20, 900, 2))
        items.append(synth.polylines(w, h, 1000 * seed + i, n_lines=int(rng.integers(1, 40))))
    items.append(np.zeros((5, 7), np.uint8))
    items.append(np.ones((3, 3000), np.uint8))
    for i in range(n3):
        w, h, d = (int(v) for v in rng.integers(8, 140, 3))
        items.append(synth.tree3d(w, h, d, 7000 * seed + i, target_fraction=0.02))
    items.append(np.zeros((4, 4, 4), np.uint8))
    return items


@pytest.mark.parametrize("tail_words", [0, 2000, 1 << 30])
def test_3d_tail_in_batch(tail_words):
    """Volumes of very different depth in one batch, finishing in the
    one-launch tail from different iterations on: each equals the oracle."""
    import begputhinning_amd as B
    from begputhinning_amd.batch import ThinningBatch
    from begputhinning_amd import synth
    vols = [synth.tree3d(70, 90, 80, 99), synth.tree3d(40, 30, 120, 5, target_fraction=0.1),
            np.ones((9, 10, 11), np.uint8), np.zeros((4, 5, 6), np.uint8),
            synth.tree3d(130, 60, 50, 7, target_fraction=0.02)]
    ins = [torch.from_numpy(a).cuda() for a in vols]
    outs = [torch.full_like(t, 9) for t in ins]
    B.set_config3d(2, 16, tail_words)
    try:
        b = ThinningBatch()
        for t, o in zip(ins, outs):
            b.add3d(t, o)
        for rep in range(2):
            b.run(use_graph=False)
            torch.cuda.synchronize()
            its = b.iterations()
            for k, a in enumerate(vols):
                ref, rit = O.thin3d(a)
                assert np.array_equal(outs[k].cpu().numpy(), ref), (tail_words, rep, k)
                assert its[k] == rit, (tail_words, rep, k, its[k], rit)
        b.close()
    finally:
        B.set_config3d()


@pytest.mark.parametrize("graph", [False, True])
def test_mixed_batch_matches_oracle(graph):
    from begputhinning_amd.batch import ThinningBatch
    items = make_items(1)
    print(f"[batch] {len(items)} items graph={graph}", flush=True)
    ins = [torch.from_numpy(a).cuda() for a in items]
    outs = [torch.full_like(t, 9) for t in ins]
    b = ThinningBatch()
    for t, o in zip(ins, outs):
        (b.add2d if t.dim() == 2 else b.add3d)(t, o)
    for rep in range(2):  # a batch can be run again
        b.run(use_graph=graph)
        torch.cuda.synchronize()
        its = b.iterations()
        for k, (a, o) in enumerate(zip(items, outs)):
            ref, rit = O.thin2d(a) if a.ndim == 2 else O.thin3d(a)
            assert np.array_equal(o.cpu().numpy(), ref), (rep, k, a.shape)
            assert its[k] == rit, (rep, k, its[k], rit)
    b.close()


def test_batch_timing_and_single_item_equality():
    import begputhinning_amd as B
    from begputhinning_amd.batch import ThinningBatch
    items = make_items(2, n2=10, n3=3)
    ins = [torch.from_numpy(a).cuda() for a in items]
    outs = [torch.empty_like(t) for t in ins]
    b = ThinningBatch()
    for t, o in zip(ins, outs):
        (b.add2d if t.dim() == 2 else b.add3d)(t, o)
    b.set_timing(True)
    b.run()
    ms2, n2, ms3, n3 = b.kernel_time()
    assert n2 > 0 and ms2 > 0 and n3 > 0 and ms3 > 0
    for t, o in zip(ins, outs):
        single, _ = (B.thin2d_device(t) if t.dim() == 2 else B.thin3d_device(t))
        assert torch.equal(single, o)
    b.close()


@pytest.mark.parametrize("kind", ["empty", "tiny", "one_band", "full"])
def test_degenerate_batches(kind):
    """Batches whose first pass finds no band (all empty), single pixels, one
    single-band image, and solid images: every item equals the oracle,
    iteration counts included, over two runs."""
    from begputhinning_amd.batch import ThinningBatch
    if kind == "empty":
        items = [np.zeros((5, 7), np.uint8), np.zeros((300, 2100), np.uint8),
                 np.zeros((4, 4, 4), np.uint8)]
    elif kind == "tiny":
        items = [np.ones((1, 1), np.uint8), np.zeros((1, 1), np.uint8), np.ones((1, 1, 1), np.uint8)]
    elif kind == "one_band":
        rng = np.random.default_rng(11)
        items = [(rng.random((40, 300)) < 0.7).astype(np.uint8)]
    else:
        items = [np.ones((500, 130), np.uint8), np.ones((64, 2000), np.uint8),
                 np.ones((20, 30, 40), np.uint8)]
    ins = [torch.from_numpy(a).cuda() for a in items]
    outs = [torch.full_like(t, 9) for t in ins]
    b = ThinningBatch()
    for t, o in zip(ins, outs):
        (b.add2d if t.dim() == 2 else b.add3d)(t, o)
    for rep in range(2):
        b.run(use_graph=rep == 1)
        torch.cuda.synchronize()
        its = b.iterations()
        for k, (a, o) in enumerate(zip(items, outs)):
            ref, rit = O.thin2d(a) if a.ndim == 2 else O.thin3d(a)
            assert np.array_equal(o.cpu().numpy(), ref), (kind, rep, k)
            assert its[k] == rit, (kind, rep, k, its[k], rit)
    b.close()


def split_items():
    """Images whose tails leave a few bands queued (a thick disc, thick lines),
    over every segment count per wave from 1 (width > 2048) to 64 (width <= 64),
    single-band and multi-band."""
    from begputhinning_amd import synth
    items = []
    for k, (w, h) in enumerate([(60, 700), (200, 90), (700, 650), (1100, 420), (1400, 500),
                                (2500, 300), (1500, 40)]):
        yy, xx = np.mgrid[0:h, 0:w]
        r = min(w, h) * 0.4
        disc = ((xx - w * 0.55) ** 2 + (yy - h * 0.45) ** 2 < r * r).astype(np.uint8)
        lines = synth.polylines(w, h, 50 + k, n_lines=6)
        items.append(np.maximum(disc, lines))
    return items


@pytest.mark.parametrize("cfg", [(4, 16, 3), (4, 8, 2), (8, 16, 6), (1, 16, 1), (2, 8, 1),
                                 (4, 16, 1), (8, 8, 4), (3, 16, 2)])
def test_split_band_tail(cfg):
    """Once few bands are queued, a band is split over several workgroups with
    short windows (or kept whole where the window is too short for parts):
    every launch shape gives the oracle's skeletons and iteration counts."""
    import begputhinning_amd as B
    from begputhinning_amd.batch import ThinningBatch
    items = split_items()
    print(f"[split] cfg={cfg}", flush=True)
    ins = [torch.from_numpy(a).cuda() for a in items]
    outs = [torch.full_like(t, 9) for t in ins]
    B.set_config2d(*cfg)
    try:
        b = ThinningBatch()
        for t, o in zip(ins, outs):
            b.add2d(t, o)
        for rep in range(2):
            b.run(use_graph=rep == 1)
            torch.cuda.synchronize()
            its = b.iterations()
            for k, (a, o) in enumerate(zip(items, outs)):
                ref, rit = O.thin2d(a)
                assert np.array_equal(o.cpu().numpy(), ref), (cfg, rep, k, a.shape)
                assert its[k] == rit, (cfg, rep, k, its[k], rit)
        b.close()
    finally:
        B.set_config2d()


@pytest.mark.parametrize("blocks_per_cu,volumes", [(1, True), (3, True), (0, True), (-1, True),
                                                   (-1, False), (4, False)])
def test_pass_grid_sizes(blocks_per_cu, volumes):
    """The batch pass kernel pulls bands from a queue: any number of
    workgroups per CU gives the same skeletons and iteration counts, also
    the automatic count (-1), which differs between batches with and
    without volumes."""
    import begputhinning_amd as B
    from begputhinning_amd.batch import ThinningBatch
    items = make_items(4, n2=16, n3=0)
    if not volumes:
        items = [a for a in items if a.ndim == 2]
    ins = [torch.from_numpy(a).cuda() for a in items]
    outs = [torch.full_like(t, 9) for t in ins]
    B.set_config2d(4, 16, 3, blocks_per_cu)
    try:
        b = ThinningBatch()
        for t, o in zip(ins, outs):
            (b.add2d if t.dim() == 2 else b.add3d)(t, o)
        b.run(use_graph=False)
        torch.cuda.synchronize()
        its = b.iterations()
        b.close()
    finally:
        B.set_config2d()
    for k, (a, o) in enumerate(zip(items, outs)):
        ref, rit = O.thin2d(a) if a.ndim == 2 else O.thin3d(a)
        assert np.array_equal(o.cpu().numpy(), ref), (blocks_per_cu, k)
        assert its[k] == rit, (blocks_per_cu, k)


def test_autotune_keeps_results():
    """The autotuner tries launch shapes on the batch itself; every shape gives
    the same skeletons, and the batch keeps the fastest one."""
    from begputhinning_amd.batch import ThinningBatch
    items = make_items(3, n2=12, n3=3)
    ins = [torch.from_numpy(a).cuda() for a in items]
    outs = [torch.full_like(t, 9) for t in ins]
    b = ThinningBatch()
    for t, o in zip(ins, outs):
        (b.add2d if t.dim() == 2 else b.add3d)(t, o)
    best = b.autotune(reps=1)
    print(f"[batch] autotune -> {best}", flush=True)
    assert 1 <= best["waves"] <= 8 and best["rows"] in (8, 16)
    assert 1 <= best["iters_per_pass"] <= 16 and best["ms"] > 0
    assert 0 <= best["blocks_per_cu_2d"] <= 32 and 0 <= best["blocks_per_cu_3d"] <= 32
    b.run(use_graph=False)
    torch.cuda.synchronize()
    its = b.iterations()
    for k, (a, o) in enumerate(zip(items, outs)):
        ref, rit = O.thin2d(a) if a.ndim == 2 else O.thin3d(a)
        assert np.array_equal(o.cpu().numpy(), ref), k
        assert its[k] == rit, k
    b.close()

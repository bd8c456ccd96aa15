"""The persistent 2D chain (bgt_set_persistent2d(1)): every pass of a batch's
band queue in one launch, passes separated by a bounded device-wide barrier.
Each item equals the oracle, iteration counts included, in mixed batches
(the 3D chain running beside it), in batches of images alone (the grid then
fills every CU), with two words per lane, under other launch shapes, and over
repeated runs of one batch."""
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
    B.set_persistent2d(True)
    yield
    B.set_persistent2d(False)


def _items(seed, n2=20, n3=3, wide=False):
    from begputhinning_amd import synth
    rng = np.random.default_rng(seed)
    items = []
    for i in range(n2):
        w, h = (int(v) for v in rng.integers(20, 1200, 2))
        items.append(synth.polylines(w, h, 900 * seed + i, n_lines=int(rng.integers(1, 30))))
    yy, xx = np.mgrid[0:800, 0:600]
    items.append(((xx - 300) ** 2 + (yy - 400) ** 2 < 260 ** 2).astype(np.uint8))  # many passes
    items.append(np.zeros((40, 50), np.uint8))
    if wide:
        items.append(synth.tree2d(6000, 120, 4))
    for i in range(n3):
        w, h, d = (int(v) for v in rng.integers(10, 120, 3))
        items.append(synth.tree3d(w, h, d, 4000 * seed + i, target_fraction=0.03))
    return items


def _check(items, runs=2):
    from begputhinning_amd.batch import ThinningBatch
    ins = [torch.from_numpy(a).cuda() for a in items]
    outs = [torch.full_like(t, 9) for t in ins]
    b = ThinningBatch()
    for t, o in zip(ins, outs):
        (b.add2d if t.dim() == 2 else b.add3d)(t, o)
    for rep in range(runs):
        for o in outs:
            o.fill_(9)
        b.run(use_graph=False)
        torch.cuda.synchronize()
        its = b.iterations()
        for k, a in enumerate(items):
            ref, rit = O.thin2d(a) if a.ndim == 2 else O.thin3d(a)
            assert np.array_equal(outs[k].cpu().numpy(), ref), (rep, k, a.shape)
            assert its[k] == rit, (rep, k, its[k], rit)
    b.close()


@pytest.mark.parametrize("kind", ["mixed", "images", "wide"])
def test_persistent_chain(kind):
    print(f"[persistent] {kind}", flush=True)
    if kind == "mixed":
        _check(_items(1))
    elif kind == "images":
        _check(_items(2, n3=0))
    else:
        _check(_items(3, n2=6, n3=1, wide=True))


@pytest.mark.parametrize("cfg", [(4, 8, 2), (8, 16, 6), (2, 16, 1)])
def test_persistent_chain_shapes(cfg):
    import begputhinning_amd as B
    print(f"[persistent] shape {cfg}", flush=True)
    B.set_config2d(cfg[0], cfg[1], cfg[2], -1)
    try:
        _check(_items(4, n2=10, n3=1), runs=1)
    finally:
        B.set_config2d()

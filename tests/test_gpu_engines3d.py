"""The three 3D engines on the MI355X give the oracle's skeleton and
iteration count, per volume and in mixed batches: the word-list engine with
its one-launch tail (0), the word lists handing over to the tile engine once
few words are active (1, the handoff exercised from the first poll on, late
and never), and the tile engine throughout (2)."""
import numpy as np
import pytest

import _oracle as O

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


@pytest.fixture(scope="module", autouse=True)
def _gpu():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")


CASES = [(0, 2048), (1, 0), (1, 64), (1, 16384), (1, 1 << 30), (2, 2048)]


@pytest.fixture
def engine(request):
    import begputhinning_amd as B
    e, tail = request.param
    B.set_engine3d(e)
    B.set_config3d(2, 16, tail)
    yield e, tail
    B.set_engine3d(1)
    B.set_config3d(2, 16, 16384)


def volumes():
    from begputhinning_amd import synth
    rng = np.random.default_rng(77)
    vols = [synth.tree3d(130, 120, 110, 31), synth.tree3d(70, 40, 200, 32, target_fraction=0.1),
            synth.tree3d(33, 17, 9, 33, target_fraction=0.3),
            (rng.random((20, 30, 70)) < 0.45).astype(np.uint8),
            np.ones((18, 19, 65), np.uint8), np.zeros((5, 6, 7), np.uint8),
            np.ones((1, 1, 1), np.uint8)]
    v = np.zeros((40, 50, 96), np.uint8)
    v[-3:, -3:, -40:] = 1  # foreground up to the far corner and the last word of the line
    v[:3, :4, :5] = 1
    vols.append(v)
    return vols


@pytest.mark.parametrize("engine", CASES, indirect=True, ids=[f"e{e}-t{t}" for e, t in CASES])
def test_per_volume(engine):
    import begputhinning_amd as B
    for k, a in enumerate(volumes()):
        print(f"[engines3d] {engine} volume {k} {a.shape}", flush=True)
        out, it = B.thin3d_device(torch.from_numpy(a).cuda())
        torch.cuda.synchronize()
        ref, rit = O.thin3d(a)
        assert it == rit, (engine, k)
        assert np.array_equal(out.cpu().numpy(), ref), (engine, k)


@pytest.mark.parametrize("engine", CASES, indirect=True, ids=[f"e{e}-t{t}" for e, t in CASES])
def test_batch(engine):
    from begputhinning_amd import synth
    from begputhinning_amd.batch import ThinningBatch
    items = volumes() + [synth.polylines(300, 200, 41), synth.tree3d(256, 256, 256, synth.item_seed(3, 0))]
    print(f"[engines3d] batch {engine}: {len(items)} items", flush=True)
    ins = [torch.from_numpy(a).cuda() for a in items]
    outs = [torch.full_like(t, 9) for t in ins]
    b = ThinningBatch()
    for t, o in zip(ins, outs):
        (b.add2d if t.dim() == 2 else b.add3d)(t, o)
    for rep in range(2):
        for o in outs:
            o.fill_(9)
        b.run(use_graph=bool(rep))
        torch.cuda.synchronize()
        its = b.iterations()
        for k, a in enumerate(items):
            ref, rit = O.thin2d(a) if a.ndim == 2 else O.thin3d(a)
            assert its[k] == rit, (engine, rep, k)
            assert np.array_equal(outs[k].cpu().numpy(), ref), (engine, rep, k)
    b.close()

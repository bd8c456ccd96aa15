"""Byte <-> bit conversion at every alignment on the MI355X: caller buffers
whose rows start at byte offsets 0..7 with odd pitches, widths around the
64-pixel run boundary; results equal the oracle and no byte outside the
caller's rows is written (sentinels survive)."""
import numpy as np
import pytest

import _oracle as O

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


@pytest.fixture(scope="module", autouse=True)
def _gpu():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")


def view2d(flat, off, h, w, pitch):
    return flat.as_strided((h, w), (pitch, 1), off)


WIDTHS = [1, 3, 4, 5, 31, 33, 63, 64, 65, 127, 129, 200]


@pytest.mark.parametrize("w", WIDTHS)
@pytest.mark.parametrize("off", [0, 1, 2, 3, 5])
def test_2d_offsets(w, off):
    import begputhinning_amd as B
    h = 37
    pitch = w + 3 + (off & 1)
    rng = np.random.default_rng(w * 13 + off)
    img = (rng.random((h, w)) < 0.6).astype(np.uint8)
    print(f"[pack2d] w={w} off={off} pitch={pitch}", flush=True)
    src = torch.full((off + h * pitch + 8,), 5, dtype=torch.uint8, device="cuda")
    dst = torch.full((off + h * pitch + 8,), 7, dtype=torch.uint8, device="cuda")
    vin = view2d(src, off, h, w, pitch)
    vin.copy_(torch.from_numpy(img).cuda())
    vout = view2d(dst, off, h, w, pitch)
    B.thin2d_device(vin, out=vout)
    torch.cuda.synchronize()
    ref, _ = O.thin2d(img)
    assert np.array_equal(vout.cpu().numpy(), ref)
    mask = np.ones(dst.numel(), bool)
    for y in range(h):
        mask[off + y * pitch: off + y * pitch + w] = False
    assert (dst.cpu().numpy()[mask] == 7).all()


@pytest.mark.parametrize("w", [1, 5, 31, 33, 64, 65, 100])
@pytest.mark.parametrize("off", [0, 1, 3, 6])
def test_3d_offsets(w, off):
    import begputhinning_amd as B
    h, d = 9, 7
    n = w * h * d
    rng = np.random.default_rng(w * 7 + off)
    vol = (rng.random((d, h, w)) < 0.6).astype(np.uint8)
    print(f"[pack3d] w={w} off={off}", flush=True)
    src = torch.full((off + n + 8,), 5, dtype=torch.uint8, device="cuda")
    dst = torch.full((off + n + 8,), 7, dtype=torch.uint8, device="cuda")
    vin = src[off:off + n].view(d, h, w)
    vin.copy_(torch.from_numpy(vol).cuda())
    vout = dst[off:off + n].view(d, h, w)
    B.thin3d_device(vin, out=vout)
    torch.cuda.synchronize()
    ref, _ = O.thin3d(vol)
    assert np.array_equal(vout.cpu().numpy(), ref)
    d_all = dst.cpu().numpy()
    assert (d_all[:off] == 7).all() and (d_all[off + n:] == 7).all()


def test_batch_offsets():
    """Batched pack/unpack with items at odd offsets and pitches."""
    from begputhinning_amd.batch import ThinningBatch
    rng = np.random.default_rng(3)
    specs2 = [(int(rng.integers(1, 300)), int(rng.integers(1, 60)), k % 8) for k in range(20)]
    specs3 = [(int(rng.integers(1, 90)), int(rng.integers(1, 20)), int(rng.integers(1, 12)), k % 4)
              for k in range(6)]
    b = ThinningBatch()
    keep = []
    for (w, h, off) in specs2:
        pitch = w + (off % 3)
        img = (rng.random((h, w)) < 0.55).astype(np.uint8)
        src = torch.full((off + h * pitch + 8,), 5, dtype=torch.uint8, device="cuda")
        dst = torch.full((off + h * pitch + 8,), 7, dtype=torch.uint8, device="cuda")
        vin = view2d(src, off, h, w, pitch)
        vin.copy_(torch.from_numpy(img).cuda())
        vout = view2d(dst, off, h, w, pitch)
        b.add2d(vin, vout)
        mask = np.ones(dst.numel(), bool)
        for y in range(h):
            mask[off + y * pitch: off + y * pitch + w] = False
        keep.append((img, vout, dst, mask))
    for (w, h, d, off) in specs3:
        n = w * h * d
        vol = (rng.random((d, h, w)) < 0.55).astype(np.uint8)
        src = torch.full((off + n + 8,), 5, dtype=torch.uint8, device="cuda")
        dst = torch.full((off + n + 8,), 7, dtype=torch.uint8, device="cuda")
        vin = src[off:off + n].view(d, h, w)
        vin.copy_(torch.from_numpy(vol).cuda())
        vout = dst[off:off + n].view(d, h, w)
        b.add3d(vin, vout)
        mask = np.ones(dst.numel(), bool)
        mask[off:off + n] = False
        keep.append((vol, vout, dst, mask))
    print(f"[pack batch] {len(keep)} items", flush=True)
    b.run(use_graph=False)
    torch.cuda.synchronize()
    for k, (a, vout, dst, mask) in enumerate(keep):
        ref, _ = O.thin2d(a) if a.ndim == 2 else O.thin3d(a)
        assert np.array_equal(vout.cpu().numpy(), ref), (k, a.shape)
        assert (dst.cpu().numpy()[mask] == 7).all(), (k, a.shape)
    b.close()

"""2D parity on the MI355X: the HIP path (through the C-ABI) against the oracle,
bit-exact, including the oracle's iteration count."""
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
    yield
    B.set_config2d(4, 16, 3)


def thin_gpu(img):
    import begputhinning_amd as B
    t = torch.from_numpy(np.ascontiguousarray(img, np.uint8)).cuda()
    out, it = B.thin2d_device(t)
    torch.cuda.synchronize()
    return out.cpu().numpy(), it


def check(img, label):
    print(f"[2d] {label} shape={img.shape}", flush=True)
    ref, rit = O.thin2d(img)
    got, it = thin_gpu(img)
    assert np.array_equal(got, ref), f"{label}: {np.argwhere(got != ref)[:5]}"
    assert it == rit, f"{label}: iterations {it} != {rit}"


def test_workload1_512():
    from begputhinning_amd import synth
    check(synth.workload1(), "polylines 512x512")


def test_workload2_1024():
    from begputhinning_amd import synth
    check(synth.workload2(), "tree 1024x1024")


@pytest.mark.parametrize("w,h", [(1, 1), (1, 7), (7, 1), (3, 3), (63, 65), (64, 64), (65, 33),
                                 (100, 300), (129, 70), (257, 129), (513, 40), (1000, 61),
                                 (2048, 37), (4096, 20)])
def test_random_shapes(w, h):
    rng = np.random.default_rng(w * 7919 + h)
    img = (rng.random((h, w)) < 0.6).astype(np.uint8)
    check(img, f"random {w}x{h}")


def test_every_4x4_pattern():
    """All 2^16 patterns of a 4x4 block, each in a zero frame of its own (6x6
    cells tiled into one 1536x1536 frame; cells never interact), so every 3x3
    neighbourhood meets every direction's rule and the result depends on the
    subiteration order."""
    codes = np.arange(1 << 16)
    cells = np.zeros((1 << 16, 6, 6), np.uint8)
    for k in range(16):
        cells[:, 1 + k // 4, 1 + k % 4] = (codes >> k) & 1
    img = cells.reshape(256, 256, 6, 6).transpose(0, 2, 1, 3).reshape(1536, 1536)
    check(img, "every 4x4 pattern")


def test_full_and_empty():
    check(np.ones((77, 150), np.uint8), "all ones")
    check(np.zeros((77, 150), np.uint8), "all zeros")


@pytest.mark.parametrize("cfg", [(1, 8, 1), (2, 8, 3), (4, 8, 4), (1, 16, 1), (3, 16, 2), (4, 16, 2),
                                 (4, 16, 4), (4, 16, 5), (8, 8, 3), (8, 16, 6), (6, 8, 2)])
def test_launch_shapes(cfg):
    import begputhinning_amd as B
    from begputhinning_amd import synth
    B.set_config2d(*cfg)
    try:
        img = synth.polylines(700, 900, 12345)
        check(img, f"config {cfg}")
        rng = np.random.default_rng(5)
        check((rng.random((300, 200)) < 0.7).astype(np.uint8), f"config {cfg} random")
    finally:
        B.set_config2d(4, 16, 3)


def test_pitched_input():
    import begputhinning_amd as B
    from begputhinning_amd import synth
    img = synth.polylines(300, 200, 777)
    big = torch.zeros((200, 512), dtype=torch.uint8, device="cuda")
    big[:, :300] = torch.from_numpy(img).cuda()
    view = big[:, :300]
    out = torch.full((200, 400), 7, dtype=torch.uint8, device="cuda")
    B.thin2d_device(view, out=out[:, :300])
    torch.cuda.synchronize()
    ref, _ = O.thin2d(img)
    got = out.cpu().numpy()
    assert np.array_equal(got[:, :300], ref)
    assert (got[:, 300:] == 7).all()


DTYPES = [np.uint8, np.int8, np.uint16, np.int16, np.uint32, np.int32, np.float32, np.float64,
          np.bool_]


@pytest.mark.parametrize("dt", DTYPES)
def test_host_api_dtypes(dt):
    """Thinning2D<T>: foreground iff (unsigned)v > 0 with saturating float
    conversion; output T(0)/T(1)."""
    import begputhinning_amd as B
    rng = np.random.default_rng(3)
    h, w = 90, 130
    base = (rng.random((h, w)) < 0.6)
    if dt == np.bool_:
        data = base.copy()
        mask = base
    elif np.issubdtype(dt, np.floating):
        vals = rng.choice([0.0, 0.5, 0.99, 1.0, 2.5, -3.0, np.nan, np.inf], size=(h, w))
        data = np.where(base, vals, 0.0).astype(dt)
        mask = np.nan_to_num(data.astype(np.float64), nan=0.0) >= 1.0
    else:
        info = np.iinfo(dt)
        vals = rng.integers(info.min, int(info.max) + 1, size=(h, w), dtype=np.int64)
        data = np.where(base, vals, 0).astype(dt)
        mask = data != 0
    ref, _ = O.thin2d(mask.astype(np.uint8))
    out = B.Thinning2D(data)
    assert out.dtype == data.dtype
    assert np.array_equal(out.astype(np.float64) != 0, ref.astype(bool))
    assert set(np.unique(out.astype(np.float64)).tolist()) <= {0.0, 1.0}

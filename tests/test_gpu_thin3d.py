"""3D parity on the MI355X: the HIP path (through the C-ABI) against the oracle,
bit-exact, including the iteration count."""
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
    B.set_config3d()
    yield
    B.set_config3d()


def check(vol, label):
    import begputhinning_amd as B
    print(f"[3d] {label} shape={vol.shape}", flush=True)
    ref, rit = O.thin3d(vol)
    t = torch.from_numpy(np.ascontiguousarray(vol, np.uint8)).cuda()
    out, it = B.thin3d_device(t)
    torch.cuda.synchronize()
    got = out.cpu().numpy()
    assert np.array_equal(got, ref), f"{label}: {np.argwhere(got != ref)[:5]}"
    assert it == rit, f"{label}: iterations {it} != {rit}"


@pytest.mark.parametrize("w,h,d", [(1, 1, 1), (5, 4, 3), (33, 17, 9), (64, 64, 64), (65, 70, 40),
                                   (128, 50, 30), (200, 40, 35), (256, 64, 20), (300, 30, 20),
                                   (384, 20, 25), (512, 12, 14)])
def test_random_shapes(w, h, d):
    rng = np.random.default_rng(w * 31 + h * 7 + d)
    vol = (rng.random((d, h, w)) < 0.5).astype(np.uint8)
    check(vol, f"random {w}x{h}x{d}")


@pytest.mark.parametrize("w,h,d", [(32, 3, 3), (64, 5, 4), (96, 6, 7), (256, 9, 5)])
def test_full_last_word_corner(w, h, d):
    """Widths that fill the last state word, with foreground up to the far
    corner (x = w-1, y = h-1, z = d-1): the corner word's +x neighbours lie in
    the zero margins after the last line."""
    rng = np.random.default_rng(w + h + d)
    vol = (rng.random((d, h, w)) < 0.8).astype(np.uint8)
    vol[-2:, -2:, -40:] = 1
    check(vol, f"corner {w}x{h}x{d}")


@pytest.mark.parametrize("side", [64, 96])
def test_tree(side):
    from begputhinning_amd import synth
    check(synth.tree3d(side, side + 8, side - 8, synth.item_seed(3, side)), f"tree {side}")


def test_workload3_256():
    from begputhinning_amd import synth
    check(synth.workload3(), "tree 256^3")


@pytest.mark.parametrize("tail_words", [0, 1, 500, 1 << 30])
def test_tail_launch(tail_words):
    """The remaining iterations in one launch (one workgroup per volume) from
    whenever the active words drop to `tail_words`: never (0), late, or from
    the first poll on; the result and iteration count stay the oracle's."""
    import begputhinning_amd as B
    from begputhinning_amd import synth
    B.set_config3d(2, 16, tail_words)
    try:
        check(synth.tree3d(70, 90, 80, 99), f"tail {tail_words}")
        check(synth.workload3(), f"tail {tail_words} 256^3")
        rng = np.random.default_rng(tail_words % 1000)
        check((rng.random((20, 33, 47)) < 0.5).astype(np.uint8), f"tail {tail_words} random")
    finally:
        B.set_config3d()


@pytest.mark.parametrize("cfg", [(1, 2), (2, 4), (8, 64), (0, 16)])
def test_launch_shapes(cfg):
    import begputhinning_amd as B
    from begputhinning_amd import synth
    B.set_config3d(*cfg)
    try:
        check(synth.tree3d(70, 90, 80, 99), f"config {cfg}")
    finally:
        B.set_config3d()


@pytest.mark.parametrize("dt", [np.uint8, np.int16, np.float32, np.bool_])
def test_host_api_dtypes(dt):
    """Thinning3D<T>: foreground iff v > 0 in T."""
    import begputhinning_amd as B
    rng = np.random.default_rng(11)
    shape = (20, 30, 40)
    base = rng.random(shape) < 0.55
    if dt == np.bool_:
        data = base
        mask = base
    elif dt == np.float32:
        vals = rng.choice([0.0, 0.25, 1.0, -2.0, np.nan], size=shape)
        data = np.where(base, vals, 0.0).astype(dt)
        mask = np.nan_to_num(data, nan=0.0) > 0
    else:
        vals = rng.integers(-5, 6, size=shape)
        data = np.where(base, vals, 0).astype(dt)
        mask = data > 0
    ref, _ = O.thin3d(mask.astype(np.uint8))
    out = B.Thinning3D(data)
    assert out.dtype == data.dtype
    assert np.array_equal(out != 0, ref.astype(bool))

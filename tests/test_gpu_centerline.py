"""Segment lists and centerlines of the full drop-in pipeline on the MI355X
(GPU thinning, host tracing and post-processing) against the oracle,
bit-exact (float32 outputs compared as bits)."""
import numpy as np
import pytest

import _oracle as O

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


@pytest.fixture(scope="module", autouse=True)
def _gpu():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")


def same_floats(a, b):
    return a.dtype == b.dtype == np.float32 and np.array_equal(a.view(np.uint32), b.view(np.uint32))


@pytest.mark.parametrize("mx,span", [(10, 3), (0, 0), (-1, 5), (1000, 2)])
def test_centerline2d_workload1(mx, span):
    import begputhinning_amd as B
    from begputhinning_amd import synth
    img = synth.workload1()
    ref = O.centerline2d(img, mx, span)
    x, y, idx = B.CenterlineExtraction(img, mx, span)
    assert np.array_equal(idx, ref["idx"])
    assert same_floats(x, ref["x"]) and same_floats(y, ref["y"])


def test_centerline2d_tree_float_input():
    import begputhinning_amd as B
    from begputhinning_amd import synth
    img = synth.tree2d(600, 500, 4242)
    ref = O.centerline2d(img, 8, 4)
    x, y, idx = B.CenterlineExtraction(img.astype(np.float32) * 3.0, 8, 4)
    assert np.array_equal(idx, ref["idx"])
    assert same_floats(x, ref["x"]) and same_floats(y, ref["y"])


def test_centerline3d_tree():
    import begputhinning_amd as B
    from begputhinning_amd import synth
    vol = synth.tree3d(90, 70, 60, 777)
    ref = O.centerline3d(vol, 5, 2)
    x, y, z, idx = B.CenterlineExtraction(vol, 5, 2)
    assert np.array_equal(idx, ref["idx"])
    assert same_floats(x, ref["x"]) and same_floats(y, ref["y"]) and same_floats(z, ref["z"])


def test_extract_segments_on_gpu_skeleton():
    import begputhinning_amd as B
    from begputhinning_amd import synth
    img = synth.workload2()
    skel = B.Thinning2D(img)
    got = B.extract_segments_flat(skel)
    ref = O.extract2d(skel)
    for k in ("x", "y", "start", "is_end"):
        assert np.array_equal(got[k], ref[k]), k


@pytest.mark.parametrize("dt", [np.float32, np.float64, np.uint32, np.int16, np.int8, np.uint16])
def test_centerline3d_pixel_types(dt):
    """3D CenterlineExtraction keeps int(v) > 0 (3D/CenterlineExtraction.cpp:44),
    evaluated on the GPU: fractions below one, negatives, NaN, infinities,
    values past int's range and unsigned values that wrap negative are
    background."""
    import begputhinning_amd as B
    from begputhinning_amd import synth
    base = synth.tree3d(60, 50, 40, 99).astype(bool)
    rng = np.random.default_rng(7)
    if np.issubdtype(dt, np.floating):
        pool = np.array([0.5, 0.99, 1.0, 1.5, 7.0, -3.0, np.nan, np.inf, -np.inf, 3e9, 2147483647.0])
        vals = rng.choice(pool, size=base.shape).astype(dt)
        mask = (vals.astype(np.float64) >= 1.0) & (vals.astype(np.float64) < 2147483648.0)
    else:
        info = np.iinfo(dt)
        vals = rng.integers(info.min, int(info.max) + 1, size=base.shape, dtype=np.int64).astype(dt)
        mask = vals.astype(np.int64).astype(np.int32) > 0 if dt == np.uint32 else vals > 0
    data = np.where(base, vals, np.zeros((), dt)).astype(dt)
    mask = mask & base
    ref = O.centerline3d(mask.astype(np.uint8), 6, 3)
    x, y, z, idx = B.CenterlineExtraction(data, 6, 3)
    assert np.array_equal(idx, ref["idx"])
    assert same_floats(x, ref["x"]) and same_floats(y, ref["y"]) and same_floats(z, ref["z"])

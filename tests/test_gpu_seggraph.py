"""Segment tracing over skeletons in device memory: the GPU classifies and
compacts the skeleton points, the host follows only them.  The segment lists
must equal the oracle's tracer (pinned against the reference's compiled
ExtractSegments.cpp) and the library's host tracer on the same zero-padded
frame, bit for bit: coordinates, segment starts and end flags."""
import numpy as np
import pytest

import _oracle as O

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


@pytest.fixture(scope="module", autouse=True)
def _gpu():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")


def padded(a, shape):
    f = np.zeros(shape, np.uint8)
    f[tuple(slice(0, n) for n in a.shape)] = a
    return f


def check(a, shape, label):
    import begputhinning_amd as B
    print(f"[seggraph] {label} {a.shape} -> {shape}", flush=True)
    frame = padded(a, shape)
    oracle = O.extract2d(frame) if frame.ndim == 2 else O.extract3d(frame)
    host = B.extract_segments_flat(frame)
    got = B.extract_segments_device(torch.from_numpy(np.ascontiguousarray(a)).cuda(), shape)
    for k in ("x", "y", "z", "start", "is_end"):
        if oracle[k] is None:
            assert got[k] is None and host[k] is None
            continue
        assert np.array_equal(oracle[k], got[k]), f"{label}: {k} differs from the oracle"
        assert np.array_equal(oracle[k], host[k]), f"{label}: host tracer {k} differs from the oracle"


@pytest.mark.parametrize("w,h,density", [(1, 1, 1.0), (3, 3, 1.0), (5, 9, 0.5), (64, 64, 0.2),
                                         (65, 33, 0.3), (130, 70, 0.08), (200, 150, 0.5),
                                         (1000, 37, 0.05)])
def test_random_masks_2d(w, h, density):
    """Dense and sparse random masks: junctions of every degree, points on
    the frame's border (skipped but counted as neighbours)."""
    rng = np.random.default_rng(w * 31 + h)
    a = (rng.random((h, w)) < density).astype(np.uint8)
    for shape in [(h, w), (h + 1, w + 2), (max(h, 3) + 64, (w + 31) // 32 * 32 + 32)]:
        check(a, shape, f"random {density}")


def test_skeletons_2d():
    from begputhinning_amd import synth
    for img, label in [(synth.workload1(), "polylines 512"), (synth.workload2(), "tree 1024")]:
        skel, _ = O.thin2d(img)
        h, w = skel.shape
        check(skel, (((h + 63) // 128) * 128 + 64, (w + 31) // 32 * 32), label)
        check(skel, skel.shape, label + " unpadded")


@pytest.mark.parametrize("w,h,d,density", [(1, 1, 1, 1.0), (3, 3, 3, 1.0), (7, 5, 4, 0.5),
                                           (33, 20, 9, 0.1), (70, 40, 12, 0.03)])
def test_random_masks_3d(w, h, d, density):
    rng = np.random.default_rng(w * 131 + h * 7 + d)
    a = (rng.random((d, h, w)) < density).astype(np.uint8)
    for shape in [(d, h, w), (d + 1, h + 2, w + 3), ((d + 31) // 32 * 32, (h + 31) // 32 * 32,
                                                     (w + 31) // 32 * 32)]:
        check(a, shape, f"random3d {density}")


def test_skeleton_3d():
    from begputhinning_amd import synth
    v = synth.tree3d(96, 80, 64, 5)
    skel, _ = O.thin3d(v)
    check(skel, (96, 96, 96), "tree3d")


def test_padded_frame_smaller_than_skeleton_is_rejected():
    import begputhinning_amd as B
    t = torch.ones((4, 4), dtype=torch.uint8, device="cuda")
    with pytest.raises(B.BGTError):
        B.extract_segments_device(t, (3, 4))


@pytest.mark.parametrize("dtype", ["uint8", "bool", "int16", "float32", "float64", "uint32"])
def test_large_host_arrays_traced_on_the_device(dtype):
    """ExtractSegmentsV on host arrays of at least half a megapixel goes
    through the device (copy in, nonzero as foreground, pre-pass, host walk):
    the lists equal the oracle's tracer on the array's nonzero mask, for every
    pixel type (negative, NaN and -0.0 values included), in 2D and 3D, on
    skeletons and on raw masks dense with junctions."""
    import begputhinning_amd as B
    from begputhinning_amd import synth
    rng = np.random.default_rng(7)
    skel2, _ = O.thin2d(synth.polylines(1024, 640, synth.item_seed(1, 5)))
    raw2 = (rng.random((600, 900)) < 0.3).astype(np.uint8)
    skel3, _ = O.thin3d(synth.tree3d(96, 88, 80, 11))
    for m in (skel2, raw2, skel3):
        assert m.size >= 1 << 19
        print(f"[seggraph] host array {dtype} {m.shape}", flush=True)
        a = m.astype(dtype)
        if dtype in ("int16", "float32", "float64"):
            a = np.where(m != 0, -3, 0).astype(dtype)  # negative values are set
        if dtype in ("float32", "float64"):
            a[m == 0] = -0.0  # negative zero is not
            nan_at = np.flatnonzero(m)[::97]
            a.flat[nan_at] = np.nan  # NaN is nonzero
        ref = O.extract2d(m) if m.ndim == 2 else O.extract3d(m)
        got = B.extract_segments_flat(a)
        for k in ("x", "y", "z", "start", "is_end"):
            if ref[k] is None:
                assert got[k] is None
                continue
            assert np.array_equal(ref[k], got[k]), f"{dtype} {m.shape}: {k} differs from the oracle"


def test_padded_frame_beyond_int16_is_rejected():
    """Device tracing refuses a padded frame with a side above 32767 (the
    reference's short coordinates) and takes one of exactly 32767."""
    import begputhinning_amd as B
    t = torch.zeros((2, 8), dtype=torch.uint8, device="cuda")
    t[1, 3] = 1
    with pytest.raises(B.BGTError, match="int16"):
        B.extract_segments_device(t, (2, 32768))
    r = B.extract_segments_device(t, (2, 32767))
    assert len(r["start"]) == 1  # the only point lies on the frame's border

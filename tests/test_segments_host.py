"""Host tracer and post-processing of the product library (no GPU needed)
against the oracle, and the oracle against the reference's own tracer compiled
on the host (oracle/_ref), on hand-made shapes and seeded skeletons."""
import json
import os

import numpy as np
import pytest

import _oracle as O

GOLD = os.path.join(os.path.dirname(__file__), "golden", "segments2d.json")
GOLD3 = os.path.join(os.path.dirname(__file__), "golden", "segments3d.json")


def shapes2d():
    """Hand-made skeletons covering the tracer's cases."""
    out = {}
    def canvas(h=12, w=14):
        return np.zeros((h, w), np.uint8)
    a = canvas(); a[2:10, 6] = 1; out["I"] = a
    a = canvas(); a[2:7, 6] = 1; a[7, 5] = a[8, 4] = a[7, 7] = a[8, 8] = 1; out["Y"] = a
    a = canvas(); a[6, 2:12] = 1; a[2:11, 7] = 1; out["X"] = a
    a = canvas(); a[3, 3:9] = 1; a[8, 3:9] = 1; a[3:9, 3] = 1; a[3:9, 8] = 1; a[8, 9:12] = 1; out["loop_on_junction"] = a
    a = canvas(); a[5, 5] = 1; out["isolated"] = a
    a = canvas(); a[5, 5] = a[5, 6] = 1; out["two_points"] = a
    a = canvas(); a[0, 3:10] = 1; a[1:6, 9] = 1; out["border_touching"] = a
    a = canvas(); a[3, 3:9] = 1; a[8, 3:9] = 1; a[3:9, 3] = 1; a[3:9, 8] = 1; out["closed_ring"] = a
    a = canvas(); a[2, 2] = a[3, 3] = a[4, 4] = a[5, 5] = a[5, 6] = a[6, 7] = a[6, 8] = 1; out["diagonal"] = a
    a = canvas(16, 16); a[8, 2:14] = 1; a[2:14, 8] = 1; a[3, 3:13] = 1; out["grid"] = a
    return out


def _cmp(a, b, keys=("x", "y", "start", "is_end")):
    for k in keys:
        assert np.array_equal(np.asarray(a[k]), np.asarray(b[k])), k


@pytest.mark.skipif(not O.ref_available("libref2d.so"), reason="reference tracer not built")
def test_oracle_vs_reference_tracer_2d():
    for name, img in shapes2d().items():
        _cmp(O.extract2d(img), O.ref_extract2d(img))
    rng = np.random.default_rng(1)
    for k in range(20):
        img = (rng.random((40, 50)) < 0.3).astype(np.uint8)
        skel, _ = O.thin2d(img)
        _cmp(O.extract2d(skel), O.ref_extract2d(skel))
        _cmp(O.extract2d(img), O.ref_extract2d(img))  # raw (non-skeleton) input too


@pytest.mark.skipif(not O.ref_available("libref3d.so"), reason="reference tracer not built")
def test_oracle_vs_reference_tracer_3d():
    rng = np.random.default_rng(2)
    for k in range(10):
        vol = (rng.random((14, 16, 18)) < 0.3).astype(np.uint8)
        skel, _ = O.thin3d(vol)
        _cmp(O.extract3d(skel), O.ref_extract3d(skel), ("x", "y", "z", "start", "is_end"))
        _cmp(O.extract3d(vol), O.ref_extract3d(vol), ("x", "y", "z", "start", "is_end"))


def test_oracle_vs_committed_golden_2d():
    with open(GOLD) as f:
        gold = json.load(f)
    for name, img in shapes2d().items():
        _cmp(O.extract2d(img), gold[name])


def test_oracle_vs_committed_golden_3d():
    with open(GOLD3) as f:
        gold = json.load(f)
    for case in gold:
        rng = np.random.default_rng(case["seed"])
        vol = (rng.random(tuple(case["shape"])) < case["p"]).astype(np.uint8)
        if case["thin"]:
            vol, _ = O.thin3d(vol)
        _cmp(O.extract3d(vol), case, ("x", "y", "z", "start", "is_end"))


def test_product_tracer_2d_matches_oracle():
    import begputhinning_amd as B
    for name, img in shapes2d().items():
        _cmp(B.extract_segments_flat(img), O.extract2d(img))
    rng = np.random.default_rng(3)
    for k in range(10):
        img = (rng.random((60, 70)) < 0.25).astype(np.uint8)
        _cmp(B.extract_segments_flat(img), O.extract2d(img))
        _cmp(B.extract_segments_flat(img.astype(np.float64) * -2.0), O.extract2d(img))


def test_product_tracer_3d_matches_oracle():
    import begputhinning_amd as B
    rng = np.random.default_rng(4)
    for k in range(6):
        vol = (rng.random((12, 15, 17)) < 0.25).astype(np.uint8)
        _cmp(B.extract_segments_flat(vol), O.extract3d(vol), ("x", "y", "z", "start", "is_end"))


def test_smoothing_closed_form():
    """or_smooth against the closed form: output k is the mean of the window
    [k - s_k, k + s_k], s_k = min(k, span, n-1-k), computed from exact integer
    sums and divided in float32 (exact while sums stay below 2^24)."""
    rng = np.random.default_rng(5)
    for n in list(range(1, 12)) + [37, 200]:
        for span in (0, 1, 2, 3, 5, 40):
            px = rng.integers(0, 2048, n).astype(np.int16)
            py = rng.integers(0, 2048, n).astype(np.int16)
            pz = rng.integers(0, 512, n).astype(np.int16)
            ox, oy, oz = O.smooth(px, py, pz, span)
            sp = min(span, (n - 1) // 2)
            for arr, out in ((px, ox), (py, oy), (pz, oz)):
                cs = np.concatenate([[0], np.cumsum(arr.astype(np.int64))])
                exp = np.zeros(n, np.float32)
                for k in range(n):
                    s = min(k, sp, n - 1 - k)
                    tot = cs[k + s + 1] - cs[k - s]
                    exp[k] = np.float32(tot) / np.float32(2 * s + 1)
                assert np.array_equal(out.view(np.uint32), exp.view(np.uint32)), (n, span)


def test_prune_rule_hand_cases():
    """Pruning keeps a segment iff size > max_segment_length or it is not an
    end segment (2D/CenterlineExtraction.cpp:51-57); a negative limit prunes
    every end segment."""
    img = np.zeros((20, 30), np.uint8)
    img[5, 3:20] = 1          # end segment of 17 points
    img[12, 3:6] = 1          # end segment of 3 points
    segs = O.extract2d(img)
    assert len(segs["is_end"]) == 2 and segs["is_end"].all()
    assert O.centerline2d(img, 3, 0)["idx"].tolist() == [0, 17]
    assert O.centerline2d(img, 2, 0)["idx"].tolist() == [0, 17, 20]
    assert O.centerline2d(img, 17, 0)["idx"].tolist() == [0]
    assert O.centerline2d(img, -1, 0)["idx"].tolist() == [0]


def test_padding_rule():
    """2D padding pw = ceil32(w), ph = max(1, (h+63)/128)*128 + 64
    (2D/CenterlineExtraction.cpp:27-28): a skeleton touching the bottom image
    row is traced when ph > h and skipped as frame border when ph == h
    (h = 192)."""
    for h, traced in ((150, True), (192, False)):
        img = np.zeros((h, 64), np.uint8)
        img[h - 1, 10:20] = 1
        r = O.centerline2d(img, 0, 0)
        assert (len(r["idx"]) > 1) == traced, h


def test_coordinates_beyond_int16_are_rejected():
    """Segment points are int16 (the reference's point2D/point3D hold shorts,
    2D/Structs.h:15-22): tracing a side longer than 32767 is refused with
    BGT_ERR_UNSUPPORTED rather than wrapping, and 32767 itself is fine.
    CenterlineExtraction refuses the same once its padded frame exceeds it."""
    import begputhinning_amd as B
    with pytest.raises(B.BGTError, match="int16"):
        B.extract_segments_flat(np.zeros((1, 32768), np.uint8))
    with pytest.raises(B.BGTError, match="int16"):
        B.extract_segments_flat(np.zeros((32768, 1, 1), np.uint8))
    r = B.extract_segments_flat(np.zeros((3, 32767), np.uint8))
    assert len(r["start"]) == 1 and len(r["x"]) == 0
    # 32740 pixels wide pads to 32768: past int16
    with pytest.raises(B.BGTError, match="int16"):
        B.CenterlineExtraction(np.zeros((4, 32740), np.uint8), 10, 3)

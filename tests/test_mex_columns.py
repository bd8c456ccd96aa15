"""The MEX gateways' output layout (2D/Matlab2DThinning.cpp:163-175,
3D/Matlab3DThinning.cpp:157-170) through bgt_mex_columns: segment i of a
centerline becomes an n_i x 2 [y x] (n_i x 3 [y x z]) double matrix stored
column-major, i.e. numpy's ravel(order="F") of the stacked columns.  Runs on
the host (no GPU): the centerlines come from the oracle."""
import ctypes

import numpy as np
import pytest

import _oracle as O


def _raw(cl, three_d):
    from begputhinning_amd._lib import check, lib
    x, y, idx = cl["x"], cl["y"], cl["idx"].astype(np.uint32)
    z = cl["z"] if three_d else None
    nc = 3 if three_d else 2
    out = np.full(max(1, int(idx[-1]) * nc), np.nan)
    p = lambda a: ctypes.c_void_p(a.ctypes.data) if a is not None else None
    check(lib().bgt_mex_columns(p(x), p(y), p(z), p(idx), len(idx) - 1, p(out)))
    return out


@pytest.mark.parametrize("three_d", [False, True])
def test_column_major_y_x_z(three_d):
    from begputhinning_amd import mex_columns, synth
    if three_d:
        cl = O.centerline3d(synth.tree3d(60, 50, 40, 5), 4, 2)
        cols = lambda s, e: [cl["y"][s:e], cl["x"][s:e], cl["z"][s:e]]
    else:
        cl = O.centerline2d(synth.workload1(), 10, 3)
        cols = lambda s, e: [cl["y"][s:e], cl["x"][s:e]]
    idx = cl["idx"]
    assert len(idx) > 10
    raw = _raw(cl, three_d)
    nc = 3 if three_d else 2
    mats = mex_columns((cl["x"], cl["y"], cl["z"], idx) if three_d else (cl["x"], cl["y"], idx))
    assert len(mats) == len(idx) - 1
    for i in range(len(idx) - 1):
        s, e = int(idx[i]), int(idx[i + 1])
        want = np.stack(cols(s, e), axis=1).astype(np.float64)  # n x nc, MATLAB's [y x (z)]
        assert np.array_equal(raw[s * nc:e * nc], want.ravel(order="F")), i
        assert mats[i].shape == (e - s, nc) and np.array_equal(mats[i], want), i


def test_empty_and_bad_starts():
    from begputhinning_amd import BGTError
    from begputhinning_amd._lib import check, lib
    x = np.zeros(1, np.float32)
    idx = np.zeros(1, np.uint32)
    out = np.zeros(1)
    p = lambda a: ctypes.c_void_p(a.ctypes.data)
    check(lib().bgt_mex_columns(p(x), p(x), None, p(idx), 0, p(out)))
    bad = np.array([0, 5, 3], np.uint32)
    with pytest.raises(BGTError):
        check(lib().bgt_mex_columns(p(np.zeros(5, np.float32)), p(np.zeros(5, np.float32)), None,
                                    p(bad), 2, p(np.zeros(10))))

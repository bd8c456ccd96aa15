"""Pins of the oracle's thinning drivers (or2d_thin / or3d_thin), the
iteration semantics of 2D/CUDA_ThinningKernelMultiBlock2D.cuh:99-154 and
3D/CUDA_ThinningKernel3D_Linear.cuh:690-782 (synchronous subiterations, zero
outside the frame, the 3D protection mask taken at the iteration's start):

  * brute force: a per-pixel restatement -- every subiteration looks each
    pixel's 3x3 (3x3x3) neighbourhood up in the predicate tables (pinned on
    their own in test_oracle_pred2d/3d), zero outside the frame, synchronous
    update, directions in the reference's order, the 3D protection mask taken
    at the start of each iteration -- run on every 4x4 pattern inside a zero
    frame (2^16 images) and on random small frames and volumes, compared with
    the word-parallel oracle including the iteration count;
  * order sensitivity: a frame where another direction order gives another
    skeleton, so the brute force really checks the order;
  * fixed point: one more iteration of the result deletes nothing;
  * topology: connected components of foreground and background are kept.
"""
import numpy as np
import pytest

import _oracle as O

# clockwise neighbour offsets (dx, dy) of x0..x7: N, NE, E, SE, S, SW, W, NW
OFF2 = [(0, -1), (1, -1), (1, 0), (1, 1), (0, 1), (-1, 1), (-1, 0), (-1, -1)]
# 3D neighbour i <-> (dx, dy, dz), raster (dz, dy, dx) order without the centre
OFF3 = [(j % 3 - 1, (j // 3) % 3 - 1, j // 9 - 1) for j in range(27) if j != 13]


def _shift(a, dx, dy, dz=None):
    """a's neighbour at (x+dx, y+dy[, z+dz]) for every pixel, zero outside."""
    p = np.pad(a, [(0, 0)] + [(1, 1)] * (a.ndim - 1))
    if dz is None:
        h, w = a.shape[1:]
        return p[:, 1 + dy:1 + dy + h, 1 + dx:1 + dx + w]
    d, h, w = a.shape[1:]
    return p[:, 1 + dz:1 + dz + d, 1 + dy:1 + dy + h, 1 + dx:1 + dx + w]


def brute2d(imgs, order=(0, 1, 2, 3), max_iter=500):
    """Batch (n, h, w) of 0/1 frames -> (skeletons, iterations incl. the last)."""
    tabs = [O.table2d(d).astype(bool) for d in range(4)]
    a = imgs.astype(np.uint16).copy()
    iters = np.zeros(len(a), np.int64)
    live = np.ones(len(a), bool)
    for _ in range(max_iter):
        before = a.copy()
        for d in order:
            idx = a << 8
            for i, (dx, dy) in enumerate(OFF2):
                idx = idx | (_shift(a, dx, dy) << i)
            a = a & ~tabs[d][idx].astype(np.uint16)
        iters += live
        changed = (a != before).reshape(len(a), -1).any(axis=1)
        live &= changed
        if not live.any():
            break
    return a.astype(np.uint8), iters


def _lookup3(table_u64, idx):
    return ((table_u64[idx >> 6] >> (idx & 63).astype(np.uint64)) & np.uint64(1)).astype(bool)


def brute3d(vols, max_iter=500):
    tabs = [O.table3d(d) for d in range(6)]  # U, N, E, D, S, W: C_d over 26 neighbours
    a = vols.astype(np.int64).copy()
    iters = np.zeros(len(a), np.int64)
    live = np.ones(len(a), bool)

    def comb(state, d):
        idx = np.zeros_like(state)
        for i, (dx, dy, dz) in enumerate(OFF3):
            idx |= _shift(state, dx, dy, dz) << i
        return _lookup3(tabs[d], idx)

    for _ in range(max_iter):
        before = a.copy()
        prot = a.astype(bool)
        for d in range(6):
            prot &= ~comb(a, d)
        for d in range(6):
            a = np.where(comb(a, d) & a.astype(bool) & ~prot, 0, a)
        iters += live
        changed = (a != before).reshape(len(a), -1).any(axis=1)
        live &= changed
        if not live.any():
            break
    return a.astype(np.uint8), iters


def test_every_4x4_pattern_in_a_zero_frame():
    codes = np.arange(1 << 16)
    imgs = np.zeros((len(codes), 6, 6), np.uint8)
    for k in range(16):
        imgs[:, 1 + k // 4, 1 + k % 4] = (codes >> k) & 1
    ref, it = brute2d(imgs)
    for n in range(len(imgs)):
        got, git = O.thin2d(imgs[n])
        assert np.array_equal(got, ref[n]) and git == it[n], n


def test_random_small_frames_2d():
    rng = np.random.default_rng(20)
    for trial in range(40):
        h, w = (int(v) for v in rng.integers(1, 14, 2))
        imgs = (rng.random((8, h, w)) < rng.uniform(0.3, 0.9)).astype(np.uint8)
        ref, it = brute2d(imgs)
        for n in range(len(imgs)):
            got, git = O.thin2d(imgs[n])
            assert np.array_equal(got, ref[n]) and git == it[n], (trial, n, h, w)


def test_direction_order_matters_2d():
    """Some frame's skeleton depends on the direction order; the oracle takes
    the reference's N, S, E, W."""
    rng = np.random.default_rng(21)
    imgs = (rng.random((400, 7, 7)) < 0.6).astype(np.uint8)
    ref, _ = brute2d(imgs)
    alt, _ = brute2d(imgs, order=(0, 2, 1, 3))
    diff = [n for n in range(len(imgs)) if not np.array_equal(ref[n], alt[n])]
    assert diff, "no order-sensitive frame found"
    n = diff[0]
    got, _ = O.thin2d(imgs[n])
    assert np.array_equal(got, ref[n]) and not np.array_equal(got, alt[n])


def test_random_small_volumes_3d():
    rng = np.random.default_rng(30)
    for trial in range(12):
        d, h, w = (int(v) for v in rng.integers(1, 7, 3))
        vols = (rng.random((6, d, h, w)) < rng.uniform(0.3, 0.9)).astype(np.uint8)
        ref, it = brute3d(vols)
        for n in range(len(vols)):
            got, git = O.thin3d(vols[n])
            assert np.array_equal(got, ref[n]) and git == it[n], (trial, n, d, h, w)


def test_fixed_point_and_subset():
    from begputhinning_amd import synth
    img = synth.polylines(160, 120, 5, n_lines=10)
    sk, _ = O.thin2d(img)
    assert not (sk & ~img).any()
    _, changed = O.iterate2d(sk, 1)
    assert changed == 0
    vol = synth.tree3d(40, 36, 30, 6)
    sk3, _ = O.thin3d(vol)
    assert not (sk3 & ~vol).any()
    _, changed3 = O.iterate3d(sk3, 1)
    assert changed3 == 0


def test_topology_preserved():
    nd = pytest.importorskip("scipy.ndimage")
    from begputhinning_amd import synth
    rng = np.random.default_rng(40)
    eight = np.ones((3, 3), int)
    four = nd.generate_binary_structure(2, 1)
    for k in range(6):
        img = synth.polylines(90, 70, 100 + k, n_lines=6)
        img[rng.random(img.shape) < 0.05] = 0  # holes
        sk, _ = O.thin2d(img)
        pad_i, pad_s = np.pad(img, 1), np.pad(sk, 1)
        assert nd.label(pad_i, eight)[1] == nd.label(pad_s, eight)[1]
        assert nd.label(1 - pad_i, four)[1] == nd.label(1 - pad_s, four)[1]
    s26 = np.ones((3, 3, 3), int)
    s6 = nd.generate_binary_structure(3, 1)
    for k in range(3):
        vol = synth.tree3d(30, 28, 26, 200 + k)
        sk, _ = O.thin3d(vol)
        pad_v, pad_s = np.pad(vol, 1), np.pad(sk, 1)
        assert nd.label(pad_v, s26)[1] == nd.label(pad_s, s26)[1]
        assert nd.label(1 - pad_v, s6)[1] == nd.label(1 - pad_s, s6)[1]

"""The drop-in C++ headers on the MI355X, called the way a user of the
reference's API calls them (tests/dropin/dropin_main.cpp, built against
include/2D and include/3D): Thinning2D/Thinning3D for every pixel type and
CenterlineExtraction, against the oracle, bit-exact (floats as bits)."""
import os
import subprocess
import tempfile

import numpy as np
import pytest

import _oracle as O

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

DRV = os.path.join(os.path.dirname(os.path.abspath(__file__)), "dropin")
NAMES = {np.uint8: "uint8", np.int8: "int8", np.uint16: "uint16", np.int16: "int16",
         np.uint32: "uint32", np.int32: "int32", np.float32: "float32", np.float64: "float64",
         np.bool_: "bool"}


@pytest.fixture(scope="module", autouse=True)
def _gpu():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    if not os.path.exists(os.path.join(DRV, "dropin2d")):
        pytest.skip("drop-in drivers not built")


def run(dim, args, tmp):
    print(f"[dropin{dim}d] {' '.join(map(str, args))}", flush=True)
    r = subprocess.run([os.path.join(DRV, f"dropin{dim}d")] + [str(a) for a in args],
                       capture_output=True, text=True, timeout=120, cwd=tmp)
    assert r.returncode == 0, r.stderr
    assert "bgt:" not in r.stderr, r.stderr


def values(rng, dt, base):
    if dt == np.bool_:
        return base.copy()
    if np.issubdtype(dt, np.floating):
        vals = rng.choice([0.0, 0.5, 1.0, 2.5, -3.0, np.nan], size=base.shape)
        return np.where(base, vals, 0.0).astype(dt)
    info = np.iinfo(dt)
    vals = rng.integers(max(info.min, -100), min(int(info.max), 100) + 1, size=base.shape)
    return np.where(base, vals, 0).astype(dt)


@pytest.mark.parametrize("dt", list(NAMES))
def test_thinning2d_header(dt):
    rng = np.random.default_rng(5)
    data = values(rng, dt, rng.random((70, 110)) < 0.6)
    if np.issubdtype(dt, np.floating):
        mask = np.nan_to_num(data.astype(np.float64), nan=0.0) >= 1.0  # (unsigned)v > 0
    else:
        mask = data != 0
    ref, _ = O.thin2d(mask.astype(np.uint8))
    with tempfile.TemporaryDirectory() as tmp:
        data.tofile(os.path.join(tmp, "in.raw"))
        run(2, ["thin", NAMES[dt], 110, 70, "in.raw", "out.raw"], tmp)
        out = np.fromfile(os.path.join(tmp, "out.raw"), dt).reshape(70, 110)
    assert np.array_equal(out.astype(np.float64) != 0, ref.astype(bool))
    assert set(np.unique(out.astype(np.float64)).tolist()) <= {0.0, 1.0}


@pytest.mark.parametrize("dt", [np.uint8, np.int8, np.int16, np.float32, np.bool_])
def test_thinning3d_header(dt):
    rng = np.random.default_rng(6)
    data = values(rng, dt, rng.random((18, 26, 34)) < 0.55)
    mask = (np.nan_to_num(data.astype(np.float64), nan=0.0) > 0) if dt != np.bool_ else data
    ref, _ = O.thin3d(mask.astype(np.uint8))
    with tempfile.TemporaryDirectory() as tmp:
        data.tofile(os.path.join(tmp, "in.raw"))
        run(3, ["thin", NAMES[dt], 34, 26, 18, "in.raw", "out.raw"], tmp)
        out = np.fromfile(os.path.join(tmp, "out.raw"), dt).reshape(18, 26, 34)
    assert np.array_equal(out.astype(np.float64) != 0, ref.astype(bool))


def parse_centerline(path, dim):
    lines = open(path).read().split("\n")
    ns = int(lines[0])
    idx = np.array([int(v) for v in lines[1].split()], np.uint32)
    assert idx.size == ns + 1
    pts = [l.split() for l in lines[2:2 + int(idx[-1])]]
    cols = [np.array([np.float32(float(p[c])) for p in pts], np.float32) for c in range(dim)]
    return ns, idx, cols


def same_floats(a, b):
    return a.dtype == b.dtype == np.float32 and np.array_equal(a.view(np.uint32), b.view(np.uint32))


@pytest.mark.parametrize("dt,mx,span", [(np.uint8, 10, 3), (np.float32, 0, 0), (np.int32, 6, 4)])
def test_centerline2d_header(dt, mx, span):
    from begputhinning_amd import synth
    img = synth.tree2d(300, 260, 31)
    data = img.astype(dt) * (3 if dt != np.bool_ else 1)
    ref = O.centerline2d(img, mx, span)
    with tempfile.TemporaryDirectory() as tmp:
        data.tofile(os.path.join(tmp, "in.raw"))
        run(2, ["centerline", NAMES[dt], 300, 260, mx, span, "in.raw", "out.txt"], tmp)
        ns, idx, (x, y) = parse_centerline(os.path.join(tmp, "out.txt"), 2)
    assert np.array_equal(idx, ref["idx"])
    assert same_floats(x, ref["x"]) and same_floats(y, ref["y"])


def test_centerline3d_header():
    from begputhinning_amd import synth
    vol = synth.tree3d(70, 60, 50, 91)
    ref = O.centerline3d(vol, 5, 2)
    with tempfile.TemporaryDirectory() as tmp:
        vol.tofile(os.path.join(tmp, "in.raw"))
        run(3, ["centerline", "uint8", 70, 60, 50, 5, 2, "in.raw", "out.txt"], tmp)
        ns, idx, (x, y, z) = parse_centerline(os.path.join(tmp, "out.txt"), 3)
    assert np.array_equal(idx, ref["idx"])
    assert same_floats(x, ref["x"]) and same_floats(y, ref["y"]) and same_floats(z, ref["z"])

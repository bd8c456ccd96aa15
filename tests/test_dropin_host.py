"""The drop-in C++ API (include/2D, include/3D) on the host, no GPU needed:
the headers compile for every pixel type of the reference's instantiations,
the tracers called through them equal the oracle and the reference's own
tracers (compiled on the host in oracle/_ref), and the thinning entry points
fail cleanly (zeroed output, message in bgt_last_error, no exit) when the GPU
is unavailable."""
import os
import shutil
import subprocess
import tempfile

import numpy as np
import pytest

import _oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
INC = os.path.join(ROOT, "include")
DRV = os.path.join(ROOT, "tests", "dropin")
TYPES = ["unsigned char", "signed char", "char", "unsigned short", "short", "unsigned int", "int",
         "float", "double", "bool"]


def _have_driver(dim):
    return os.path.exists(os.path.join(DRV, f"dropin{dim}d"))


def _run(dim, args, tmp):
    exe = os.path.join(DRV, f"dropin{dim}d")
    return subprocess.run([exe] + [str(a) for a in args], capture_output=True, text=True,
                          timeout=120, cwd=tmp)


def _parse_segments(path, dim):
    xs, ys, zs, start, ends = [], [], [], [0], []
    for line in open(path):
        parts = line.split()
        ends.append(1 if parts[0] == "E" else 0)
        for p in parts[1:]:
            c = [int(v) for v in p.split(",")]
            xs.append(c[0]); ys.append(c[1])
            if dim == 3:
                zs.append(c[2])
        start.append(len(xs))
    return {"x": np.array(xs, np.int16), "y": np.array(ys, np.int16),
            "z": np.array(zs, np.int16) if dim == 3 else None,
            "start": np.array(start, np.uint32), "is_end": np.array(ends, np.uint8)}


def _same(a, b, keys):
    for k in keys:
        assert np.array_equal(np.asarray(a[k]), np.asarray(b[k])), k


@pytest.mark.skipif(shutil.which("g++") is None, reason="no host C++ compiler")
@pytest.mark.parametrize("dim", [2, 3])
def test_headers_compile_for_all_pixel_types(dim, tmp_path):
    lines = []
    if dim == 2:
        lines += ['#include "2D/LibThinning2D.h"', '#include "2D/ExtractSegments.h"',
                  '#include "2D/CenterlineExtraction.h"']
        for t in TYPES:
            lines += [f"template void Thinning2D<{t}>(int, int, {t} *, {t} *);",
                      f"template std::vector<SSegment *> ExtractSegmentsV<{t}>({t} *, int, int);",
                      f"template unsigned CenterlineExtraction<{t}>({t} *, int, int, int, int, "
                      "float **, float **, unsigned **);"]
    else:
        lines += ['#include "3D/LibThinning3D.h"', '#include "3D/ExtractSegments.h"',
                  '#include "3D/CenterlineExctraction.h"', '#include "3D/CenterlineExtraction.h"']
        for t in TYPES:
            lines += [f"template void Thinning3D<{t}>(int, int, int, {t} *, {t} *);",
                      f"template std::vector<SSegment *> ExtractSegmentsV<{t}>({t} *, int, int, int);",
                      f"template std::vector<std::vector<point3D>> ExtractSegments<{t}>({t} *, int, "
                      "int, int);",
                      f"template unsigned CenterlineExtraction<{t}>({t} *, int, int, int, int, int, "
                      "float **, float **, float **, unsigned **);"]
    src = tmp_path / "inst.cpp"
    src.write_text("\n".join(lines) + "\n")
    r = subprocess.run(["g++", "-std=c++14", "-Wall", "-Wextra", "-Werror", "-fsyntax-only",
                        "-I" + INC, str(src)], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


def _skeleton2d():
    from begputhinning_amd import synth
    img = synth.polylines(180, 150, synth.item_seed(1, 7), n_lines=12)
    return O.thin2d(img)[0]


def _skeleton3d():
    from begputhinning_amd import synth
    vol = synth.tree3d(48, 40, 36, synth.item_seed(3, 7))
    return O.thin3d(vol)[0]


@pytest.mark.skipif(not _have_driver(2), reason="drop-in drivers not built")
def test_trace2d_through_dropin_header():
    sk = _skeleton2d()
    h, w = sk.shape
    with tempfile.TemporaryDirectory() as tmp:
        sk.astype(np.uint8).tofile(os.path.join(tmp, "in.raw"))
        r = _run(2, ["trace", w, h, "in.raw", "out.txt"], tmp)
        assert r.returncode == 0, r.stderr
        got = _parse_segments(os.path.join(tmp, "out.txt"), 2)
    _same(got, O.extract2d(sk), ("x", "y", "start", "is_end"))
    if O.ref_available("libref2d.so"):
        _same(got, O.ref_extract2d(sk), ("x", "y", "start", "is_end"))


@pytest.mark.skipif(not _have_driver(3), reason="drop-in drivers not built")
@pytest.mark.parametrize("plain", [False, True])
def test_trace3d_through_dropin_header(plain):
    sk = _skeleton3d()
    d, h, w = sk.shape
    with tempfile.TemporaryDirectory() as tmp:
        sk.astype(np.uint8).tofile(os.path.join(tmp, "in.raw"))
        r = _run(3, ["trace_plain" if plain else "trace", w, h, d, "in.raw", "out.txt"], tmp)
        assert r.returncode == 0, r.stderr
        got = _parse_segments(os.path.join(tmp, "out.txt"), 3)
    keys = ("x", "y", "z", "start") if plain else ("x", "y", "z", "start", "is_end")
    _same(got, O.extract3d(sk), keys)
    if O.ref_available("libref3d.so"):
        ref = O.ref_extract3d_plain(sk) if plain else O.ref_extract3d(sk)
        _same(got, ref, keys)


@pytest.mark.skipif(not _have_driver(2), reason="drop-in drivers not built")
def test_thinning_without_gpu_fails_cleanly():
    """No GPU here: Thinning2D through the header returns, zero-fills out and
    leaves the message in bgt_last_error (the reference has no error path)."""
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    img = (np.random.default_rng(0).random((20, 30)) < 0.5).astype(np.uint8)
    with tempfile.TemporaryDirectory() as tmp:
        img.tofile(os.path.join(tmp, "in.raw"))
        np.full(img.shape, 7, np.uint8).tofile(os.path.join(tmp, "out.raw"))
        r = _run(2, ["thin", "uint8", 30, 20, "in.raw", "out.raw"], tmp)
        assert r.returncode == 0, r.stderr
        assert "bgt:" in r.stderr
        out = np.fromfile(os.path.join(tmp, "out.raw"), np.uint8)
        assert out.size == img.size and not out.any()

"""Pins of the oracle's 3D predicates (3D/CUDA_SubiterationKernels3D.cuh:14-234).

  * exhaustive 2^26-entry tables against the reference header compiled on the
    host (oracle/_ref/libref_pred3d.so, built by oracle/Makefile where the
    reference tree is present);
  * committed SHA-256 + popcount of each table and 4096 sampled entries
    (tests/golden/pred3d_tables.json), so the check also runs where the
    reference build is absent;
  * geometry: every direction is the Up rule seen through a proper rotation
    of the cube that takes the direction's face neighbour to z-1 (the face
    each rule requires to be background: U z-1, N y-1, E x-1, D z+1, S y+1,
    W x+1).  The six rules are transcribed separately, so a typo in one
    breaks this.
"""
import hashlib
import json
import os

import numpy as np
import pytest

import _oracle as O

GOLD = os.path.join(os.path.dirname(__file__), "golden", "pred3d_tables.json")
NAMES = "UNEDSW"
# neighbour i <-> offset (dx, dy, dz), raster (dz, dy, dx) order without centre
OFFS = [(j % 3 - 1, (j // 3) % 3 - 1, j // 9 - 1) for j in range(27) if j != 13]
FACE = {"U": (0, 0, -1), "N": (0, -1, 0), "E": (-1, 0, 0),
        "D": (0, 0, 1), "S": (0, 1, 0), "W": (1, 0, 0)}


def rotation_to_up(face):
    """A proper rotation matrix R (det +1) with R @ face = (0, 0, -1)."""
    import itertools
    for perm in itertools.permutations(range(3)):
        for signs in itertools.product((1, -1), repeat=3):
            R = np.zeros((3, 3), int)
            for k in range(3):
                R[k, perm[k]] = signs[k]
            if round(np.linalg.det(R)) == 1 and tuple(R @ np.array(face)) == (0, 0, -1):
                return R
    raise AssertionError(face)


def bits(table_u64):
    return np.unpackbits(table_u64.view(np.uint8), bitorder="little")


@pytest.fixture(scope="module")
def tables():
    return [O.table3d(d) for d in range(6)]


def test_committed_hashes(tables):
    with open(GOLD) as f:
        gold = json.load(f)
    for d, name in enumerate(NAMES):
        t = tables[d]
        assert hashlib.sha256(t.tobytes()).hexdigest() == gold[name]["sha256"]
        assert int(bits(t).sum()) == gold[name]["popcount"]
        b = bits(t)
        for c, v in gold[name]["samples"]:
            assert b[c] == v


@pytest.mark.skipif(not O.ref_available("libref_pred3d.so"),
                    reason="reference predicates not built (oracle/_ref)")
def test_against_reference_header(tables):
    for d in range(6):
        assert np.array_equal(tables[d].view(np.uint32), O.ref_table3d(d)), NAMES[d]


def test_rotations_of_up(tables):
    up = bits(tables[0])
    idx = {o: i for i, o in enumerate(OFFS)}
    for d, name in enumerate(NAMES):
        if name == "U":
            continue
        R = rotation_to_up(FACE[name])
        pi = [idx[tuple(int(v) for v in R @ np.array(o))] for o in OFFS]
        td = bits(tables[d])
        chunk = 1 << 22
        for base in range(0, 1 << 26, chunk):
            c = np.arange(base, base + chunk, dtype=np.int64)
            m = np.zeros_like(c)
            for i in range(26):
                m |= ((c >> i) & 1) << pi[i]
            assert np.array_equal(td[base:base + chunk], up[m]), name


def test_face_neighbour_background(tables):
    for d, name in enumerate(NAMES):
        face_i = OFFS.index(FACE[name])
        b = bits(tables[d]).reshape(-1)
        # any configuration with the face neighbour set is never deletable
        c = np.arange(1 << 26, dtype=np.int64)
        assert not b[(c >> face_i) & 1 == 1].any(), name

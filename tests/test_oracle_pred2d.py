"""Pins of the oracle's 2D predicates (2D/CUDA_SubiterationKernels2D.cuh:15-119).

The reference header needs CUDA lane shuffles for ShiftLeft/ShiftRight
(2D/CUDA_DeviceHelperFunctions.cuh:34-50), so it cannot be compiled on the host
without stand-ins; the four directions are pinned instead by:
  * rotation: E, S, W are N with the clockwise neighbour labels x0..x7 advanced
    by 2, 4, 6 -- the four directions are transcribed separately, so a dropped
    term, flipped sign or swapped operand in one of them breaks the relation;
  * topology: every configuration a direction deletes is a simple point
    (8-connected foreground, 4-connected background) whose face neighbour in
    that direction is background;
  * the committed tables (regression fixture).
"""
import json
import os

import numpy as np
import pytest

import _oracle as O

GOLD = os.path.join(os.path.dirname(__file__), "golden", "pred2d_tables.json")
# clockwise neighbour offsets (dx, dy) of x0..x7: N, NE, E, SE, S, SW, W, NW
OFF = [(0, -1), (1, -1), (1, 0), (1, 1), (0, 1), (-1, 1), (-1, 0), (-1, -1)]


def rot(c, k):
    """configuration whose x_i is c's x_(i+k)"""
    out = c & 0x100
    for i in range(8):
        out |= ((c >> ((i + k) % 8)) & 1) << i
    return out


def grid(c):
    g = np.zeros((3, 3), np.uint8)
    g[1, 1] = (c >> 8) & 1
    for i, (dx, dy) in enumerate(OFF):
        g[1 + dy, 1 + dx] = (c >> i) & 1
    return g


def components(cells, adj):
    cells = set(cells)
    seen, n = set(), 0
    for c in cells:
        if c in seen:
            continue
        n += 1
        stack = [c]
        seen.add(c)
        while stack:
            y, x = stack.pop()
            for dy, dx in adj:
                q = (y + dy, x + dx)
                if q in cells and q not in seen:
                    seen.add(q)
                    stack.append(q)
    return n


A8 = [(dy, dx) for dy in (-1, 0, 1) for dx in (-1, 0, 1) if dy or dx]
A4 = [(-1, 0), (1, 0), (0, -1), (0, 1)]


def is_simple(c):
    """Local simplicity in the 3x3 window, computed by flood fill."""
    g = grid(c)
    nb = [(y, x) for y in range(3) for x in range(3) if (y, x) != (1, 1)]
    fg = [p for p in nb if g[p]]
    bg = [p for p in nb if not g[p]]
    # 8-components of foreground neighbours (all are 8-adjacent to the centre)
    c8 = components(fg, A8)
    # 4-components of background that contain a 4-neighbour of the centre
    bgs = set(bg)
    seen, c4 = set(), 0
    for p in [(0, 1), (1, 0), (1, 2), (2, 1)]:
        if p in bgs and p not in seen:
            c4 += 1
            stack = [p]
            seen.add(p)
            while stack:
                y, x = stack.pop()
                for dy, dx in A4:
                    q = (y + dy, x + dx)
                    if q in bgs and q not in seen:
                        seen.add(q)
                        stack.append(q)
    return c8 == 1 and c4 == 1


def tables():
    return [O.table2d(d) for d in range(4)]  # N, S, E, W


def test_rotation_relations():
    N, S, E, W = tables()
    for c in range(512):
        assert E[c] == N[rot(c, 2)], c
        assert S[c] == N[rot(c, 4)], c
        assert W[c] == N[rot(c, 6)], c


def test_centre_required():
    for t in tables():
        assert not t[:256].any()


@pytest.mark.parametrize("d,face", [(0, 0), (1, 4), (2, 2), (3, 6)])
def test_deleted_points_are_simple_border_points(d, face):
    t = O.table2d(d)
    assert t.sum() > 0
    for c in np.nonzero(t)[0]:
        c = int(c)
        assert not (c >> face) & 1, (d, c)
        assert is_simple(c), (d, c)


def test_committed_tables():
    with open(GOLD) as f:
        gold = json.load(f)
    for name, d in (("N", 0), ("S", 1), ("E", 2), ("W", 3)):
        assert O.table2d(d).tolist() == gold[name]

"""The 2D deletion rule as the pass kernel computes it: the three-input gate
network of north_keep in csrc/thin2d_kernels.h, read from the header as it
ships (table constants and wiring), evaluated on all 512 3x3 configurations
for each direction through keep<D>'s label rotation, equals the oracle's
deletion table of the reference template (2D/CUDA_SubiterationKernels2D.cuh:
15-119).  A wrong table bit, a swapped operand or a wrong rotation fails it."""
import os
import re

import numpy as np

import _oracle as O

HDR = os.path.join(os.path.dirname(__file__), "..", "begputhinning_amd", "csrc", "thin2d_kernels.h")


def _function_body(src, name):
    start = src.index(name + "(")
    open_ = src.index("{", start)
    depth, i = 0, open_
    while True:
        if src[i] == "{":
            depth += 1
        elif src[i] == "}":
            depth -= 1
            if depth == 0:
                return src[open_ + 1:i]
        i += 1


def _network():
    src = open(HDR).read()
    body = re.sub(r"//[^\n]*", "", _function_body(src, "uint32_t north_keep"))
    steps = []
    for line in body.split(";"):
        m = re.search(r"(?:const uint32_t (\w+) =|return)\s*lut3<(0x[0-9a-fA-F]+)>\((\w+),\s*(\w+),\s*(\w+)\)",
                      line)
        if m:
            steps.append((m.group(1) or "return", int(m.group(2), 16), m.group(3), m.group(4), m.group(5)))
    assert steps and steps[-1][0] == "return", steps
    return steps


def _keep(steps, env):
    env = dict(env)
    for out, table, a, b, c in steps:
        v = (table >> ((env[a] << 2) | (env[b] << 1) | env[c])) & 1
        if out == "return":
            return v
        env[out] = v
    raise AssertionError("no return")


def _rotations():
    """keep<D>'s argument order: the labels north_keep sees for direction D."""
    src = open(HDR).read()
    body = _function_body(src, "uint32_t keep")
    rot = {}
    for m in re.finditer(r"(?:if \(D == (\d)\) )?return north_keep\(ctr, ([x0-9, ]+)\)", body):
        d = int(m.group(1)) if m.group(1) else 3
        rot[d] = [int(t.strip()[1:]) for t in m.group(2).split(",")]
    assert sorted(rot) == [0, 1, 2, 3], rot
    return rot


def test_gate_count():
    # eight gates for F and one for the centre: nine instructions per 32
    # pixels and subiteration
    assert len(_network()) == 9


def test_network_equals_reference_rule_all_directions():
    steps = _network()
    rot = _rotations()
    for d in range(4):
        table = O.table2d(d)
        got = np.zeros(512, np.uint8)
        for c in range(512):
            x = [(c >> i) & 1 for i in range(8)]
            p = (c >> 8) & 1
            env = {"ctr": p}
            for k, lab in enumerate(rot[d]):
                env[f"x{k}"] = x[lab]
            got[c] = p & (1 - _keep(steps, env))
        bad = np.flatnonzero(got != table)
        assert bad.size == 0, (d, bad[:8])

"""Seeded synthetic inputs of the thinning workloads (generated by the C++
generators in libbgt.so so that tests, benches and fixtures agree bit-for-bit).

Workloads (base seed 0xBE6E0000 + workload number, item i):
  1  512x512 polylines: 60 polylines of 3-8 vertices, tube width U[3,15),
     salt 0.0005
  2  1024x1024 branching trees: depth 6, length U[40,200), radius 8 -> 1,
     added until 5% foreground (about 6% realised), salt 0.001
  3  256^3 vessel trees: depth 7, length U[10,60), radius 8 -> 1, plus 20
     axis-aligned ellipsoids with semi-axes U[3,12), trees added until 3%
     foreground
  4  mixed batch: 512 images with sides U{128..2048} (workload-1 content) and
     16 volumes with sides U{64..384} (workload-3 content)
  5  serving mix: 80% images with sides U{256..1024} (workload 1 or 2 content),
     20% volumes with sides U{64..512} (workload-3 content)
"""
import ctypes

import numpy as np

from ._lib import check, lib


def item_seed(workload, item=0):
    return int(lib().bgt_synth_item_seed(workload, item))


def polylines(w, h, seed, n_lines=60, p_salt=0.0005):
    a = np.empty((h, w), np.uint8)
    check(lib().bgt_synth_polylines(ctypes.c_void_p(a.ctypes.data), w, h, seed, n_lines, p_salt))
    return a


def tree2d(w, h, seed, depth=6, p_salt=0.001, target_fraction=0.05):
    a = np.empty((h, w), np.uint8)
    check(lib().bgt_synth_tree2d(ctypes.c_void_p(a.ctypes.data), w, h, seed, depth, p_salt,
                                 target_fraction))
    return a


def tree3d(w, h, d, seed, depth=7, n_ellipsoids=20, target_fraction=0.03):
    a = np.empty((d, h, w), np.uint8)
    check(lib().bgt_synth_tree3d(ctypes.c_void_p(a.ctypes.data), w, h, d, seed, depth,
                                 n_ellipsoids, target_fraction))
    return a


def workload1(item=0):
    return polylines(512, 512, item_seed(1, item))


def workload2(item=0):
    return tree2d(1024, 1024, item_seed(2, item))


def workload3(item=0, side=256):
    return tree3d(side, side, side, item_seed(3, item))


def batch_dims(n2=512, lo2=128, hi2=2048, n3=16, lo3=64, hi3=384, workload=4):
    d2 = np.zeros(2 * n2, np.int32)
    d3 = np.zeros(3 * n3, np.int32)
    seed = item_seed(workload, 1 << 20)
    check(lib().bgt_synth_batch_dims(seed, n2, lo2, hi2,
                                     d2.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), n3, lo3,
                                     hi3, d3.ctypes.data_as(ctypes.POINTER(ctypes.c_int))))
    return d2.reshape(n2, 2), d3.reshape(n3, 3)


def workload4_item(i, dims2, dims3):
    """Item i of the mixed batch: images first (w, h), then volumes (w, h, d)."""
    if i < len(dims2):
        w, h = (int(v) for v in dims2[i])
        return polylines(w, h, item_seed(4, i))
    w, h, d = (int(v) for v in dims3[i - len(dims2)])
    return tree3d(w, h, d, item_seed(4, i))


def arrivals(n, qps, seed):
    gaps = np.zeros(n, np.float64)
    mix = np.zeros(n, np.float64)
    check(lib().bgt_synth_arrivals(seed, n, qps, gaps.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                   mix.ctypes.data_as(ctypes.POINTER(ctypes.c_double))))
    return gaps, mix

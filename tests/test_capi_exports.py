"""The C-ABI library loads on a host without a GPU and exports every function
that include/bgt.h declares (no compute calls: those need the GPU), and the
Python binding declares a signature for each of them."""
import ctypes
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "bgt.h")
LIB = os.path.join(ROOT, "begputhinning_amd", "libbgt.so")


def declared():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"\b(bgt_[a-z0-9_]+)\s*\(", text)
    return sorted(set(names))


def test_header_declares_the_api():
    names = declared()
    for must in ("bgt_thinning2d", "bgt_thinning3d", "bgt_thin2d_device", "bgt_thin3d_device",
                 "bgt_extract_segments2d", "bgt_extract_segments3d", "bgt_centerline2d",
                 "bgt_centerline3d", "bgt_batch_run", "bgt_last_error"):
        assert must in names


@pytest.mark.skipif(not os.path.exists(LIB), reason="library not built")
def test_library_exports_every_declared_symbol():
    L = ctypes.CDLL(LIB)
    missing = [n for n in declared() if not hasattr(L, n)]
    assert not missing, missing
    r = subprocess.run(["nm", "-D", "--defined-only", LIB], capture_output=True, text=True)
    if r.returncode == 0:
        exported = set(re.findall(r"\sT\s(bgt_[a-z0-9_]+)$", r.stdout, flags=re.M))
        assert set(declared()) <= exported


@pytest.mark.skipif(not os.path.exists(LIB), reason="library not built")
def test_binding_covers_the_header():
    from begputhinning_amd import _lib
    bound = set(_lib.SIGNATURES) | set(_lib.EXTRA_SIGNATURES)
    assert set(declared()) <= bound, sorted(set(declared()) - bound)
    assert _lib.lib().bgt_version() >= 1
    assert isinstance(_lib.lib().bgt_last_error(), bytes)

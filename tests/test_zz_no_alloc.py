"""The allocation-free 2D entry point (bgt_thin2d_device_noalloc) on the
MI355X, through the C-ABI: its outputs and iteration counts equal those of
bgt_thin2d_device (and the oracle's) on two calls in a row, the second with a
smaller image, so the second call runs over what the first left in the
per-GPU array; pitched buffers; an image beyond the array's size is
rejected; device free memory does not move across calls."""
import ctypes

import numpy as np
import pytest

import _oracle as O

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


@pytest.fixture(scope="module", autouse=True)
def _gpu():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")


def _call(name, img, out):
    from begputhinning_amd._lib import check, lib
    it = ctypes.c_int(-1)
    s = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    check(getattr(lib(), name)(ctypes.c_void_p(img.data_ptr()), img.stride(0), img.shape[1],
                               img.shape[0], ctypes.c_void_p(out.data_ptr()), out.stride(0), s,
                               ctypes.byref(it)))
    return it.value


def test_noalloc_equals_device_path_two_calls():
    from begputhinning_amd import synth
    imgs = [synth.polylines(1500, 1200, 91), synth.tree2d(640, 480, 92), synth.polylines(33, 17, 93)]
    for k, a in enumerate(imgs):
        print(f"[no_alloc] call {k}: {a.shape}", flush=True)
        t = torch.from_numpy(a).cuda()
        ref_out = torch.full_like(t, 9)
        got = torch.full_like(t, 7)
        rit = _call("bgt_thin2d_device", t, ref_out)
        git = _call("bgt_thin2d_device_noalloc", t, got)
        torch.cuda.synchronize()
        assert git == rit, k
        assert torch.equal(got, ref_out), k
        oref, oit = O.thin2d(a)
        assert oit == git and np.array_equal(got.cpu().numpy(), oref), k


def test_noalloc_pitched_views():
    from begputhinning_amd import synth
    a = synth.polylines(700, 300, 94)
    base = torch.zeros((300, 1000), dtype=torch.uint8, device="cuda")
    base[:, 100:800] = torch.from_numpy(a).cuda()
    src = base[:, 100:800]
    obase = torch.full((300, 900), 5, dtype=torch.uint8, device="cuda")
    dst = obase[:, 50:750]
    it = _call("bgt_thin2d_device_noalloc", src, dst)
    torch.cuda.synchronize()
    ref, rit = O.thin2d(a)
    assert it == rit and np.array_equal(dst.cpu().numpy(), ref)
    # bytes around the output view are untouched
    assert bool((obase[:, :50] == 5).all()) and bool((obase[:, 750:] == 5).all())


def test_noalloc_rejects_too_large_and_keeps_memory_flat():
    from begputhinning_amd import BGTError, synth
    from begputhinning_amd._lib import lib
    cap = lib().bgt_thin2d_noalloc_max_bytes()
    assert cap == 16 << 20
    # 4096 wide: 64 words per row, 1 KiB of state per row over both buffers
    big = torch.zeros((cap // 1024 + 64, 4096), dtype=torch.uint8, device="cuda")
    with pytest.raises(BGTError) as e:
        _call("bgt_thin2d_device_noalloc", big, torch.empty_like(big))
    assert e.value.code == 4  # BGT_ERR_UNSUPPORTED
    del big
    a = torch.from_numpy(synth.polylines(800, 600, 95)).cuda()
    o = torch.empty_like(a)
    _call("bgt_thin2d_device_noalloc", a, o)
    torch.cuda.synchronize()
    free0 = torch.cuda.mem_get_info()[0]
    for _ in range(5):
        _call("bgt_thin2d_device_noalloc", a, o)
    torch.cuda.synchronize()
    assert torch.cuda.mem_get_info()[0] == free0

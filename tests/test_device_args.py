"""Argument checks of the device-resident Python entry points: every bad
(input, out) pair raises before any kernel runs.  Host tensors are checked
without a GPU; shape, dtype and layout mismatches of GPU tensors in the
`gpu`-marked tests."""
import pytest

torch = pytest.importorskip("torch")

gpu = pytest.mark.gpu


def _entry_points():
    import begputhinning_amd as B
    from begputhinning_amd.batch import ThinningBatch
    from begputhinning_amd.serve import ThinningServer

    def batch2(i, o):
        b = ThinningBatch()
        try:
            b.add2d(i, o)
        finally:
            b.close()

    def batch3(i, o):
        b = ThinningBatch()
        try:
            b.add3d(i, o)
        finally:
            b.close()

    def serve(i, o):
        # the check runs before the server is touched
        ThinningServer.submit(None, i, o)

    return {"thin2d": (2, lambda i, o: B.thin2d_device(i, out=o)),
            "thin3d": (3, lambda i, o: B.thin3d_device(i, out=o)),
            "batch2d": (2, batch2), "batch3d": (3, batch3), "serve2d": (2, serve),
            "serve3d": (3, serve)}


@pytest.mark.parametrize("name", ["thin2d", "thin3d", "batch2d", "batch3d", "serve2d", "serve3d"])
def test_host_tensors_rejected(name):
    nd, f = _entry_points()[name]
    shape = (5, 7) if nd == 2 else (3, 5, 7)
    t = torch.zeros(shape, dtype=torch.uint8)
    with pytest.raises(TypeError):
        f(t, t.clone())


@gpu
@pytest.mark.parametrize("name", ["thin2d", "thin3d", "batch2d", "batch3d", "serve2d", "serve3d"])
def test_bad_out_rejected(name):
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    nd, f = _entry_points()[name]
    shape = (6, 9) if nd == 2 else (4, 6, 9)
    smaller = (6, 8) if nd == 2 else (4, 6, 8)
    t = torch.zeros(shape, dtype=torch.uint8, device="cuda")
    cases = [
        (TypeError, torch.zeros(shape, dtype=torch.uint8)),                  # host out
        (TypeError, torch.zeros(shape, dtype=torch.int32, device="cuda")),   # other dtype
        (ValueError, torch.zeros(smaller, dtype=torch.uint8, device="cuda")),  # smaller
        (ValueError, torch.zeros(shape[::-1], dtype=torch.uint8, device="cuda").transpose(0, -1)),
    ]
    for exc, o in cases:
        with pytest.raises(exc):
            f(t, o)
    with pytest.raises(TypeError):  # a non-uint8 input
        f(torch.zeros(shape, dtype=torch.float32, device="cuda"), torch.zeros_like(t))

"""ssg_frame_sharpness_noalloc: the allocation-free form of ssg_frame_sharpness, called
through the C ABI, equals the allocating entry point bit for bit on two calls in a row
(the second with fewer, smaller frames, so stale scratch contents would show), and sizes
beyond its scratch limits are refused before any GPU work."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _weights(lib, fn, frames):
    from serstacker_amd import _lib

    views = _lib.images([_lib.device_image(f) for f in frames])
    out = np.full(len(frames), np.nan, np.float32)
    rc = getattr(lib, fn)(len(frames), views, out.ctypes.data_as(C.POINTER(C.c_float)))
    assert rc == 0, lib.ssg_last_error()
    return out


def test_noalloc_matches_allocating_entry_point():
    import torch

    from serstacker_amd import _lib

    lib = _lib.load()
    g = torch.Generator(device="cpu").manual_seed(7)
    big = torch.rand((24, 1080, 1920), generator=g).cuda()
    small = torch.rand((5, 241, 333), generator=g).cuda() * 3.0
    for frames in ([big[i] for i in range(24)], [small[i] for i in range(5)]):
        a = _weights(lib, "ssg_frame_sharpness", frames)
        b = _weights(lib, "ssg_frame_sharpness_noalloc", frames)
        assert np.isfinite(a).all()
        assert np.array_equal(a.view(np.uint32), b.view(np.uint32)), (a, b)
    # a strided view (row step > width) as well
    v = big[0][:, 100:700]
    a = _weights(lib, "ssg_frame_sharpness", [v, v])
    b = _weights(lib, "ssg_frame_sharpness_noalloc", [v, v])
    assert np.array_equal(a, b)


def test_noalloc_refuses_sizes_beyond_its_scratch():
    import torch

    from serstacker_amd import _lib

    lib = _lib.load()
    f = torch.zeros((4000, 16), device="cuda")  # ceil(4000 / 16) = 250 partials per frame
    views = _lib.images([_lib.device_image(f)] * 1100)  # 24576 + 8 * 1100 * 250 > 2 MiB
    out = np.zeros(1100, np.float32)
    rc = lib.ssg_frame_sharpness_noalloc(1100, views, out.ctypes.data_as(C.POINTER(C.c_float)))
    assert rc == -1 and b"scratch" in lib.ssg_last_error()
    views = _lib.images([_lib.device_image(f[:16])] * 2049)
    out = np.zeros(2049, np.float32)
    rc = lib.ssg_frame_sharpness_noalloc(2049, views, out.ctypes.data_as(C.POINTER(C.c_float)))
    assert rc == -1

"""GPU master-frame quality: compute_local_variance_map on 8/16-bit frames
(c_local_variance_sharpness_measure.cc:193-249, integer pyrDown and depthScale = 20 / maxval)
and select_master_frame's per-frame metric with average_bayer_planes for Bayer frames
(c_image_stacking_pipeline_base.cc:350-359), against the oracle."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

# Q: float64 sums of g^4 in two orders; map: float(0.05 Q) may straddle one rounding
Q_REL = 1e-12
MAP_RTOL = 3e-7


@pytest.fixture(scope="module")
def S():
    import torch  # noqa: F401

    import serstacker_amd as S

    S._lib.load()
    return S


def gpu(a):
    import torch

    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def textured(r, h, w, dtype, frac=0.8):
    """A smooth random field with noise, spanning frac of the dtype's range (or [0, frac])."""
    f = r.normal(size=(h // 4 + 2, w // 4 + 2))
    f = np.kron(f, np.ones((4, 4)))[:h, :w]
    f = (f - f.min()) / max(np.ptp(f), 1e-9)
    f = np.clip(f + r.normal(0, 0.03, f.shape), 0, 1) * frac
    if dtype == np.float32:
        return f.astype(np.float32)
    return np.round(f * np.iinfo(dtype).max).astype(dtype)


@pytest.mark.parametrize("dtype", [np.uint8, np.uint16])
@pytest.mark.parametrize("shape", [(161, 113), (320, 240)])
@pytest.mark.parametrize("kradius,dscale", [(1, 1), (1, 0), (2, 3)])
def test_local_variance_map_integer_depths(S, oracle32, dtype, shape, kradius, dscale):
    from oracle import lvm

    w, h = shape
    r = np.random.default_rng(w + kradius + 5 * dscale + (dtype == np.uint16))
    img = textured(r, h, w, dtype)
    om, oq = lvm.local_variance_map(oracle32, img, kradius, dscale)
    assert oq > 0
    for src in (img, gpu(img)):
        m, q = S.local_variance_map(src, kradius, dscale)
        assert abs(q - oq) <= Q_REL * abs(oq)
        np.testing.assert_allclose(m, om, rtol=MAP_RTOL, atol=0)


def test_local_variance_map_integer_flat(S):
    m, q = S.local_variance_map(np.full((40, 48), 1234, np.uint16), 1, 1)
    assert q == 0 and not m.any()


@pytest.mark.parametrize("colorid,dtype", [(8, np.uint16), (11, np.uint8), (9, np.float32), (0, np.uint16),
                                           (0, np.uint8), (0, np.float32)])
@pytest.mark.parametrize("device", [False, True])
def test_master_frame_quality(S, oracle32, colorid, dtype, device):
    from oracle import master

    r = np.random.default_rng(colorid * 3 + np.dtype(dtype).itemsize + device)
    h, w = 192, 256
    frames = []
    for i in range(5):
        f = textured(r, h, w, dtype, frac=0.5 + 0.08 * i)
        frames.append(f)
    q = S.master_frame_quality([gpu(f) for f in frames] if device else frames, colorid, 1, 1)
    want = [master.frame_quality(oracle32, f, colorid, 1, 1) for f in frames]
    for a, b in zip(q, want):
        assert b > 0 and abs(a - b) <= Q_REL * abs(b)
    start = 37
    assert S.select_master_frame(q, start) == master.select(want, start)

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


def test_create_reference_frame_chain(S, oracle32):
    """Master-frame generation as create_reference_frame runs it (c_image_stacking_pipeline.cc:
    1112-1312): the best-of scan on raw 16-bit frames picks the master frame, the stacking window
    around it is registered against it and averaged, and linear_interpolation_inpaint fills the
    stack's holes (:1277-1279).  Against the oracle chain fed with the GPU's transforms."""
    import torch

    from oracle import finish, master
    from tests import synth_np

    w, h, n = 192, 160, 12
    ref, frames, _ = synth_np.frames(n, w, h, radius=0.38 * min(w, h), seed=91, shift=5.0, rot_deg=0.5,
                                     scale_dev=0.003, shear=0.002, blur=(0.6, 2.6), noise=0.002)
    frames = [np.clip(f, 0, 1).astype(np.float32) for f in frames]
    raw = [np.round(f * 65535).astype(np.uint16) for f in frames]
    # select_master_frame: best of the scan window
    start, end = S.master_scan_window(n, 8)
    assert (start, end) == master.scan_window(n, 8)
    q = S.master_frame_quality([gpu(r) for r in raw[start:end]], 0, 1, 1)
    want_q = [master.frame_quality(oracle32, r, 0, 1, 1) for r in raw[start:end]]
    np.testing.assert_allclose(q, want_q, rtol=Q_REL, atol=0)
    mpos = S.select_master_frame(q, start)
    assert mpos == master.select(want_q, start)
    # the stacking window around the master frame, registered against it and averaged
    sp, ep = S.master_stack_window(n, 0, mpos, 6)
    assert (sp, ep) == master.stack_window(n, 0, mpos, 6)
    reg = S.c_frame_registration(S.c_image_registration_options(motion_type=3, scale=1.0, eps=0.02, ecch_max_level=3,
                                                                max_iterations=50))
    assert reg.setup_reference_frame(gpu(frames[mpos]))
    win = [gpu(frames[i]) for i in range(sp, ep)]
    st = reg.register_batch(win)
    assert all(s.ok for s in st)
    acc = S.c_weigthed_average(w, h, 1)
    acc.add_registered_batch(win, st, 3)
    avg, mask = acc.compute()
    A = np.zeros((h, w), np.float32)
    W = np.zeros((h, w), np.float32)
    for k, i in enumerate(range(sp, ep)):
        rmap = oracle32.create_map(oracle32.xform(3, list(st[k].params)), w, h)
        d, m = oracle32.base_remap(frames[i], rmap, interp=1, border=4)
        wr, _ = oracle32.base_remap(np.ones((h, w), np.float32), rmap, interp=1, border=0, want_mask=False)
        oracle32.weighted_average_update(d, A, W, weights=(wr * (m.astype(np.float32) / 255)).astype(np.float32))
    want_mask = np.where(W > 0, 255, 0).astype(np.uint8)
    np.testing.assert_array_equal(mask, want_mask)
    # linear_interpolation_inpaint(reference_frame, reference_mask); the master frame covers its
    # own grid, so holes are punched into the mask (as a masked master region would leave them)
    holes = mask.copy()
    holes[40:61, 50:91] = 0
    holes[::17, ::13] = 0
    got = S.inpaint_linear(avg, holes)
    want = finish.inpaint_linear(A, holes)
    assert np.linalg.norm(got - want) / np.linalg.norm(want) <= 1e-5
    # the stack lies on the master frame's grid
    inner = (slice(12, -12), slice(12, -12))
    assert np.abs(got - frames[mpos])[inner].mean() < 0.05

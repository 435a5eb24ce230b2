"""GPU local-variance sharpness map (compute_local_variance_map,
c_local_variance_sharpness_measure.cc:193-249) against its oracle, and the
pipeline's weighted stacking with per-pixel weight maps: weights computed on the
unaligned frames, remapped with BORDER_CONSTANT, gated by the base_remap mask
(c_image_stacking_pipeline.cc:1466-1472, 1653-1660, 1704-1712)."""
import numpy as np
import pytest

from tests import synth_np

pytestmark = pytest.mark.gpu

STACK_REL_L2 = 1e-5


@pytest.fixture(scope="module")
def ssg():
    import torch  # noqa: F401

    import serstacker_amd as S

    S._lib.load()
    return S


def gpu(a):
    import torch

    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def rel_l2(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


@pytest.mark.parametrize("shape", [(161, 113), (320, 240)])
@pytest.mark.parametrize("kradius,dscale", [(1, 1), (1, 0), (2, 2)])
@pytest.mark.parametrize("cn", [1, 3])
def test_local_variance_map_matches_oracle(ssg, oracle32, shape, kradius, dscale, cn):
    from oracle import lvm

    S = ssg
    w, h = shape
    r = np.random.default_rng(w + 7 * kradius + dscale + cn)
    img = r.uniform(0, 1, (h, w) if cn == 1 else (h, w, 3)).astype(np.float32)
    m, q = S.local_variance_map(gpu(img), kradius, dscale)
    om, oq = lvm.local_variance_map(oracle32, img, kradius, dscale)
    assert abs(q - oq) <= 1e-12 * abs(oq)
    # Q enters the map as float(0.05 Q): equal unless the two sum orders straddle a rounding
    np.testing.assert_allclose(m, om, rtol=3e-7, atol=0)


def test_stack_with_local_variance_weight_maps(ssg, oracle32, oracle64):
    S = ssg
    w, h, n = 320, 240, 6
    ref, frames, _ = synth_np.frames(n, w, h, radius=0.38 * min(w, h), seed=31, shift=4.0, rot_deg=0.5,
                                     scale_dev=0.003, shear=0.002, blur=(0.8, 2.0), noise=0.002)
    reg = S.c_frame_registration(S.c_image_registration_options(motion_type=3, scale=1.0, eps=0.02, ecch_max_level=3,
                                                                max_iterations=50))
    assert reg.setup_reference_frame(ref)
    dev = [gpu(f) for f in frames]
    st = reg.register_batch(dev)
    wmaps = [gpu(S.local_variance_map(d, 1, 1)[0]) for d in dev]
    acc = S.c_weigthed_average(w, h, 1)
    acc.add_registered_batch(dev, st, 3, weight_maps=wmaps)
    avg, mask = acc.compute()
    wsum = acc.get_acc_counters()
    S64 = np.zeros((h, w))
    W64 = np.zeros((h, w))
    A32 = np.zeros((h, w), np.float32)
    W32 = np.zeros((h, w), np.float32)
    cnt = np.zeros((h, w), np.int64)
    for i, f in enumerate(frames):
        wm = wmaps[i].cpu().numpy()
        t = oracle32.xform(3, list(st[i].params))
        rmap = oracle32.create_map(t, w, h)
        d, m = oracle32.base_remap(f, rmap, interp=1, border=4)
        wr, _ = oracle32.base_remap(wm, rmap, interp=1, border=0, want_mask=False)
        wts = np.where(m > 0, wr, np.float32(0)).astype(np.float32)
        oracle32.weighted_average_update(d, A32, W32, weights=wts)
        S64 += wts.astype(np.float64) * d
        W64 += wts
        cnt += m > 0
    np.testing.assert_array_equal(acc.counts(), cnt)
    ref64 = np.where(W64 > 0, S64 / np.maximum(W64, 1e-300), 0)
    assert rel_l2(avg, ref64) <= STACK_REL_L2
    assert rel_l2(avg, A32) <= STACK_REL_L2
    assert rel_l2(wsum, W64) <= 1e-6


@pytest.mark.parametrize("shape", [(161, 113), (320, 240), (1920, 1080)])
def test_frame_sharpness_matches_oracle(ssg, oracle32, shape):
    """The C1 per-frame weight, mean squared 3x3 laplacian (c_image_stacking_pipeline.cc:1466-1467):
    per-pixel L in float in the oracle's order, summed in double (in another order), returned as
    float."""
    S = ssg
    w, h = shape
    r = np.random.default_rng(w)
    frames = [r.random((h, w)).astype(np.float32) for _ in range(3)]
    got = S.frame_sharpness([gpu(f) for f in frames])
    want = np.array([oracle32.sharpness(f) for f in frames], np.float64)
    np.testing.assert_allclose(got, want, rtol=2e-7, atol=0)

"""GPU upscale of registration remaps (upscale_remap, c_image_stacking_pipeline.cc:1869-1905)
against the oracle, bit for bit, and stacking at x1.5 after align (the frame loop's
frame_upscale_after_align step, :1633-1641): registered remaps upscaled, the frames warped
onto the larger grid and averaged."""
import numpy as np
import pytest

from oracle import finish as F
from tests import synth_np

pytestmark = pytest.mark.gpu

STACK_REL_L2 = 1e-5
OPTIONS = [F.UPSCALE_X15, F.UPSCALE_X2, F.UPSCALE_X3]


@pytest.fixture(scope="module")
def S():
    import torch  # noqa: F401

    import serstacker_amd as S

    S._lib.load()
    return S


def gpu(a):
    import torch

    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def smooth_map(h, w, seed):
    r = np.random.default_rng(seed)
    yy, xx = np.mgrid[0:h, 0:w].astype(np.float32)
    a = r.uniform(-0.02, 0.02, 4)
    mx = xx * (1 + a[0]) + yy * a[1] + r.uniform(-3, 3) + r.normal(0, 0.05, (h, w))
    my = xx * a[2] + yy * (1 + a[3]) + r.uniform(-3, 3) + r.normal(0, 0.05, (h, w))
    return np.stack([mx, my], -1).astype(np.float32)


@pytest.mark.parametrize("option", OPTIONS)
@pytest.mark.parametrize("shape", [(37, 53), (64, 48), (5, 3)])
def test_upscale_remaps_bitexact(S, oracle32, option, shape):
    import torch

    h, w = shape
    maps = [smooth_map(h, w, 10 * h + w + i) for i in range(5)]
    want = [F.upscale_map(oracle32, m, option) for m in maps]
    # host maps (staged one by one)
    got = S.upscale_remaps(option, maps)
    for g, x in zip(got, want):
        np.testing.assert_array_equal(g, x)
    # device maps, one launch for the batch
    dw, dh = S.upscale_size(option, w, h)
    outs = [torch.empty((dh, dw, 2), dtype=torch.float32, device="cuda") for _ in maps]
    S.upscale_remaps(option, [gpu(m) for m in maps], out=outs)
    for g, x in zip(outs, want):
        np.testing.assert_array_equal(g.cpu().numpy(), x)


def test_upscale_none_copies_and_rejects(S):
    m = smooth_map(9, 11, 1)
    np.testing.assert_array_equal(S.upscale_remaps(F.UPSCALE_NONE, [m])[0], m)
    with pytest.raises(S.SsgError):
        S.upscale_remaps(F.UPSCALE_X15, [m], out=[np.empty((9, 11, 2), np.float32)])
    with pytest.raises(S.SsgError):
        S.upscale_size(7, 10, 10)


@pytest.mark.parametrize("option", [F.UPSCALE_X15, F.UPSCALE_X2])
def test_stack_upscaled_after_align(S, oracle32, option):
    import torch

    w, h, n = 160, 128, 5
    ref, frames, _ = synth_np.frames(n, w, h, radius=0.38 * min(w, h), seed=77, shift=4.0, rot_deg=0.5,
                                     scale_dev=0.003, shear=0.002, blur=(0.8, 2.0), noise=0.002)
    reg = S.c_frame_registration(S.c_image_registration_options(motion_type=3, scale=1.0, eps=0.02, ecch_max_level=2,
                                                                max_iterations=50))
    assert reg.setup_reference_frame(gpu(ref))
    dev = [gpu(f) for f in frames]
    maps = [torch.empty((h, w, 2), dtype=torch.float32, device="cuda") for _ in range(n)]
    st = reg.register_batch_remaps(dev, maps)
    assert all(s.ok for s in st)
    dw, dh = S.upscale_size(option, w, h)
    up = [torch.empty((dh, dw, 2), dtype=torch.float32, device="cuda") for _ in range(n)]
    S.upscale_remaps(option, maps, out=up)
    wts = [0.6 + 0.1 * i for i in range(n)]
    acc = S.c_weigthed_average(dw, dh, 1)
    acc.add_remapped_batch(dev, up, frame_weights=wts)
    avg, mask = acc.compute()
    A = np.zeros((dh, dw), np.float32)
    W = np.zeros((dh, dw), np.float32)
    S64 = np.zeros((dh, dw))
    W64 = np.zeros((dh, dw))
    cnt = np.zeros((dh, dw), np.int64)
    for i in range(n):
        um = F.upscale_map(oracle32, maps[i].cpu().numpy(), option)
        np.testing.assert_array_equal(up[i].cpu().numpy(), um)
        d, dm = oracle32.base_remap(frames[i], um, interp=1, border=4)
        # the scalar weight as the pipeline applies it: a constant image remapped with
        # BORDER_CONSTANT 0, times mask / 255
        wr, _ = oracle32.base_remap(np.full((h, w), np.float32(wts[i])), um, interp=1, border=0, want_mask=False)
        wt = (wr * (dm.astype(np.float32) / 255)).astype(np.float32)
        oracle32.weighted_average_update(d, A, W, weights=wt)
        mb = dm > 0
        S64 += np.where(mb, np.float64(np.float32(wts[i])) * d, 0)
        W64 += np.where(mb, np.float64(np.float32(wts[i])), 0)
        cnt += mb
    np.testing.assert_array_equal(acc.counts(), cnt)
    ref64 = np.where(W64 > 0, S64 / np.maximum(W64, 1e-300), 0)
    assert avg.shape == (dh, dw)
    assert np.linalg.norm(avg - ref64) / np.linalg.norm(ref64) <= STACK_REL_L2
    assert np.linalg.norm(avg - A) / np.linalg.norm(A) <= STACK_REL_L2

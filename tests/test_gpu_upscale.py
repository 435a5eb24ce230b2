"""GPU upscales of stacking at a larger size against the oracle, bit for bit: upscale_remap of the
registration remaps (c_image_stacking_pipeline.cc:1869-1905) and upscale_image of frames, weight
maps and their 8-bit masks (:1949-2000).  Then both stages end to end:
  * after align (the frame loop's frame_upscale_after_align step, :1633-1641): registered remaps
    upscaled, the frames warped onto the larger grid and averaged;
  * before align (:914-918, :1474-1492): the reference frame with its mask and every frame
    upscaled, registered at the upscaled size and stacked.  The upscaled frames are smoother
    than the originals and some pyramid levels end at max_iterations without meeting eps, where
    the FP32 build of the oracle itself ends up to 0.025 px from the FP64 one; each frame's bar
    is therefore 0.01 px or twice that frame's FP32 spread (the farthest the FP32 oracle lands
    from FP64 over the frame and 8 one-ulp perturbed copies), as for the other registrations
    run to convergence."""
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


def binary_mask(h, w, seed, frac=0.06):
    r = np.random.default_rng(seed)
    return np.where(r.random((h, w)) < frac, 0, 255).astype(np.uint8)


@pytest.mark.parametrize("option", OPTIONS)
@pytest.mark.parametrize("shape", [(37, 53), (64, 48), (5, 3), (1, 6)])
@pytest.mark.parametrize("cn", [1, 3])
def test_upscale_images_bitexact(S, oracle32, option, shape, cn):
    """Frames and masks, host and device: bit-exact against the oracle.  Masks are 0 / 255 (as
    the pipeline's) and arbitrary 8-bit values (every path of the fixed-point arithmetic); one
    frame has no mask."""
    import torch

    h, w = shape
    r = np.random.default_rng(100 * h + w + cn + option)
    frames = [r.random((h, w, cn) if cn == 3 else (h, w)).astype(np.float32) for _ in range(5)]
    masks = [binary_mask(h, w, 7 + i) for i in range(3)] + [r.integers(0, 256, (h, w), dtype=np.uint8), None]
    want = [F.upscale_image(oracle32, f, m, option) for f, m in zip(frames, masks)]
    print(f"[upscale images] option {option} shape {shape} cn {cn}", flush=True)
    got, gm = S.upscale_images(option, frames, masks)
    for i, ((wf, wm), g, m) in enumerate(zip(want, got, gm)):
        np.testing.assert_array_equal(g, wf, err_msg=f"frame {i}")
        if wm is None:
            assert m is None
        else:
            np.testing.assert_array_equal(m, wm, err_msg=f"mask {i}")
    dw, dh = S.upscale_size(option, w, h)
    outs = [torch.empty((dh, dw) + frames[0].shape[2:], dtype=torch.float32, device="cuda") for _ in frames]
    mouts = [None if m is None else torch.empty((dh, dw), dtype=torch.uint8, device="cuda") for m in masks]
    S.upscale_images(option, [gpu(f) for f in frames], [None if m is None else gpu(m) for m in masks], out=outs,
                     out_masks=mouts)
    for i, ((wf, wm), g, m) in enumerate(zip(want, outs, mouts)):
        np.testing.assert_array_equal(g.cpu().numpy(), wf, err_msg=f"device frame {i}")
        if m is not None:
            np.testing.assert_array_equal(m.cpu().numpy(), wm, err_msg=f"device mask {i}")


def test_upscale_images_none_and_rejects(S):
    f = np.random.default_rng(2).random((9, 11, 3)).astype(np.float32)
    m = binary_mask(9, 11, 3)
    g, gm = S.upscale_images(F.UPSCALE_NONE, [f], [m])
    assert np.array_equal(g[0], f) and np.array_equal(gm[0], m)
    g, gm = S.upscale_images(F.UPSCALE_X3, [f[..., 0]])
    assert g[0].shape == (27, 33) and gm is None
    with pytest.raises(S.SsgError):  # 2-channel images are remaps, not frames
        S.upscale_images(F.UPSCALE_X15, [np.zeros((9, 11, 2), np.float32)])
    with pytest.raises(S.SsgError):  # a mask of another size
        S.upscale_images(F.UPSCALE_X15, [f], [binary_mask(8, 11, 1)])
    with pytest.raises(S.SsgError):  # an output of the source size
        S.upscale_images(F.UPSCALE_X2, [f], out=[np.empty_like(f)])


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


def test_stack_upscaled_before_align(S, oracle32, oracle64):
    """frame_upscale_before_align at x1.5: the reference frame and its mask, and every frame,
    upscaled on the GPU (bit-exact against the oracle); registration at the upscaled size against
    the FP64 oracle registering the oracle's upscaled frames (the per-frame FP32-spread bar); the
    stack of the upscaled frames at the upscaled size, counts bit-exact and within 1e-5 of the
    FP64 sums."""
    import torch

    from oracle.oracle import reg_options

    option = F.UPSCALE_X15
    w, h, n = 160, 128, 4
    ref, frames, _ = synth_np.frames(n, w, h, radius=0.38 * min(w, h), seed=91, shift=3.0, rot_deg=0.4,
                                     scale_dev=0.002, shear=0.001, blur=(0.8, 2.0), noise=0.002)
    ref_mask = binary_mask(h, w, 5, frac=0.01)
    dw, dh = S.upscale_size(option, w, h)
    (ref_up,), (ref_mask_up,) = S.upscale_images(option, [gpu(ref)], [gpu(ref_mask)],
                                                 out=[torch.empty((dh, dw), dtype=torch.float32, device="cuda")],
                                                 out_masks=[torch.empty((dh, dw), dtype=torch.uint8, device="cuda")])
    o_ref, o_mask = F.upscale_image(oracle32, ref, ref_mask, option)
    np.testing.assert_array_equal(ref_up.cpu().numpy(), o_ref)
    np.testing.assert_array_equal(ref_mask_up.cpu().numpy(), o_mask)
    up = [torch.empty((dh, dw), dtype=torch.float32, device="cuda") for _ in range(n)]
    S.upscale_images(option, [gpu(f) for f in frames], out=up)
    o_frames = [F.upscale_image(oracle32, f, None, option)[0] for f in frames]
    for g, o in zip(up, o_frames):
        np.testing.assert_array_equal(g.cpu().numpy(), o)

    kw = dict(motion_type=3, scale=1.0, eps=0.02, ecch_max_level=3, max_iterations=50)
    reg = S.c_frame_registration(S.c_image_registration_options(**kw))
    assert reg.setup_reference_frame(ref_up, ref_mask_up)
    st = reg.register_batch(up)
    oreg = oracle64.registration(reg_options(**kw))
    oreg.setup_reference_frame(o_ref, o_mask)
    oreg32 = oracle32.registration(reg_options(**kw))
    oreg32.setup_reference_frame(o_ref, o_mask)
    for i in range(n):
        o = oreg.register_frame(o_frames[i])
        assert st[i].ok == o.ok == 1
        d = synth_np.max_displacement(st[i].matrix, o.matrix, dw, dh)
        d32 = synth_np.fp32_spread(oreg32, o, o_frames[i], None, dw, dh, seed=200 + i)
        print(f"  frame {i}: GPU {d:.4f} px from FP64, FP32 oracle spread {d32:.4f} px", flush=True)
        assert d <= synth_np.warp_bar(d32), (i, d, d32)

    acc = S.c_weigthed_average(dw, dh, 1)
    acc.add_registered_batch(up, st, 3)
    avg, _ = acc.compute()
    S64 = np.zeros((dh, dw))
    W64 = np.zeros((dh, dw))
    cnt = np.zeros((dh, dw), np.int64)
    for i in range(n):
        t = oracle32.xform(3, list(st[i].params))
        d, dm = oracle32.base_remap(o_frames[i], oracle32.create_map(t, dw, dh), interp=1, border=4)
        wt = np.float64(dm.astype(np.float32) / np.float32(255))  # the weight of an unweighted frame: mask / 255
        mb = dm > 0
        S64 += np.where(mb, wt * d, 0)
        W64 += np.where(mb, wt, 0)
        cnt += mb
    np.testing.assert_array_equal(acc.counts(), cnt)
    ref64 = np.where(W64 > 0, S64 / np.maximum(W64, 1e-300), 0)
    assert np.linalg.norm(avg - ref64) / np.linalg.norm(ref64) <= STACK_REL_L2

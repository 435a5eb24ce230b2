"""The registration pyramid by column strips (SSG_TUNE_PYR_STRIP 1, the default: the smoothed
level 0 and level 1 in one pass, the other levels by pyrDown strips) against the tiled build
(knob 0: a smoothing pass, then pyrDown per level), through the C ABI. The reference frame's and
every batch frame's pyramid feed the ECC iterations, so identical statuses and parameters bit for
bit -- on sizes whose edge strips and row chunks are partial, at every smoothing width (3..9 taps
take the fused pass, 15 and none the strip pyrDown alone), at ecc.scale 1 (the strips read the
caller's frames, one of them not 16-byte aligned) and 0.5 (they read the half-size copies) --
say that both builds give the same levels."""
import numpy as np
import pytest

from tests import synth_np

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ssg():
    import torch  # noqa: F401

    import serstacker_amd as S
    from serstacker_amd import autotune

    S._lib.load()
    yield S
    autotune.set_knob("pyr_strip", 1)


def register(S, kw, ref, dev, fused):
    from serstacker_amd import autotune

    autotune.set_knob("pyr_strip", fused)
    reg = S.c_frame_registration(S.c_image_registration_options(**kw))
    assert reg.setup_reference_frame(ref)
    return reg.register_batch(dev)


def misaligned(t):
    """A copy of t whose data pointer sits 4 bytes past a 16-byte boundary."""
    import torch

    flat = torch.empty(t.numel() + 1, dtype=t.dtype, device=t.device)
    v = flat[1:].view(t.shape)
    v.copy_(t)
    assert v.data_ptr() % 16 == 4
    return v


@pytest.mark.parametrize("w,h", [(240, 180), (260, 132), (196, 70), (128, 128), (68, 65), (250, 101)])
@pytest.mark.parametrize("sigma", [1.0, 0.5, 0.7, 1.4, 2.4, 0.0])
@pytest.mark.parametrize("scale", [1.0, 0.5])
def test_strip_pyramid_same_registration(ssg, w, h, sigma, scale):
    import torch

    S = ssg
    ref, frames, _ = synth_np.frames(4, w, h, radius=min(w, h) * 0.35, seed=7 + w + h, shift=2.0, rot_deg=0.4,
                                     scale_dev=0.002, blur=(0.8, 1.4), noise=0.002)
    dev = [torch.from_numpy(np.ascontiguousarray(f)).cuda() for f in frames]
    dev.append(misaligned(dev[0]))
    for motion in (0, 3):
        kw = dict(motion_type=motion, scale=scale, eps=0.02, ecch_max_level=3, max_iterations=20,
                  input_smooth_sigma=sigma, reference_smooth_sigma=sigma)
        st0 = register(S, kw, ref, dev, 0)
        for variant in (1,):
            st1 = register(S, kw, ref, dev, variant)
            assert len(st0) == len(st1) == len(dev)
            for a, b in zip(st0, st1):
                assert (a.ok, a.code, a.iterations, a.iterations_translation, a.levels_failed) == \
                    (b.ok, b.code, b.iterations, b.iterations_translation, b.levels_failed)
                assert list(a.params) == list(b.params), variant
            # the misaligned copy of frame 0 takes the gathered loads: the same result as frame 0
            assert list(st1[-1].params) == list(st1[0].params)


def test_strip_pyramid_full_chunks(ssg):
    """A batch large enough for the strips' full 128-row chunks (and the batched half-size step
    over 1080p frames): 96 synthetic 1920x1080 frames, translation-first affine ECC at ecc.scale
    1 and 0.5, strip build against tiled build, bit for bit."""
    import torch

    S = ssg
    n, W, H = 96, 1920, 1080
    p = S.synth_params(width=W, height=H, disk_radius=420.0)
    fr = torch.empty((n, H, W), dtype=torch.float32, device="cuda")
    S.synth_render(p, 0, [fr[i] for i in range(n)])
    ref = torch.empty((H, W), dtype=torch.float32, device="cuda")
    S.synth_render(p, -1, [ref])
    frames = [fr[i] for i in range(n)]
    for scale in (1.0, 0.5):
        kw = dict(motion_type=S.MOTION_AFFINE, scale=scale, eps=0.02, ecch_max_level=4, max_iterations=50)
        st0 = register(S, kw, ref, frames, 0)
        st1 = register(S, kw, ref, frames, 1)
        assert sum(s.ok for s in st1) == n
        for a, b in zip(st0, st1):
            assert (a.ok, a.code, a.iterations, a.iterations_translation, a.rho) == \
                (b.ok, b.code, b.iterations, b.iterations_translation, b.rho)
            assert list(a.params) == list(b.params)

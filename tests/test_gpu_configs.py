"""Parity at the workloads' own shapes and on the registration failure paths.

* the plain-method workload at full size: 64 uint16 640x480 frames, reference = the
  mean of frames 0-7, EUCLIDEAN 3-level ECC (eps 1e-4, <= 50 iterations), bilinear
  base_remap, per-frame weight = mean squared laplacian, weighted average;
* the batched affine workload's geometry at half size: 4-level pyramid, translation
  first, per-frame blur (16 frames of 960x540);
* forward-additive ECC at ecc.scale 0.5 (the reference default), both with a fixed
  iteration budget (identical iteration counts) and to convergence;
* frames that make the Hessian singular (constant, zero) or carry NaN pixels: the
  status, iteration counts and level failures of the FP64 oracle.

Bars: warps <= 0.01 px (max displacement over the frame) vs the FP64 oracle; counts and
accumulated frames bit-exact; stacks <= 1e-5 relative L2.
"""
import numpy as np
import pytest

from tests import synth_np

pytestmark = pytest.mark.gpu

WARP_TOL_PX = 0.01
STACK_REL_L2 = 1e-5


@pytest.fixture(scope="module")
def ssg():
    import torch  # noqa: F401

    import serstacker_amd as S

    S._lib.load()
    return S


def warp_bar(d32):
    """The warp bar of one frame: 0.01 px from the FP64 oracle, or -- on frames where the
    FP32 reference-faithful oracle itself lands farther from FP64 (a level that stops on
    eps < max_eps near the optimum, or runs out of iterations while oscillating, is
    sensitive to FP32 rounding) -- within twice that oracle's own distance."""
    return max(WARP_TOL_PX, 2.0 * d32)


def rel_l2(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def gpu_tensor(a):
    import torch

    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def oracle_stack(oracle32, frames, st, motion, weights, w, h, border=4):
    """The reference pipeline's accumulation of `frames` with the GPU's parameters:
    base_remap (bilinear, `border`), the frame's scalar weight remapped CONSTANT and
    multiplied by mask/255 (c_image_stacking_pipeline.cc:1653-1714), then
    c_weigthed_average.  Returns (counts, FP64 sum-form average, FP32 incremental mean,
    number of frames added)."""
    npar = {0: 2, 1: 3, 2: 4, 3: 6}[motion]
    S_sum = np.zeros((h, w))
    W_sum = np.zeros((h, w))
    cnt = np.zeros((h, w), np.int64)
    A32 = np.zeros((h, w), np.float32)
    W32 = np.zeros((h, w), np.float32)
    added = 0
    for i, f in enumerate(frames):
        if not st[i].ok:
            continue
        added += 1
        t = oracle32.xform(motion, list(st[i].params)[:npar])
        rmap = oracle32.create_map(t, w, h)
        d, m = oracle32.base_remap(f, rmap, interp=1, border=border)
        wimg = np.full((h, w), np.float32(weights[i]))
        wr, _ = oracle32.base_remap(wimg, rmap, interp=1, border=0, want_mask=False)
        wts = (wr * (m.astype(np.float32) / 255)).astype(np.float32)
        oracle32.weighted_average_update(d, A32, W32, weights=wts)
        mb = m > 0
        S_sum += np.where(mb, np.float64(np.float32(weights[i])) * d, 0)
        W_sum += np.where(mb, np.float64(np.float32(weights[i])), 0)
        cnt += mb
    avg = np.where(W_sum > 0, S_sum / np.maximum(W_sum, 1e-300), 0)
    return cnt, avg, A32, added


# ------------------------------------------------------------------------------------------------
def test_plain_method_64x640x480_u16_euclidean(ssg, oracle32, oracle64):
    """64 u16 640x480 disk frames (radius 180, euclidean shift U(-8,8) px, rotation
    U(-0.5,0.5) deg, Poisson 2000 e- + read noise 4 DN), converted by the raw-frame
    conversion (/2^16), registered EUCLIDEAN (3 levels, eps 1e-4, <= 50 iterations,
    ecc.scale 1) against the mean of frames 0-7, stacked with the mean-squared-laplacian
    weight through the bilinear base_remap."""
    import torch

    S = ssg
    W, H, N = 640, 480, 64
    p = S.synth_params(width=W, height=H, disk_radius=180.0, max_shift=8.0, max_rotation_deg=0.5, max_scale_dev=0.0,
                       max_shear=0.0, blur_sigma_min=0.0, blur_sigma_max=0.0, peak_electrons=2000.0, read_noise=4.0,
                       output_scale=1.0, seed=641480)
    raw = torch.empty((N, H, W), dtype=torch.int16, device="cuda")
    raws = [raw[i] for i in range(N)]
    S.synth_render(p, 0, raws)
    f32 = torch.empty((N, H, W), dtype=torch.float32, device="cuda")
    frames_dev = [f32[i] for i in range(N)]
    S.read_input_frames(raws, S._lib.COLORID_MONO, 16, outs=frames_dev)
    raw_h = raw.cpu().numpy().view(np.uint16)
    frames = [raw_h[i].astype(np.float32) / np.float32(65536.0) for i in range(N)]
    got = f32.cpu().numpy()
    assert all(np.array_equal(got[i], frames[i]) for i in range(N)), "u16 -> f32 conversion"
    assert raw_h.max() > 1000, "the photon scale reaches the disk brightness"

    ref = (np.sum(np.stack(frames[:8]).astype(np.float64), axis=0) / 8).astype(np.float32)
    kw = dict(motion_type=S.MOTION_EUCLIDEAN, scale=1.0, eps=1e-4, ecch_max_level=3, max_iterations=50)
    reg = S.c_frame_registration(S.c_image_registration_options(**kw))
    assert reg.setup_reference_frame(gpu_tensor(ref))
    st = reg.register_batch(frames_dev)

    from oracle.oracle import reg_options
    oreg = oracle64.registration(reg_options(**kw))
    oreg.setup_reference_frame(ref)
    worst = 0.0
    for i, f in enumerate(frames):
        o = oreg.register_frame(f)
        assert (st[i].ok, st[i].code) == (o.ok, o.code), i
        assert list(st[i].level_iterations[:3]) == list(o.level_iterations[:3]), i
        assert st[i].iterations_translation == o.iterations_translation, i
        d = synth_np.max_displacement(st[i].matrix, o.matrix, W, H)
        worst = max(worst, d)
        assert d <= WARP_TOL_PX, (i, d)
    print(f"plain-method workload: worst warp difference vs FP64 {worst:.5f} px over {N} frames")

    wts = S.frame_sharpness(frames_dev)
    ow = np.array([oracle64.sharpness(f) for f in frames])
    assert np.abs(wts / ow - 1).max() <= 2e-7

    acc = S.c_weigthed_average(W, H, 1)
    acc.add_registered_batch(frames_dev, st, S.MOTION_EUCLIDEAN, frame_weights=wts)
    avg, mask = acc.compute()
    cnt, ref_avg, A32, added = oracle_stack(oracle32, frames, st, S.MOTION_EUCLIDEAN, wts, W, H)
    assert np.array_equal(acc.counts(), cnt)
    assert acc.accumulated_frames() == added == sum(s.ok for s in st)
    assert np.array_equal(mask > 0, cnt > 0)
    assert rel_l2(avg, ref_avg) <= STACK_REL_L2
    assert rel_l2(avg, A32) <= STACK_REL_L2


# ------------------------------------------------------------------------------------------------
def test_batched_affine_geometry_4_levels(ssg, oracle32, oracle64):
    """The batched affine workload's generator and options (affine shift/rotation/scale/
    shear, per-frame blur U(0.8,2.5), 4-level pyramid, translation first, eps 0.02) at
    half size: 16 frames of 960x540 (levels 960x540, 480x270, 240x134, 120x66)."""
    import torch

    S = ssg
    W, H, N = 960, 540, 16
    p = S.synth_params(width=W, height=H, disk_radius=210.0, seed=9605)
    dev = torch.empty((N, H, W), dtype=torch.float32, device="cuda")
    frames_dev = [dev[i] for i in range(N)]
    truth = S.synth_render(p, 0, frames_dev)
    ref_t = torch.empty((H, W), dtype=torch.float32, device="cuda")
    S.synth_render(p, -1, [ref_t])
    ref = ref_t.cpu().numpy()
    frames = [f for f in dev.cpu().numpy()]
    kw = dict(motion_type=S.MOTION_AFFINE, scale=1.0, eps=0.02, ecch_max_level=4, max_iterations=50)
    reg = S.c_frame_registration(S.c_image_registration_options(**kw))
    assert reg.setup_reference_frame(ref_t)
    st = reg.register_batch(frames_dev)

    from oracle.oracle import reg_options
    oreg = oracle64.registration(reg_options(**kw))
    oreg.setup_reference_frame(ref)
    oreg32 = oracle32.registration(reg_options(**kw))
    oreg32.setup_reference_frame(ref)
    assert oreg.num_levels() == 4
    assert [oreg.level_size(l) for l in range(4)] == [(960, 540), (480, 270), (240, 134), (120, 66)]
    worst, strict = 0.0, 0
    for i, f in enumerate(frames):
        o = oreg.register_frame(f)
        o32 = oreg32.register_frame(f)
        assert (st[i].ok, st[i].code) == (o.ok, o.code) == (1, 0), i
        d = synth_np.max_displacement(st[i].matrix, o.matrix, W, H)
        d32 = synth_np.max_displacement(o32.matrix, o.matrix, W, H)
        print(f"frame {i:2d}: |gpu - fp64| {d:.5f} px, |fp32 oracle - fp64| {d32:.5f} px")
        worst = max(worst, d)
        strict += d <= WARP_TOL_PX
        assert d <= warp_bar(d32), (i, d, d32)
        # sanity of the generator's convention: the true map (reference grid -> frame) is found to
        # within the bias of blurred, noisy frames (the corners lie ~1100 px from the centre)
        assert synth_np.max_displacement(st[i].matrix, truth[i], W, H) < 1.0, i
    print(f"4-level affine: worst warp difference vs FP64 {worst:.5f} px over {N} frames, "
          f"{strict} within {WARP_TOL_PX} px")
    assert strict >= N - 2

    wts = S.frame_sharpness(frames_dev)
    acc = S.c_weigthed_average(W, H, 1)
    acc.add_registered_batch(frames_dev, st, S.MOTION_AFFINE, frame_weights=wts)
    avg, mask = acc.compute()
    cnt, ref_avg, A32, added = oracle_stack(oracle32, frames, st, S.MOTION_AFFINE, wts, W, H)
    assert np.array_equal(acc.counts(), cnt)
    assert acc.accumulated_frames() == added == N
    assert rel_l2(avg, ref_avg) <= STACK_REL_L2
    assert rel_l2(avg, A32) <= STACK_REL_L2


# ------------------------------------------------------------------------------------------------
def _half_frames(motion):
    return synth_np.frames(4, 320, 240, radius=90, seed=31 + motion, shift=4.0, rot_deg=0.5,
                           scale_dev=0.003 if motion == 3 else 0.0, shear=0.002 if motion == 3 else 0.0,
                           blur=(0.8, 1.6), noise=0.002)


@pytest.mark.parametrize("motion", [0, 1, 3])
def test_forward_additive_at_scale_half_one_step(ssg, oracle64, motion):
    """ecc.scale 0.5 (c_frame_registration.h:49): the frames are pyrDown'ed once
    (BORDER_REFLECT101) before ECC, the transform scaled by 0.5 and back by 1/0.5
    (c_frame_registration.cc:617-623,797-802,870-872).  One level, one iteration: the
    half-size images, the solve and the rescale, at the 0.01 px bar for every motion."""
    S = ssg
    ref, frames, _ = _half_frames(motion)
    kw = dict(motion_type=motion, scale=0.5, eps=1e-9, ecch_max_level=1, max_iterations=1,
              ecch_estimate_translation_first=0)
    reg = S.c_frame_registration(S.c_image_registration_options(**kw))
    assert reg.setup_reference_frame(ref)
    st = reg.register_batch([gpu_tensor(f) for f in frames])
    from oracle.oracle import reg_options
    oreg = oracle64.registration(reg_options(**kw))
    oreg.setup_reference_frame(ref)
    for i, f in enumerate(frames):
        o = oreg.register_frame(f)
        assert (st[i].ok, st[i].code) == (o.ok, o.code) == (1, 0)
        assert list(st[i].level_iterations[:1]) == list(o.level_iterations[:1]) == [2]
        d = synth_np.max_displacement(st[i].matrix, o.matrix, 320, 240)
        assert d <= WARP_TOL_PX, (i, d)


@pytest.mark.parametrize("motion", [0, 3])
@pytest.mark.parametrize("budget", ["fixed", "converged"])
def test_forward_additive_at_scale_half(ssg, oracle32, oracle64, motion, budget):
    """The same path over a 2-level pyramid with translation first, for a fixed budget
    (identical iteration counts) and to convergence.  At half size the affine iteration is
    sensitive to FP32 rounding: the FP32 reference-faithful oracle itself lands up to
    ~0.04 px from the FP64 one on these frames, so the GPU is held to 0.01 px or, per
    batch, to 2.5x that oracle's largest distance; translation stays at 0.01 px.  (The
    fixed-scale euclidean iteration at half size is chaotic -- FP32 runs of the oracle end
    up to 0.15 px from FP64 -- and is pinned by the one-step test above.)"""
    S = ssg
    ref, frames, _ = _half_frames(motion)
    if budget == "fixed":
        kw = dict(motion_type=motion, scale=0.5, eps=1e-9, ecch_max_level=2, max_iterations=3)
    else:
        kw = dict(motion_type=motion, scale=0.5, eps=0.02, ecch_max_level=2, max_iterations=50)
    reg = S.c_frame_registration(S.c_image_registration_options(**kw))
    assert reg.setup_reference_frame(ref)
    st = reg.register_batch([gpu_tensor(f) for f in frames])
    from oracle.oracle import reg_options
    oreg = oracle64.registration(reg_options(**kw))
    oreg.setup_reference_frame(ref)
    oreg32 = oracle32.registration(reg_options(**kw))
    oreg32.setup_reference_frame(ref)
    d64, d32 = [], []
    for i, f in enumerate(frames):
        o = oreg.register_frame(f)
        o32 = oreg32.register_frame(f)
        assert (st[i].ok, st[i].code) == (o.ok, o.code) == (1, 0)
        if budget == "fixed":
            assert list(st[i].level_iterations[:2]) == list(o.level_iterations[:2]) == [4, 4]
            assert st[i].iterations_translation == o.iterations_translation
        d64.append(synth_np.max_displacement(st[i].matrix, o.matrix, 320, 240))
        d32.append(synth_np.max_displacement(o32.matrix, o.matrix, 320, 240))
        print(f"frame {i}: |gpu - fp64| {d64[-1]:.5f} px, |fp32 oracle - fp64| {d32[-1]:.5f} px")
    bar = WARP_TOL_PX if motion == 0 else max(WARP_TOL_PX, 2.5 * max(d32))
    assert max(d64) <= bar, (d64, d32)


# ------------------------------------------------------------------------------------------------
@pytest.mark.parametrize("motion", [0, 1, 3])
def test_singular_and_nan_frames_follow_the_oracle(ssg, oracle32, oracle64, motion):
    """A constant or zero frame makes every level's Hessian singular: cv::invert fails
    (ecc2.cc:1331-1335), the level is counted failed after one iteration and the frame
    keeps the identity; its rho is NaN (zero variance), which the `rho < min_rho` gate
    lets through, as in the reference.  A frame with a NaN block drives the parameters to
    NaN; its later iterations sample nothing and fail on the empty Hessian."""
    S = ssg
    w, h = 320, 240
    ref, frames, _ = synth_np.frames(1, w, h, radius=90, seed=5, shift=4.0, rot_deg=0.5)
    const = np.full((h, w), 0.5, np.float32)
    zero = np.zeros((h, w), np.float32)
    nanf = frames[0].copy()
    nanf[100:110, 150:170] = np.nan
    cases = [const, zero, nanf, frames[0]]
    kw = dict(motion_type=motion, scale=1.0, eps=0.02, ecch_max_level=3, max_iterations=20)
    reg = S.c_frame_registration(S.c_image_registration_options(**kw))
    assert reg.setup_reference_frame(ref)
    st = reg.register_batch([gpu_tensor(f) for f in cases])
    from oracle.oracle import reg_options
    oreg = oracle64.registration(reg_options(**kw))
    oreg.setup_reference_frame(ref)
    for i, f in enumerate(cases[:3]):
        o = oreg.register_frame(f)
        g = st[i]
        assert (g.ok, g.code) == (o.ok, o.code), i
        assert list(g.level_iterations[:3]) == list(o.level_iterations[:3]), i
        assert g.iterations_translation == o.iterations_translation, i
        assert g.levels_failed == o.levels_failed, i
        gm, om = np.array(g.matrix[:]), np.array(o.matrix[:])
        assert np.array_equal(np.isnan(gm), np.isnan(om)), (i, gm, om)
        fin = ~np.isnan(om)
        assert np.array_equal(gm[fin], om[fin]), (i, gm, om)
        assert np.isnan(g.rho) == np.isnan(o.rho), i
    # the ordinary frame of the same batch registers as it does alone
    o = oreg.register_frame(cases[3])
    o32 = oracle32.registration(reg_options(**kw))
    o32.setup_reference_frame(ref)
    d32 = synth_np.max_displacement(o32.register_frame(cases[3]).matrix, o.matrix, w, h)
    assert st[3].ok == o.ok == 1
    assert synth_np.max_displacement(st[3].matrix, o.matrix, w, h) <= warp_bar(d32)


# ------------------------------------------------------------------------------------------------
@pytest.mark.parametrize("channel,pick", [(104, None), (103, 0), (101, 2), (1, 1)])
def test_colour_frames_register_on_their_channel(ssg, channel, pick):
    """BGR f32 frames register on ecc_registration_channel (create_ecc_image's extract_channel,
    core/proc/extract_channel.cc:693-697): gray = cv::cvtColor(COLOR_BGR2GRAY) as
    (b * 0.114 + g * 0.587) + r * 0.299 in float, or one channel; the result equals registering
    those gray images, bit for bit."""
    S = ssg
    ref, frames, _ = _register_colour_setup()
    kw = dict(motion_type=3, scale=1.0, eps=0.02, ecch_max_level=2, max_iterations=30, ecc_registration_channel=channel)

    def gray(im):
        if pick is not None:
            return np.ascontiguousarray(im[..., pick])
        return ((im[..., 0] * np.float32(0.114) + im[..., 1] * np.float32(0.587)) +
                im[..., 2] * np.float32(0.299)).astype(np.float32)

    col = S.c_frame_registration(S.c_image_registration_options(**kw))
    assert col.setup_reference_frame(gpu_tensor(ref))
    st_c = col.register_batch([gpu_tensor(f) for f in frames])
    one = S.c_frame_registration(S.c_image_registration_options(**kw))
    assert one.setup_reference_frame(gray(ref))
    st_g = one.register_batch([gpu_tensor(gray(f)) for f in frames])
    for a, b in zip(st_c, st_g):
        assert (a.ok, a.code, a.iterations) == (b.ok, b.code, b.iterations)
        assert list(a.params) == list(b.params)
    # the per-frame form with a host colour frame agrees too
    assert col.register_frame(frames[0])
    assert list(col.status().params) == list(st_g[0].params)


def _register_colour_setup():
    ref, frames, _ = synth_np.frames(3, 200, 150, radius=55, seed=12, shift=3.0, rot_deg=0.5, noise=0.002)
    tint = np.array([0.8, 1.0, 1.15], np.float32)
    return ((ref[..., None] * tint).astype(np.float32),
            [(f[..., None] * tint).astype(np.float32) for f in frames], None)

"""c_frame_accumulation::reinitialize (c_frame_accumulation.h:25, c_frame_accumulation.cc:171-180)
and accumulator checkpoint / resume (ssg_accumulation_state_save / _load), through the C ABI.

* reinitialize restarts a weighted average from a mean image and its weights: compute()
  returns the mean bit for bit where the weight is positive; the next add() follows the
  reference's incremental mean (FP32 oracle, <= 1e-5 relative L2) and the FP64 sum form;
* a stack split at a checkpoint (save, a fresh handle, load, the remaining adds) equals the
  stack made in one handle by the same calls, bit for bit (weighted average with registered
  batches, Bayer average, focus stacking), counts and frame counts included;
* states that do not fit the accumulator are refused; Bayer and focus accumulators refuse
  reinitialize, as the reference's return false.
"""
import numpy as np
import pytest

from tests import synth_np

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ssg():
    import torch  # noqa: F401

    import serstacker_amd as S

    S._lib.load()
    return S


def gpu(a):
    import torch

    a = np.ascontiguousarray(a)
    if a.dtype == np.uint16:
        return torch.from_numpy(a.view(np.int16)).cuda()
    return torch.from_numpy(a).cuda()


def rel_l2(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


@pytest.mark.parametrize("cn", [1, 3])
def test_reinitialize_then_add(ssg, oracle32, cn):
    S = ssg
    r = np.random.default_rng(3 + cn)
    h, w = 45, 67
    shape = (h, w) if cn == 1 else (h, w, cn)
    mean = r.uniform(0, 2, shape).astype(np.float32)
    accw = r.uniform(0.5, 9, (h, w)).astype(np.float32)
    accw[::7, ::5] = 0
    acc = S.c_weigthed_average(w, h, cn)
    assert acc.reinitialize(mean, accw)
    assert acc.accumulated_frames() == 1
    avg, mask = acc.compute()
    pos = accw > 0
    assert np.array_equal(mask, np.where(pos, 255, 0).astype(np.uint8))
    want = np.where(pos[..., None] if cn == 3 else pos, mean, 0)
    assert np.array_equal(avg, want)
    assert np.array_equal(acc.get_acc_counters(), accw)
    assert np.array_equal(acc.counts(), pos.astype(np.int32))

    frame = r.uniform(0, 2, shape).astype(np.float32)
    wts = r.uniform(0.1, 3, (h, w)).astype(np.float32)
    assert acc.add(gpu(frame), gpu(wts))
    avg, _ = acc.compute()
    assert acc.accumulated_frames() == 2
    A = mean.copy()
    W = accw.copy()
    oracle32.weighted_average_update(frame, A, W, weights=wts)
    assert rel_l2(avg[pos], A[pos]) <= 1e-5
    wb = wts[..., None] if cn == 3 else wts
    ab = accw[..., None] if cn == 3 else accw
    s64 = (mean.astype(np.float64) * ab + frame.astype(np.float64) * wb) / (ab.astype(np.float64) + wb)
    assert rel_l2(avg, s64) <= 1e-6


def test_weighted_stack_resumes_from_a_checkpoint(ssg):
    S = ssg
    ref, frames, _ = synth_np.frames(12, 200, 150, radius=55, seed=8, shift=3.0, rot_deg=0.5, noise=0.002)
    h, w = ref.shape
    kw = dict(motion_type=S.MOTION_AFFINE, scale=1.0, eps=0.02, ecch_max_level=2, max_iterations=30)
    reg = S.c_frame_registration(S.c_image_registration_options(**kw))
    assert reg.setup_reference_frame(ref)
    dev = [gpu(f) for f in frames]
    st = reg.register_batch(dev)
    wts = list(S.frame_sharpness(dev))
    first = (S.RegistrationStatus * 7)(*st[:7])
    second = (S.RegistrationStatus * 5)(*st[7:])

    whole = S.c_weigthed_average(w, h, 1)
    whole.add_registered_batch(dev[:7], first, S.MOTION_AFFINE, frame_weights=wts[:7])
    whole.add_registered_batch(dev[7:], second, S.MOTION_AFFINE, frame_weights=wts[7:])

    a = S.c_weigthed_average(w, h, 1)
    a.add_registered_batch(dev[:7], first, S.MOTION_AFFINE, frame_weights=wts[:7])
    state = a.save_state()
    del a
    b = S.c_weigthed_average(w, h, 1)
    b.load_state(state)
    assert b.accumulated_frames() == 7
    b.add_registered_batch(dev[7:], second, S.MOTION_AFFINE, frame_weights=wts[7:])
    for x, y in zip(b.compute(), whole.compute()):
        assert np.array_equal(x, y)
    assert np.array_equal(b.counts(), whole.counts())
    assert b.accumulated_frames() == whole.accumulated_frames() == 12
    # the saved state round-trips unchanged
    assert b.save_state() == whole.save_state()

    with pytest.raises(S.SsgError):
        S.c_weigthed_average(w + 1, h, 1).load_state(state)
    with pytest.raises(S.SsgError):
        S.c_weigthed_average(w, h, 3).load_state(state)
    with pytest.raises(S.SsgError):
        S.c_weigthed_average(w, h, 1).load_state(state[:-4])


def test_bayer_stack_resumes_from_a_checkpoint(ssg):
    S = ssg
    r = np.random.default_rng(12)
    h, w = 64, 96
    raws = [r.integers(0, 4096, (h, w)).astype(np.uint16) for _ in range(6)]
    whole = S.c_bayer_average(w, h, colorid=8, bpp=12)
    part = S.c_bayer_average(w, h, colorid=8, bpp=12)
    for i, f in enumerate(raws):
        assert whole.add(gpu(f))
        if i < 4:
            assert part.add(gpu(f))
    state = part.save_state()
    resumed = S.c_bayer_average(w, h, colorid=8, bpp=12)
    resumed.load_state(state)
    for f in raws[4:]:
        assert resumed.add(gpu(f))
    for x, y in zip(resumed.compute(), whole.compute()):
        assert np.array_equal(x, y)
    assert np.array_equal(resumed.counts(), whole.counts())
    assert resumed.accumulated_frames() == 6
    with pytest.raises(S.SsgError):
        S.c_bayer_average(w, h, colorid=9, bpp=12).load_state(state)
    assert not resumed.reinitialize(np.zeros((h, w), np.float32), np.ones((h, w), np.float32))


@pytest.mark.parametrize("policy", [0, 1, 2])
def test_focus_stack_resumes_from_a_checkpoint(ssg, policy):
    from tests.test_gpu_focus import focus_frames

    S = ssg
    frames = focus_frames(6, 90, 70, 3, seed=4)
    kw = dict(fusing_policy=policy)
    whole = S.c_laplacian_pyramid_focus_stacking(90, 70, 3, **kw)
    part = S.c_laplacian_pyramid_focus_stacking(90, 70, 3, **kw)
    for i, f in enumerate(frames):
        assert whole.add(gpu(f))
        if i < 3:
            assert part.add(gpu(f))
    state = part.save_state()
    resumed = S.c_laplacian_pyramid_focus_stacking(90, 70, 3, **kw)
    resumed.load_state(state)
    for f in frames[3:]:
        assert resumed.add(gpu(f))
    assert np.array_equal(resumed.compute(), whole.compute())
    assert resumed.accumulated_frames() == 6
    assert not resumed.reinitialize(np.zeros((70, 90, 3), np.float32), np.ones((70, 90), np.float32))

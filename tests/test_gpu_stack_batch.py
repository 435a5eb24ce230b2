"""ssg_stack_batch: weights + registration + fused warp/accumulate of a batch with every per-frame
decision on the device, eagerly and as a replayed hipGraph, through the C ABI.  It must give what
the separate calls give (ssg_frame_sharpness, register_batch, add_registered_batch with those
weights): identical statuses and parameters, the same stack bit for bit and the same counts --
on every replay of the graph, with eager work and synchronisations between the replays."""
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


def gpu_tensor(a):
    import torch

    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def separate_calls(S, kw, ref, dev):
    h, w = ref.shape
    reg = S.c_frame_registration(S.c_image_registration_options(**kw))
    assert reg.setup_reference_frame(ref)
    wts = S.frame_sharpness(dev)
    st = reg.register_batch(dev)
    acc = S.c_weigthed_average(w, h, 1)
    acc.add_registered_batch(dev, st, kw["motion_type"], frame_weights=wts)
    return st, acc


def same_status(a, b):
    assert (a.ok, a.code, a.iterations, a.iterations_translation, a.levels_failed) == \
        (b.ok, b.code, b.iterations, b.iterations_translation, b.levels_failed)
    assert list(a.params) == list(b.params)


@pytest.mark.parametrize("motion,scale", [(3, 1.0), (1, 1.0), (0, 1.0), (3, 0.5)])
def test_stack_batch_equals_separate_calls(ssg, motion, scale):
    import torch

    S = ssg
    ref, frames, _ = synth_np.frames(6, 240, 180, radius=70, seed=41 + motion, shift=3.0, rot_deg=0.5,
                                     scale_dev=0.003 if motion == 3 else 0.0, blur=(0.8, 1.6), noise=0.002)
    noise = np.random.default_rng(1).random(ref.shape).astype(np.float32)  # rejected by the rho gate
    frames = frames + [noise]
    dev = [gpu_tensor(f) for f in frames]
    kw = dict(motion_type=motion, scale=scale, eps=0.02, ecch_max_level=3, max_iterations=30)
    st0, acc0 = separate_calls(S, kw, ref, dev)
    want_avg, want_mask = acc0.compute()
    want_cnt = acc0.counts()
    assert st0[6].ok == 0

    reg = S.c_frame_registration(S.c_image_registration_options(**kw))
    assert reg.setup_reference_frame(ref)
    h, w = ref.shape
    acc = S.c_weigthed_average(w, h, 1)
    for rep, graph in enumerate([False, True, True, True]):
        acc.clear()
        st = reg.stack_batch(acc, dev, use_graph=graph)
        for a, b in zip(st, st0):
            same_status(a, b)
        avg, mask = acc.compute()
        assert np.array_equal(avg, want_avg), (rep, graph)
        assert np.array_equal(mask, want_mask)
        assert np.array_equal(acc.counts(), want_cnt)
        assert acc.accumulated_frames() == acc0.accumulated_frames() == 6
        # eager work on the GPU and a synchronisation between the replays
        x = torch.rand((512, 512), device="cuda")
        (x @ x).sum().item()


def test_stack_batch_graph_follows_new_frames_and_accumulates(ssg):
    """Another set of frame buffers re-captures the graph; two batches into one accumulator add up."""
    S = ssg
    ref, frames, _ = synth_np.frames(8, 200, 150, radius=55, seed=9, shift=3.0, rot_deg=0.5, noise=0.002)
    h, w = ref.shape
    kw = dict(motion_type=3, scale=1.0, eps=0.02, ecch_max_level=2, max_iterations=30)
    a_dev = [gpu_tensor(f) for f in frames[:4]]
    b_dev = [gpu_tensor(f) for f in frames[4:]]
    reg0 = S.c_frame_registration(S.c_image_registration_options(**kw))
    assert reg0.setup_reference_frame(ref)
    acc0 = S.c_weigthed_average(w, h, 1)
    for dev in (a_dev, b_dev):
        st = reg0.register_batch(dev)
        acc0.add_registered_batch(dev, st, 3, frame_weights=S.frame_sharpness(dev))
    reg = S.c_frame_registration(S.c_image_registration_options(**kw))
    assert reg.setup_reference_frame(ref)
    acc = S.c_weigthed_average(w, h, 1)
    reg.stack_batch(acc, a_dev)
    reg.stack_batch(acc, b_dev)
    for x, y in zip(acc.compute(), acc0.compute()):
        assert np.array_equal(x, y)
    assert np.array_equal(acc.counts(), acc0.counts())
    assert acc.accumulated_frames() == 8
    # a new reference drops the captured graph (its levels and constants are baked in)
    assert reg.setup_reference_frame(frames[0])
    assert reg0.setup_reference_frame(frames[0])
    acc.clear()
    acc0.clear()
    st = reg.stack_batch(acc, a_dev)
    st0 = reg0.register_batch(a_dev)
    for p, q in zip(st, st0):
        same_status(p, q)


def test_stack_batch_rejects_what_it_does_not_take(ssg):
    S = ssg
    ref, frames, _ = synth_np.frames(1, 120, 90, radius=35, seed=2)
    reg = S.c_frame_registration(S.c_image_registration_options(motion_type=3, scale=1.0, ecch_max_level=2))
    assert reg.setup_reference_frame(ref)
    with pytest.raises(S.SsgError):
        reg.stack_batch(S.c_weigthed_average(121, 90, 1), [gpu_tensor(frames[0])])
    with pytest.raises(S.SsgError):
        reg.stack_batch(S.c_weigthed_average(120, 90, 3), [gpu_tensor(frames[0])])
    with pytest.raises(S.SsgError):
        reg.stack_batch(S.c_weigthed_average(120, 90, 1), [frames[0]])  # host frame

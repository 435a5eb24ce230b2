"""GPU parity of the HIP path against the CPU oracle, through the C ABI.

Bars (from the hot-path contract):
  * per-pixel image operations with the canonical arithmetic: bit-exact;
  * integer counts (per-pixel sample counts, accumulated frames): bit-exact;
  * warps: max displacement over the frame <= 0.01 px vs the FP64 oracle;
  * stacks: relative L2 <= 1e-5 vs the oracle.
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


def rel_l2(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def gpu_tensor(a):
    import torch

    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


# ------------------------------------------------------------------------------------------------
@pytest.mark.parametrize("motion,params", [
    (3, [1.0012, -0.0071, 3.37, 0.0069, 0.9991, -2.61]),
    (0, [2.71875, -1.40625, 0, 0, 0, 0]),
    (1, [1.3, -0.7, 0.0123, 0, 0, 0]),
])
@pytest.mark.parametrize("interp", [1, 2])
@pytest.mark.parametrize("border", [0, 1, 4])
def test_remap_and_mask_bitexact(ssg, oracle32, motion, params, interp, border):
    rng = np.random.default_rng(5)
    h, w = 61, 97
    src = rng.random((h, w)).astype(np.float32)
    S = ssg
    st = S.RegistrationStatus()
    for i, v in enumerate(params):
        st.params[i] = v
    opts = S.c_image_registration_options(motion_type=motion, scale=1.0, ecch_max_level=1)
    reg = S.c_frame_registration(opts)
    assert reg.setup_reference_frame(src)
    dst, dm = reg.custom_remap(src, interpolation=interp, border_mode=border, border_value=0.25, status=st)
    t = oracle32.xform(motion, params[: {0: 2, 1: 3, 3: 6}[motion]])
    rmap = oracle32.create_map(t, w, h)
    ref_dst, ref_mask = oracle32.base_remap(src, rmap, interp=interp, border=border, cval=0.25)
    assert np.array_equal(dst, ref_dst), np.abs(dst - ref_dst).max()
    assert np.array_equal(dm, ref_mask)
    # the remap field the library reports is the oracle's map, bit for bit
    assert np.array_equal(reg.current_remap(st), rmap)


# ------------------------------------------------------------------------------------------------
def _register_setup(w=320, h=240, n=6, seed=11):
    ref, frames, truth = synth_np.frames(n, w, h, radius=0.38 * min(w, h), seed=seed, shift=4.0, rot_deg=0.5,
                                         scale_dev=0.003, shear=0.002, blur=(0.8, 2.0), noise=0.002)
    return ref, frames, truth


def test_register_affine_matches_fp64_oracle(ssg, oracle64):
    S = ssg
    ref, frames, truth = _register_setup()
    h, w = ref.shape
    kw = dict(motion_type=3, scale=1.0, eps=0.02, ecch_max_level=3, max_iterations=50)
    reg = S.c_frame_registration(S.c_image_registration_options(**kw))
    assert reg.setup_reference_frame(ref)
    st = reg.register_batch([gpu_tensor(f) for f in frames])
    from oracle.oracle import reg_options
    oreg = oracle64.registration(reg_options(**kw))
    oreg.setup_reference_frame(ref)
    for i, f in enumerate(frames):
        o = oreg.register_frame(f)
        assert st[i].ok == o.ok == 1
        d = synth_np.max_displacement(st[i].matrix, o.matrix, w, h)
        assert d <= WARP_TOL_PX, (i, d)
        assert abs(st[i].rho - o.rho) < 1e-4
        # the recovered warp is close to the known one (noise limited)
        assert synth_np.max_displacement(st[i].matrix, truth[i], w, h) < 0.2


@pytest.mark.parametrize("motion", [0, 1, 3])
def test_plain_and_batch_are_identical(ssg, motion):
    """The plain per-frame form and the batched form run the same kernels and
    reductions in the same order: identical parameters and iteration counts."""
    S = ssg
    ref, frames, _ = _register_setup(n=3, seed=3)
    kw = dict(motion_type=motion, scale=1.0, eps=0.02, ecch_max_level=3, max_iterations=30)
    reg = S.c_frame_registration(S.c_image_registration_options(**kw))
    assert reg.setup_reference_frame(ref)
    st = reg.register_batch([gpu_tensor(f) for f in frames])
    for i, f in enumerate(frames):
        reg.register_frame(f)
        p = reg.status()
        assert list(p.params) == list(st[i].params)
        assert p.iterations == st[i].iterations and p.iterations_translation == st[i].iterations_translation
        assert p.ok == st[i].ok


def test_register_euclidean_fp64_oracle_and_iteration_rule(ssg, oracle64):
    """EUCLIDEAN (fixed scale) never meets eps (reference eps uses the scale, 1):
    every level runs max_iterations and reports max+1, as in the reference."""
    S = ssg
    ref, frames, _ = synth_np.frames(3, 320, 240, radius=90, seed=21, shift=4.0, rot_deg=0.5, noise=0.002)
    kw = dict(motion_type=1, scale=1.0, eps=1e-4, ecch_max_level=3, max_iterations=12)
    reg = S.c_frame_registration(S.c_image_registration_options(**kw))
    assert reg.setup_reference_frame(ref)
    st = reg.register_batch([gpu_tensor(f) for f in frames])
    from oracle.oracle import reg_options
    oreg = oracle64.registration(reg_options(**kw))
    oreg.setup_reference_frame(ref)
    for i, f in enumerate(frames):
        o = oreg.register_frame(f)
        assert list(st[i].level_iterations[:3]) == [13, 13, 13] == list(o.level_iterations[:3])
        assert synth_np.max_displacement(st[i].matrix, o.matrix, 320, 240) <= WARP_TOL_PX


def test_rho_gate_rejects_unrelated_frame(ssg, oracle64):
    S = ssg
    ref, frames, _ = _register_setup(n=1)
    noise = np.random.default_rng(0).random(ref.shape).astype(np.float32)
    kw = dict(motion_type=3, scale=1.0, eps=0.02, ecch_max_level=3, max_iterations=20)
    reg = S.c_frame_registration(S.c_image_registration_options(**kw))
    assert reg.setup_reference_frame(ref)
    st = reg.register_batch([gpu_tensor(noise), gpu_tensor(frames[0])])
    from oracle.oracle import reg_options
    o = oracle64.registration(reg_options(**kw))
    o.setup_reference_frame(ref)
    on = o.register_frame(noise)
    assert st[0].ok == 0 and st[0].code == on.code == 3
    assert st[1].ok == 1


# ------------------------------------------------------------------------------------------------
def test_accumulate_counts_bitexact_and_stack(ssg, oracle32, oracle64):
    S = ssg
    ref, frames, _ = _register_setup(n=8, seed=17)
    h, w = ref.shape
    kw = dict(motion_type=3, scale=1.0, eps=0.02, ecch_max_level=3, max_iterations=50)
    reg = S.c_frame_registration(S.c_image_registration_options(**kw))
    assert reg.setup_reference_frame(ref)
    dev = [gpu_tensor(f) for f in frames]
    st = reg.register_batch(dev)
    weights = [oracle64.sharpness(f) for f in frames]
    acc = S.c_weigthed_average(w, h, 1)
    acc.add_registered_batch(dev, st, 3, frame_weights=weights)
    avg, mask = acc.compute()
    counts = acc.counts()
    # oracle: base_remap of each frame with the GPU's parameters, then the
    # reference pipeline's weights (constant weight remapped, * mask/255)
    S_sum = np.zeros((h, w))
    W_sum = np.zeros((h, w))
    n_cnt = np.zeros((h, w), np.int64)
    A32 = np.zeros((h, w), np.float32)
    W32 = np.zeros((h, w), np.float32)
    for i, f in enumerate(frames):
        t = oracle32.xform(3, list(st[i].params))
        rmap = oracle32.create_map(t, w, h)
        d, m = oracle32.base_remap(f, rmap, interp=1, border=4)
        wimg = np.full((h, w), np.float32(weights[i]))
        wr, _ = oracle32.base_remap(wimg, rmap, interp=1, border=0, want_mask=False)
        wts = (wr * (m.astype(np.float32) / 255)).astype(np.float32)
        oracle32.weighted_average_update(d, A32, W32, weights=wts)
        mb = m > 0
        S_sum += np.where(mb, np.float64(np.float32(weights[i])) * d, 0)
        W_sum += np.where(mb, np.float64(np.float32(weights[i])), 0)
        n_cnt += mb
    assert np.array_equal(counts, n_cnt)
    assert acc.accumulated_frames() == sum(s.ok for s in st) == len(frames)
    assert np.array_equal(mask > 0, W_sum > 0)
    ref_avg = np.where(W_sum > 0, S_sum / np.maximum(W_sum, 1e-300), 0)
    assert rel_l2(avg, ref_avg) <= STACK_REL_L2
    # and against the reference-faithful FP32 incremental mean
    assert rel_l2(avg, A32) <= STACK_REL_L2


def test_accumulate_identical_frames_is_idempotent(ssg):
    """Stacking k copies of a frame with identity warps returns the frame."""
    S = ssg
    rng = np.random.default_rng(2)
    f = rng.random((128, 200)).astype(np.float32)
    st = (S.RegistrationStatus * 5)()
    for s in st:
        s.ok = 1
        s.params[0], s.params[4] = 1.0, 1.0
    acc = S.c_weigthed_average(200, 128, 1)
    acc.add_registered_batch([gpu_tensor(f)] * 5, st, 3, frame_weights=[0.5, 1, 2, 3, 4])
    avg, mask = acc.compute()
    assert (mask == 255).all()
    assert np.array_equal(acc.counts(), np.full((128, 200), 5))
    assert np.abs(avg - f).max() <= 2e-7


def test_synth_is_deterministic(ssg):
    import torch

    S = ssg
    p = S.synth_params(width=320, height=240, disk_radius=90, seed=99)
    a = [torch.empty((240, 320), dtype=torch.float32, device="cuda") for _ in range(3)]
    b = [torch.empty((240, 320), dtype=torch.float32, device="cuda") for _ in range(3)]
    t1 = S.synth_render(p, 5, a)
    t2 = S.synth_render(p, 5, b)
    assert np.array_equal(t1, t2)
    for x, y in zip(a, b):
        assert torch.equal(x, y)
    # frame 6 rendered alone equals frame 6 of the batch
    c = [torch.empty((240, 320), dtype=torch.float32, device="cuda")]
    S.synth_render(p, 6, c)
    assert torch.equal(c[0], a[1])

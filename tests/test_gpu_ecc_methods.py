"""GPU parity of the Levenberg-Marquardt and inverse-compositional ECC aligners
(ECC_ALIGN_LM, ECC_ALIGN_INVERSE_COMPOSITIONAL, ECC_ALIGN_INVERSE_COMPOSITIONAL_LM;
c_ecclm, c_ecc_inverse_compositional, c_ecclm_inverse_compositional, ecc2.cc:1371-2086)
against the FP64 oracle, through the C ABI.

Two bars, because these aligners stop on rules that end the iteration in a region
whose size is set by the arithmetic, not by the data: the plain inverse-compositional
loop stops when the error stops decreasing (ecc2.cc:1776-1778), the damped variants
accept a step when the error decreases (ecc2.cc:2056-2067) and the fixed-scale
euclidean eps never falls below max(W, H), so near the optimum every decision is a
comparison of two errors that differ in their last bits.  On these scenes the
oracle's FP32 build (the reference's own precision) ends up to 0.09 px from its FP64
build with the same algorithm.
  * Fixed budget (a few iterations per level, eps out of reach): every step is far
    from that region, so the GPU follows the FP64 oracle's trajectory: the same
    iteration counts and level-failure flags, warps within 0.01 px.
  * To convergence: frame by frame, the GPU lies within 0.01 px of the FP64 oracle or, on
    frames where the reference algorithm in FP32 is less determined than that, within
    twice its FP32 spread: the farthest the FP32 build of the oracle (the reference's own
    precision) lands from the FP64 end point, over the frame itself and 8 copies of it
    with every pixel moved by one FP32 ulp (a different rounding of the same input, as a
    different summation order is).  Each frame prints both distances.  Correlation within
    1e-4.
"""
import numpy as np
import pytest

from tests import synth_np

pytestmark = pytest.mark.gpu

WARP_TOL_PX = 0.01
LM, IC, ICLM = 2, 1, 3


@pytest.fixture(scope="module")
def ssg():
    import torch  # noqa: F401

    import serstacker_amd as S

    S._lib.load()
    return S


def gpu_tensor(a):
    import torch

    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def _frames(n=4, w=320, h=240, seed=11, scale=1.0):
    ref, frames, truth = synth_np.frames(n, w, h, radius=0.38 * min(w, h), seed=seed, shift=4.0, rot_deg=0.5,
                                         scale_dev=0.003, shear=0.002, blur=(0.8, 2.0), noise=0.002, scale=scale)
    return ref, frames, truth


warp_bar = synth_np.warp_bar
fp32_spread = synth_np.fp32_spread


def _compare(S, oracle64, kw, ref, frames, masks=None, ref_mask=None, same_counts=False, oracle32=None):
    """GPU batch vs the FP64 oracle frame by frame.  With oracle32 (runs to convergence) each
    frame's bar is warp_bar of its FP32 spread; without it (fixed budgets) the bar is 0.01 px."""
    h, w = ref.shape
    reg = S.c_frame_registration(S.c_image_registration_options(**kw))
    assert reg.setup_reference_frame(ref, ref_mask)
    st = reg.register_batch([gpu_tensor(f) for f in frames], [gpu_tensor(m) for m in masks] if masks else None)
    from oracle.oracle import reg_options
    oreg = oracle64.registration(reg_options(**kw))
    oreg.setup_reference_frame(ref, ref_mask)
    oreg32 = None
    if oracle32 is not None:
        oreg32 = oracle32.registration(reg_options(**kw))
        oreg32.setup_reference_frame(ref, ref_mask)
    for i, f in enumerate(frames):
        o = oreg.register_frame(f, masks[i] if masks else None)
        d = synth_np.max_displacement(st[i].matrix, o.matrix, w, h)
        d32, bar = 0.0, WARP_TOL_PX
        if oreg32 is not None and o.ok:
            d32 = fp32_spread(oreg32, o, f, masks[i] if masks else None, w, h, seed=100 + i)
            bar = warp_bar(d32)
        print(f"  frame {i}: gpu ok {st[i].ok} it {st[i].iterations}/{st[i].iterations_translation} "
              f"failed {st[i].levels_failed}; fp64 ok {o.ok} it {o.iterations}/{o.iterations_translation} "
              f"failed {o.levels_failed}; |gpu-fp64| {d:.5f} px, fp32 spread {d32:.5f} px, bar {bar:.5f}",
              flush=True)
        assert st[i].ok == o.ok and st[i].code == o.code
        assert st[i].levels_failed == o.levels_failed
        if same_counts:
            assert list(st[i].level_iterations[:4]) == list(o.level_iterations[:4])
            assert st[i].iterations_translation == o.iterations_translation
        if o.ok:
            assert d <= bar, (i, d, d32)
            assert abs(st[i].rho - o.rho) < 1e-4
    return st


@pytest.mark.parametrize("method", [LM, IC, ICLM])
@pytest.mark.parametrize("motion", [0, 1, 2, 3])
def test_methods_fixed_budget_trajectory(ssg, oracle64, method, motion):
    """One level, no translation-first pass, three iterations from the 4-px start, eps out
    of reach: every step is large, and the GPU takes the FP64 oracle's steps."""
    print(f"[ecc trajectory] method {method} motion {motion}", flush=True)
    ref, frames, _ = _frames(scale=100.0)
    kw = dict(motion_type=motion, scale=1.0, eps=1e-7, ecch_max_level=1, max_iterations=3, ecc_method=method,
              min_rho=-1.0, ecch_estimate_translation_first=0)
    _compare(ssg, oracle64, kw, ref, frames, same_counts=True)


@pytest.mark.parametrize("method", [LM, IC, ICLM])
@pytest.mark.parametrize("motion", [0, 1, 2, 3])
@pytest.mark.parametrize("scale", [1.0, 100.0])
def test_methods_converged_match_oracle(ssg, oracle64, oracle32, method, motion, scale):
    """Pyramid registration (translation-first for the non-translation motions) of blurred
    noisy frames to convergence; scale 1 keeps the reference's [0, 1] range, where
    c_ecclm also stops on |rhs|^2 < 1, scale 100 lets it run to its step criteria."""
    print(f"[ecc methods] method {method} motion {motion} scale {scale}", flush=True)
    ref, frames, _ = _frames(scale=scale)
    kw = dict(motion_type=motion, scale=1.0, eps=0.02, ecch_max_level=3, max_iterations=50, ecc_method=method)
    _compare(ssg, oracle64, kw, ref, frames, oracle32=oracle32)


@pytest.mark.parametrize("method", [LM, IC, ICLM])
def test_methods_half_scale_defaults(ssg, oracle64, oracle32, method):
    """The registration defaults (ecc.scale 0.5, one level, no translation-first pass)."""
    print(f"[ecc methods defaults] method {method}", flush=True)
    ref, frames, _ = _frames(n=3, w=384, h=256, seed=5, scale=100.0)
    kw = dict(motion_type=3, ecc_method=method)
    _compare(ssg, oracle64, kw, ref, frames, oracle32=oracle32)


@pytest.mark.parametrize("method", [LM, IC, ICLM])
@pytest.mark.parametrize("budget", [3, 40])
def test_methods_with_masks(ssg, oracle64, oracle32, method, budget):
    """Current and reference masks: eroded for c_ecclm's reference only, used raw by the
    inverse-compositional aligners (nearest-neighbour mask for the LM variant)."""
    print(f"[ecc methods masks] method {method}", flush=True)
    ref, frames, _ = _frames(n=3, seed=9, scale=100.0)
    h, w = ref.shape
    rng = np.random.default_rng(4)
    ref_mask = np.full((h, w), 255, np.uint8)
    ref_mask[40:90, 30:120] = 0
    masks = []
    for _ in frames:
        m = np.full((h, w), 255, np.uint8)
        y, x = rng.integers(20, h - 60), rng.integers(20, w - 80)
        m[y:y + 40, x:x + 60] = 0
        masks.append(m)
    fixed = budget < 10
    kw = dict(motion_type=3, scale=1.0, eps=1e-7 if fixed else 0.02, ecch_max_level=1 if fixed else 2,
              max_iterations=budget, ecc_method=method, min_rho=-1.0 if fixed else 0.8,
              ecch_estimate_translation_first=0 if fixed else 1)
    _compare(ssg, oracle64, kw, ref, frames, masks=masks, ref_mask=ref_mask, same_counts=fixed,
             oracle32=None if fixed else oracle32)


@pytest.mark.parametrize("tfirst", [0, 1])
@pytest.mark.parametrize("method", [IC, ICLM])
def test_inverse_compositional_j_record(ssg, oracle64, method, tfirst):
    """Euclidean J depends on the parameters of its precompute: per frame when the
    translation-first pass swaps the transform, else fixed by the first frame after
    setup_reference_frame and kept across later batches."""
    print(f"[ecc J record] method {method} translation-first {tfirst}", flush=True)
    ref, frames, _ = _frames(n=4, seed=13, scale=100.0)
    S = ssg
    h, w = ref.shape
    kw = dict(motion_type=2, scale=1.0, eps=1e-7, ecch_max_level=2, max_iterations=2, ecc_method=method,
              ecch_estimate_translation_first=tfirst, min_rho=-1.0)
    reg = S.c_frame_registration(S.c_image_registration_options(**kw))
    assert reg.setup_reference_frame(ref)
    st = list(reg.register_batch([gpu_tensor(f) for f in frames[:2]]))
    st += list(reg.register_batch([gpu_tensor(f) for f in frames[2:]]))
    from oracle.oracle import reg_options
    oreg = oracle64.registration(reg_options(**kw))
    oreg.setup_reference_frame(ref)
    for i, f in enumerate(frames):
        o = oreg.register_frame(f)
        d = synth_np.max_displacement(st[i].matrix, o.matrix, w, h)
        print(f"  frame {i}: it {list(st[i].level_iterations[:2])} vs {list(o.level_iterations[:2])} d {d:.5f}",
              flush=True)
        assert list(st[i].level_iterations[:2]) == list(o.level_iterations[:2])
        assert d <= WARP_TOL_PX


@pytest.mark.parametrize("method", [LM, IC, ICLM])
def test_methods_plain_and_batch_identical(ssg, method):
    S = ssg
    ref, frames, _ = _frames(n=3, seed=3, scale=100.0)
    kw = dict(motion_type=3, scale=1.0, eps=0.02, ecch_max_level=3, max_iterations=30, ecc_method=method)
    reg = S.c_frame_registration(S.c_image_registration_options(**kw))
    assert reg.setup_reference_frame(ref)
    st = reg.register_batch([gpu_tensor(f) for f in frames])
    for i, f in enumerate(frames):
        reg.register_frame(f)
        p = reg.status()
        assert list(p.params) == list(st[i].params)
        assert p.iterations == st[i].iterations


def test_lm_level_failure_skips_the_rescale(ssg, oracle64):
    """c_ecclm returns false when the iteration budget ends a level; c_ecch then keeps the
    transform's scale for the finer level (ecc2.cc:1161-1170) -- same flags, same warp."""
    ref, frames, _ = _frames(n=2, seed=7, scale=100.0)
    kw = dict(motion_type=0, scale=1.0, eps=1e-6, ecch_max_level=2, max_iterations=2, ecc_method=LM)
    st = _compare(ssg, oracle64, kw, ref, frames, same_counts=True)
    assert all(s.levels_failed > 0 for s in st)

"""GPU focus stacking against the FP32 oracle (oracle/focus.py), through the C ABI:
the lpg sharpness map (core/proc/lpg.cc), laplacian pyramids and the fusion
policies of c_laplacian_pyramid_focus_stacking, plus the lpg-weighted
composition.  Same canonical arithmetic on both sides: compared bit for bit."""
import numpy as np
import pytest
from scipy import ndimage

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ssg():
    import torch  # noqa: F401

    import serstacker_amd as S

    S._lib.load()
    return S


def gpu(a):
    import torch

    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def planes(img):
    return [img] if img.ndim == 2 else [np.ascontiguousarray(img[..., c]) for c in range(img.shape[2])]


def focus_frames(n, w, h, cn, seed):
    """A textured scene focused at different depths: frame k is sharp where the
    depth map is near k and blurred elsewhere."""
    r = np.random.default_rng(seed)
    scene = ndimage.gaussian_filter(r.uniform(0, 1, (h, w, cn)), (0.8, 0.8, 0)).astype(np.float32)
    depth = ndimage.gaussian_filter(r.uniform(0, n, (h, w)), 6)
    depth = (depth - depth.min()) / max(np.ptp(depth), 1e-9) * (n - 1)
    out = []
    for k in range(n):
        blurred = ndimage.gaussian_filter(scene, (2.5, 2.5, 0))
        t = np.clip(np.abs(depth - k), 0, 1)[..., None]
        f = (scene * (1 - t) + blurred * t).astype(np.float32)
        out.append(f if cn > 1 else f[..., 0])
    return out


@pytest.mark.parametrize("cn", [1, 3])
@pytest.mark.parametrize("k,p", [(6.0, 2.0), (3.0, 1.0)])
def test_lpg_map_bitexact(ssg, cn, k, p):
    from oracle import focus

    S = ssg
    img = focus_frames(1, 77, 53, cn, seed=5)[0]
    m, sharp = S.lpg(gpu(img), k, p)
    want = focus.lpg(focus.gray(planes(img)), k, p)
    np.testing.assert_array_equal(m, want)
    assert abs(sharp - float(want.astype(np.float64).mean())) <= 1e-9 * abs(sharp)


@pytest.mark.parametrize("cn", [1, 3])
@pytest.mark.parametrize("policy,avgc,kradius,ksigma", [
    (0, 1, 0, 0.0), (0, 1, 2, 1.0), (1, 1, 0, 0.0), (1, 0, 0, 0.0), (1, 1, 2, 1.0), (2, 1, 0, 0.0)])
def test_focus_stack_bitexact(ssg, oracle32, cn, policy, avgc, kradius, ksigma):
    from oracle import focus

    S = ssg
    w, h, n = 96, 80, 4
    frames = focus_frames(n, w, h, cn, seed=7 + policy)
    print(f"[focus] cn {cn} policy {policy} avgc {avgc} kradius {kradius} ksigma {ksigma}", flush=True)
    acc = S.c_laplacian_pyramid_focus_stacking(w, h, cn, fusing_policy=policy, avgchannel=avgc, kradius=kradius,
                                               ksigma=ksigma, lpg_sigma=2.0)
    fs = focus.FocusStack(oracle32, policy, avgchannel=bool(avgc), kradius=kradius, ksigma=ksigma, lpg_sigma=2.0)
    for f in frames:
        assert acc.add(gpu(f))
        fs.add(planes(f))
    assert acc.accumulated_frames() == n
    got = acc.compute()
    want = fs.compute()
    want = want[0] if cn == 1 else np.stack(want, -1)
    np.testing.assert_array_equal(got, want)


def test_focus_lpg_weights_large_rgb(ssg, oracle32):
    """The focus configuration's shape class: RGB 1024x768, lpg(k=6, p=2) smoothed with
    sigma 5 (41 taps), 6 levels."""
    from oracle import focus

    S = ssg
    w, h, n = 1024, 768, 3
    frames = focus_frames(n, w, h, 3, seed=11)
    acc = S.c_laplacian_pyramid_focus_stacking(w, h, 3, fusing_policy=S._lib.FOCUS_LPG_WEIGHTS, lpg_k=6.0, lpg_p=2.0,
                                               lpg_sigma=5.0, max_levels=6)
    fs = focus.FocusStack(oracle32, focus.LPG_WEIGHTS, lpg_k=6.0, lpg_p=2.0, lpg_sigma=5.0, max_levels=6)
    for f in frames:
        assert acc.add(gpu(f))
        fs.add(planes(f))
    np.testing.assert_array_equal(acc.compute(), np.stack(fs.compute(), -1))


@pytest.mark.parametrize("cn", [1, 3])
@pytest.mark.parametrize("policy,avgc,kradius,ksigma", [
    (0, 1, 0, 0.0), (0, 1, 2, 1.0), (1, 1, 0, 0.0), (1, 0, 0, 0.0), (2, 1, 0, 0.0)])
def test_focus_add_frames_equals_single_adds(ssg, cn, policy, avgc, kradius, ksigma):
    """add_frames (pyramids of a batch built together, levels fused frame after frame in one
    launch) gives the stack of the same frames added one by one, bit for bit: a first batch,
    a single add between, and a second batch that continues the stack."""
    S = ssg
    w, h = 130, 98
    frames = focus_frames(7, w, h, cn, seed=23 + policy)
    print(f"[focus batch] cn {cn} policy {policy} avgc {avgc} kradius {kradius}", flush=True)
    kw = dict(fusing_policy=policy, avgchannel=avgc, kradius=kradius, ksigma=ksigma, lpg_sigma=1.5)
    one = S.c_laplacian_pyramid_focus_stacking(w, h, cn, **kw)
    for f in frames:
        assert one.add(gpu(f))
    bat = S.c_laplacian_pyramid_focus_stacking(w, h, cn, **kw)
    bat.add_frames([gpu(f) for f in frames[:3]])
    assert bat.add(gpu(frames[3]))
    bat.add_frames(frames[4:])  # host frames
    assert bat.accumulated_frames() == 7
    np.testing.assert_array_equal(bat.compute(), one.compute())


def test_focus_add_frames_errors(ssg):
    S = ssg
    acc = S.c_laplacian_pyramid_focus_stacking(64, 48, 1)
    with pytest.raises(RuntimeError):
        acc.add_frames([np.zeros((48, 64, 3), np.float32)])
    with pytest.raises(RuntimeError):
        acc.add_frames([np.zeros((40, 64), np.float32)])
    acc.add_frames([])
    assert acc.accumulated_frames() == 0

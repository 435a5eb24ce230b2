"""GPU dense ECC flow (c_eccflow, ecc2.cc:2103-2865) against its FP32 oracle
(oracle/eccflow.py) through the C ABI: the reference pyramid planes (level images,
Ix, Iy and the D cells) bit for bit, and the refined remaps of whole compute() calls
for every downscale method, with and without an initial remap and masks, single
and batched."""
import numpy as np
import pytest

from tests import synth_np

pytestmark = pytest.mark.gpu

# The GPU runs the oracle's canonical arithmetic: planes and refined remaps are
# bit-exact (a zero bound, in px).
FLOW_ATOL = 0.0


@pytest.fixture(scope="module")
def ssg():
    import torch  # noqa: F401

    import serstacker_amd as S

    S._lib.load()
    return S


def gpu(a):
    import torch

    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def scene(w, h, seed):
    return synth_np.blob_scene(w, h, seed, n=int(w * h / 50), smin=3.0, smax=6.0)


def frames_pair(w, h, seed, shift=(1.3, -0.8), amp=1.0):
    S = scene(w, h, seed)
    yy, xx = np.mgrid[0:h, 0:w].astype(np.float64)
    r = np.random.default_rng(seed)
    p = r.uniform(0, 2 * np.pi, 2)
    ex = shift[0] + amp * np.sin(2 * np.pi * xx / (1.3 * w) + p[0])
    ey = shift[1] + amp * np.cos(2 * np.pi * yy / (1.1 * h) + p[1])
    ref = S([1, 0, 0, 0, 1, 0]).astype(np.float32)
    cur = S.at(xx - ex, yy - ey).astype(np.float32)
    return ref, cur


OPTS = {
    "pyramid": dict(downscale=2, support_scale=3, max_iterations=3, min_image_size=-1, max_pyramid_level=-1),
    "pyramid_ss2": dict(downscale=2, support_scale=2, max_iterations=2, min_image_size=-1, max_pyramid_level=3),
    "pyramid_ss4": dict(downscale=2, support_scale=4, max_iterations=2, min_image_size=-1, max_pyramid_level=2),
    "pyramid_ss5": dict(downscale=2, support_scale=5, max_iterations=2, min_image_size=-1, max_pyramid_level=1),
    "recursive": dict(downscale=0, scale_factor=0.75, support_scale=3, max_iterations=2, min_image_size=8),
    "full": dict(downscale=1, scale_factor=0.5, support_scale=2, max_iterations=2, min_image_size=4),
}


def make_pair(ssg, oracle32, name):
    from oracle import eccflow as EF

    o = OPTS[name]
    g = ssg.c_eccflow(**o)
    f = EF.EccFlow(oracle32, **o)
    return g, f


@pytest.mark.parametrize("name", list(OPTS))
@pytest.mark.parametrize("shape", [(160, 128), (173, 97)])
def test_reference_pyramid_is_bit_exact(ssg, oracle32, name, shape):
    w, h = shape
    ref, _ = frames_pair(w, h, 11)
    r = np.random.default_rng(3)
    mask = (r.uniform(0, 1, (h, w)) > 0.1).astype(np.uint8) * 255
    g, f = make_pair(ssg, oracle32, name)
    g.set_reference_image(gpu(ref), gpu(mask))
    f.set_reference_image(ref, mask)
    assert g.num_levels() == len(f.levels)
    for lvl, e in enumerate(f.levels):
        assert g.level_size(lvl) == f.sizes[lvl]
        np.testing.assert_array_equal(g.reference_plane(lvl, "image"), e["ref"], err_msg=f"level {lvl} image")
        for k in ("Ix", "Iy", "a00", "a01", "a11", "idet"):
            np.testing.assert_array_equal(g.reference_plane(lvl, k), e[k], err_msg=f"level {lvl} {k}")


@pytest.mark.parametrize("name", list(OPTS))
@pytest.mark.parametrize("init", [False, True])
def test_compute_matches_oracle(ssg, oracle32, name, init):
    w, h = 160, 128
    ref, cur = frames_pair(w, h, 21)
    g, f = make_pair(ssg, oracle32, name)
    g.set_reference_image(gpu(ref))
    f.set_reference_image(ref)
    rmap = None
    if init:
        yy, xx = np.mgrid[0:h, 0:w].astype(np.float32)
        rmap = np.stack([xx + np.float32(1.1), yy - np.float32(0.6)], -1)
    want = f.compute(cur, rmap)
    got = g.compute(gpu(cur), None if rmap is None else gpu(rmap))
    d = np.abs(got - want)
    print(f"{name} init={init}: max |d| {d.max():.3g} px, exact {np.mean(d == 0):.4f}")
    assert d.max() <= FLOW_ATOL


def test_compute_with_masks_matches_oracle(ssg, oracle32):
    w, h = 160, 128
    ref, cur = frames_pair(w, h, 31)
    r = np.random.default_rng(4)
    rmask = np.full((h, w), 255, np.uint8)
    rmask[40:60, 30:90] = 0
    cmask = (r.uniform(0, 1, (h, w)) > 0.05).astype(np.uint8) * 255
    g, f = make_pair(ssg, oracle32, "pyramid")
    g.set_reference_image(gpu(ref), gpu(rmask))
    f.set_reference_image(ref, rmask)
    want = f.compute(cur, None, cmask)
    got = g.compute(gpu(cur), None, gpu(cmask))
    assert np.abs(got - want).max() <= FLOW_ATOL


def test_host_images_and_in_place_rmap(ssg, oracle32):
    w, h = 96, 80
    ref, cur = frames_pair(w, h, 41)
    g, f = make_pair(ssg, oracle32, "pyramid_ss2")
    g.set_reference_image(ref)
    f.set_reference_image(ref)
    yy, xx = np.mgrid[0:h, 0:w].astype(np.float32)
    rmap = np.stack([xx + np.float32(0.5), yy], -1)
    want = f.compute(cur, rmap)
    got = g.compute(cur, rmap)  # host views
    assert np.abs(got - want).max() <= FLOW_ATOL
    dev = gpu(rmap)
    g.compute(gpu(cur), dev, out=dev)  # rmap_out aliasing rmap_in
    assert np.abs(dev.cpu().numpy() - want).max() <= FLOW_ATOL


def test_batch_equals_single_calls(ssg, oracle32):
    import torch

    w, h, n = 160, 128, 5
    ref, _ = frames_pair(w, h, 51)
    curs = [frames_pair(w, h, 51, shift=(0.3 * i, -0.2 * i))[1] for i in range(n)]
    g, _ = make_pair(ssg, oracle32, "pyramid")
    g.set_reference_image(gpu(ref))
    yy, xx = np.mgrid[0:h, 0:w].astype(np.float32)
    rmaps = [gpu(np.stack([xx + np.float32(0.1 * i), yy], -1)) for i in range(n)]
    outs = [torch.zeros((h, w, 2), dtype=torch.float32, device="cuda") for _ in range(n)]
    dev = [gpu(c) for c in curs]
    g.compute_batch(dev, outs, rmaps=rmaps)
    for i in range(n):
        single = g.compute(dev[i], rmaps[i])
        np.testing.assert_array_equal(outs[i].cpu().numpy(), single)


def test_single_level_and_zero_support(ssg, oracle32):
    from oracle import eccflow as EF

    w, h = 64, 48
    ref, cur = frames_pair(w, h, 61, shift=(0.4, 0.2), amp=0.2)
    o = dict(downscale=2, support_scale=0, max_iterations=2, max_pyramid_level=0)
    g = ssg.c_eccflow(**o)
    f = EF.EccFlow(oracle32, **o)
    g.set_reference_image(gpu(ref))
    f.set_reference_image(ref)
    assert g.num_levels() == 1
    yy, xx = np.mgrid[0:h, 0:w].astype(np.float32)
    rmap = np.stack([xx + np.float32(0.3), yy + np.float32(0.1)], -1)
    assert np.abs(g.compute(gpu(cur), gpu(rmap)) - f.compute(cur, rmap)).max() <= FLOW_ATOL


def test_argument_errors(ssg):
    import torch

    g = ssg.c_eccflow(downscale=2, support_scale=2)
    with pytest.raises(ssg.SsgError):
        g.compute(gpu(np.zeros((32, 32), np.float32)))  # no reference yet
    g.set_reference_image(gpu(np.zeros((32, 32), np.float32)))
    with pytest.raises(ssg.SsgError):
        g.compute(gpu(np.zeros((32, 40), np.float32)))  # size mismatch
    with pytest.raises(ssg.SsgError):
        ssg.c_eccflow(downscale=7)
    with pytest.raises(ssg.SsgError):
        g.compute_batch([gpu(np.zeros((32, 32), np.float32))], [torch.zeros((32, 32), device="cuda")])


# ---------------------------------------------------------------- registration with the eccflow refine
def reg_options(ssg):
    o = ssg.c_image_registration_options(motion_type=3, scale=1.0, eps=0.02, ecch_max_level=2, max_iterations=50)
    o.enable_eccflow_registration = 1
    o.eccflow.downscale_method = 2
    o.eccflow.support_scale = 3
    o.eccflow.max_iterations = 2
    o.eccflow.min_image_size = -1
    o.eccflow.max_pyramid_level = 3
    return o


def eccflow_oracle_for(oracle32, o):
    from oracle import eccflow as EF

    e = o.eccflow
    return EF.EccFlow(oracle32, update_multiplier=e.update_multiplier, noise_level=e.noise_level,
                      max_iterations=e.max_iterations, support_scale=e.support_scale, min_image_size=e.min_image_size,
                      max_pyramid_level=e.max_pyramid_level, downscale=e.downscale_method, scale_factor=e.scale_factor)


def registration_frames(w, h, n, seed):
    S = scene(w, h, seed)
    r = np.random.default_rng(seed)
    yy, xx = np.mgrid[0:h, 0:w].astype(np.float64)
    ref = S([1, 0, 0, 0, 1, 0]).astype(np.float32)
    out = []
    for _ in range(n):
        a = np.array([1 + r.uniform(-0.004, 0.004), r.uniform(-0.01, 0.01), r.uniform(-3, 3),
                      r.uniform(-0.01, 0.01), 1 + r.uniform(-0.004, 0.004), r.uniform(-3, 3)])
        p = r.uniform(0, 2 * np.pi, 2)
        u = a[0] * xx + a[1] * yy + a[2] + 0.6 * np.sin(2 * np.pi * xx / (1.2 * w) + p[0])
        v = a[3] * xx + a[4] * yy + a[5] + 0.6 * np.cos(2 * np.pi * yy / (1.3 * h) + p[1])
        out.append(S.at(u, v).astype(np.float32))
    return ref, out


def test_register_frame_refines_the_current_remap(ssg, oracle32):
    w, h = 160, 128
    ref, frames = registration_frames(w, h, 2, 71)
    o = reg_options(ssg)
    reg = ssg.c_frame_registration(o)
    assert reg.setup_reference_frame(gpu(ref))
    f = eccflow_oracle_for(oracle32, o)
    f.set_reference_image(ref)
    for fr in frames:
        assert reg.register_frame(gpu(fr))
        st = reg.status()
        t = oracle32.xform(3, list(st.params))
        want = f.compute(fr, oracle32.create_map(t, w, h))
        got = reg.current_remap()
        np.testing.assert_array_equal(got, want)
        # c_frame_registration::remap warps with that field
        d, m = reg.remap(gpu(fr))
        cd, cm = reg.custom_remap(fr, rmap=want)
        np.testing.assert_array_equal(d, cd)
        np.testing.assert_array_equal(m, cm)


def test_batched_remaps_equal_per_frame_registration(ssg):
    import torch

    w, h, n = 160, 128, 4
    ref, frames = registration_frames(w, h, n, 81)
    o = reg_options(ssg)
    reg = ssg.c_frame_registration(o)
    assert reg.setup_reference_frame(gpu(ref))
    dev = [gpu(fr) for fr in frames]
    maps = [torch.zeros((h, w, 2), dtype=torch.float32, device="cuda") for _ in range(n)]
    st = reg.register_batch_remaps(dev, maps)
    for i in range(n):
        assert st[i].ok
        assert reg.register_frame(dev[i])
        np.testing.assert_array_equal(maps[i].cpu().numpy(), reg.current_remap())
    # without the refine the maps are the transforms' fields
    o.enable_eccflow_registration = 0
    plain = ssg.c_frame_registration(o)
    assert plain.setup_reference_frame(gpu(ref))
    st2 = plain.register_batch_remaps(dev, maps)
    for i in range(n):
        np.testing.assert_array_equal(maps[i].cpu().numpy(), plain.current_remap(st2[i]))


def test_flow_endpoint_error_vs_fp64_oracle(ssg, oracle64):
    """Reported figure: endpoint error of the GPU (FP32) flow against the same algorithm in FP64."""
    from oracle import eccflow as EF

    w, h = 256, 256
    ref, cur = frames_pair(w, h, 91, shift=(2.1, -1.4), amp=1.2)
    o = dict(downscale=2, support_scale=3, max_iterations=3, min_image_size=-1, max_pyramid_level=-1)
    g = ssg.c_eccflow(**o)
    g.set_reference_image(gpu(ref))
    got = g.compute(gpu(cur))
    f = EF.EccFlow(oracle64, **o)
    f.set_reference_image(ref.astype(np.float64))
    want = f.compute(cur.astype(np.float64))
    epe = np.hypot(got[..., 0] - want[..., 0], got[..., 1] - want[..., 1])
    print(f"flow EPE vs FP64 oracle: mean {epe.mean():.2e} px, p99 {np.percentile(epe, 99):.2e}, max {epe.max():.2e}")
    assert epe.mean() < 5e-3 and np.percentile(epe, 99) < 2e-2


# ---------------------------------------------------------------- dense-field warp + accumulate
def rel_l2(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def test_remapped_batch_equals_registered_batch_for_transform_fields(ssg):
    w, h, n = 160, 128, 5
    ref, frames = registration_frames(w, h, n, 101)
    reg = ssg.c_frame_registration(ssg.c_image_registration_options(motion_type=3, scale=1.0, eps=0.02,
                                                                     ecch_max_level=2, max_iterations=50))
    assert reg.setup_reference_frame(gpu(ref))
    dev = [gpu(f) for f in frames]
    st = reg.register_batch(dev)
    maps = [gpu(reg.current_remap(st[i])) for i in range(n)]
    wts = [0.5 + 0.1 * i for i in range(n)]
    for interp in (ssg.INTER_LINEAR, ssg.INTER_CUBIC):
        a = ssg.c_weigthed_average(w, h, 1)
        a.add_registered_batch(dev, st, 3, frame_weights=wts, interpolation=interp)
        b = ssg.c_weigthed_average(w, h, 1)
        b.add_remapped_batch(dev, maps, frame_weights=wts, interpolation=interp)
        for x, y in zip(a.compute(), b.compute()):
            np.testing.assert_array_equal(x, y)
        np.testing.assert_array_equal(a.counts(), b.counts())
        np.testing.assert_array_equal(a.get_acc_counters(), b.get_acc_counters())


def test_turbulence_pipeline_phasecorr_eccflow_bicubic_stack(ssg, oracle32):
    """Phase-correlation pre-alignment -> eccflow -> per-pixel bicubic remap -> average, against
    the oracle chain fed with the GPU's translation estimates."""
    from oracle import eccflow as EF

    w, h, n = 192, 160, 4
    S = scene(w, h, 111)
    r = np.random.default_rng(111)
    yy, xx = np.mgrid[0:h, 0:w].astype(np.float64)
    ref = S([1, 0, 0, 0, 1, 0]).astype(np.float32)
    frames = []
    for i in range(n):
        T = r.uniform(-6, 6, 2)
        p = r.uniform(0, 2 * np.pi, 2)
        ex = T[0] + 1.0 * np.sin(2 * np.pi * xx / (1.1 * w) + p[0])
        ey = T[1] + 1.0 * np.cos(2 * np.pi * yy / (1.2 * h) + p[1])
        frames.append(S.at(xx - ex, yy - ey).astype(np.float32))
    dev = [gpu(f) for f in frames]
    pc = ssg.phase_correlation(w, h, 1)
    pc.set_reference(gpu(ref))
    est = pc.estimate(dev)
    o = dict(downscale=2, support_scale=3, max_iterations=3, min_image_size=-1, max_pyramid_level=3)
    g = ssg.c_eccflow(**o)
    g.set_reference_image(gpu(ref))
    fx, fy = xx.astype(np.float32), yy.astype(np.float32)
    init = [np.stack([fx + np.float32(est[i, 0]), fy + np.float32(est[i, 1])], -1) for i in range(n)]
    maps = [gpu(m) for m in init]
    g.compute_batch(dev, maps, rmaps=maps)
    acc = ssg.c_weigthed_average(w, h, 1)
    acc.add_remapped_batch(dev, maps, interpolation=ssg.INTER_CUBIC)
    avg, mask = acc.compute()
    f = EF.EccFlow(oracle32, **o)
    f.set_reference_image(ref)
    A = np.zeros((h, w), np.float32)
    W = np.zeros((h, w), np.float32)
    cnt = np.zeros((h, w), np.int64)
    for i in range(n):
        m = f.compute(frames[i], init[i])
        np.testing.assert_array_equal(maps[i].cpu().numpy(), m)
        d, dm = oracle32.base_remap(frames[i], m, interp=2, border=4)
        oracle32.weighted_average_update(d, A, W, mask=dm)
        cnt += dm > 0
    np.testing.assert_array_equal(acc.counts(), cnt)
    assert rel_l2(avg, A) <= 1e-5
    # the flow brings the frames onto the reference
    inner = (slice(24, -24), slice(24, -24))
    assert rel_l2(avg[inner], ref[inner]) < 0.02

"""GPU FFT phase correlation against its numpy float64 oracle (oracle/phasecorr.py),
against known translations of the planetary generator, and chained into the
Bayer accumulator (the raw-frame pipeline: 2x2-binned phase correlation ->
c_bayer_average batch)."""
import numpy as np
import pytest

from tests import synth_np

pytestmark = pytest.mark.gpu

SUBPIXEL_TOL = 2e-3  # px: float32 FFT vs the float64 oracle, through the 3-point peak refinement


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


def _scene_frames(w, h, n, seed, max_shift=9.0):
    render = synth_np.blob_scene(w, h, seed=seed, n=int(0.012 * w * h), smin=1.2, smax=3.0)
    r = np.random.default_rng(seed + 1)
    ref = (render([1, 0, 0, 0, 1, 0]) + 0.01 * r.normal(size=(h, w))).astype(np.float32)
    T = r.uniform(-max_shift, max_shift, (n, 2))
    frames = [(render([1, 0, -t[0], 0, 1, -t[1]]) + 0.01 * r.normal(size=(h, w))).astype(np.float32) for t in T]
    return ref, frames, T


@pytest.mark.parametrize("binning", [1, 2])
@pytest.mark.parametrize("shape", [(256, 192), (320, 240)])
def test_phasecorr_matches_oracle(ssg, binning, shape):
    from oracle import phasecorr as pc

    S = ssg
    w, h = shape
    ref, frames, T = _scene_frames(w, h, 6, seed=int(w + binning))
    print(f"[phasecorr] {w}x{h} binning {binning}", flush=True)
    p = S.phase_correlation(w, h, binning)
    p.set_reference(gpu(ref))
    out = p.estimate([gpu(f) for f in frames])
    spec = pc.spectrum(ref, binning)
    for f, o, t in zip(frames, out, T):
        tx, ty, peak = pc.estimate(spec, f, binning)
        assert abs(o[0] - tx) < SUBPIXEL_TOL and abs(o[1] - ty) < SUBPIXEL_TOL, (o, tx, ty)
        assert abs(o[2] - peak) < 1e-5
        assert abs(o[0] - t[0]) < 0.35 and abs(o[1] - t[1]) < 0.35


def test_phasecorr_recovers_generator_translations(ssg):
    """Planetary frames (1024x1024, translation only, blur + Poisson + read noise)."""
    import torch

    S = ssg
    W = H = 1024
    p = S.synth_params(width=W, height=H, disk_radius=380.0, max_shift=10.0, max_rotation_deg=0.0, max_scale_dev=0.0,
                       max_shear=0.0, blur_sigma_min=0.8, blur_sigma_max=1.6, seed=77)
    frames = [torch.empty((H, W), dtype=torch.float32, device="cuda") for _ in range(8)]
    truth = S.synth_render(p, 0, frames)
    ref = torch.empty((H, W), dtype=torch.float32, device="cuda")
    S.synth_render(p, -1, [ref])
    pcor = S.phase_correlation(W, H, 1)
    pcor.set_reference(ref)
    out = pcor.estimate(frames)
    for o, m in zip(out, truth):
        # truth maps the reference grid into the frame: x' = x + t, so ref(x) = frame(x + t);
        # per-frame blur differs from the reference's and the noise is Poisson: a pre-alignment
        # grade estimate (the ECC that follows refines it)
        assert abs(o[0] - m[2]) < 0.15 and abs(o[1] - m[5]) < 0.15, (o, m)


def test_raw_bayer_pipeline_phasecorr_then_bayer_stack(ssg, oracle32):
    """u16 raw frames: 2x2-binned phase correlation gives translations; the Bayer
    batch stacks with them; the stack equals the oracle's per-frame sequence bit for bit."""
    import torch

    S = ssg
    W, H, n = 512, 384, 6
    p = S.synth_params(width=W, height=H, disk_radius=150.0, max_shift=8.0, max_rotation_deg=0.0, max_scale_dev=0.0,
                       max_shear=0.0, blur_sigma_min=0.8, blur_sigma_max=1.2, output_scale=1.0, seed=78)
    raw = [torch.empty((H, W), dtype=torch.int16, device="cuda") for _ in range(n)]
    truth = S.synth_render(p, 0, raw)
    ref = torch.empty((H, W), dtype=torch.int16, device="cuda")
    S.synth_render(p, -1, [ref])
    pcor = S.phase_correlation(W, H, 2)
    pcor.set_reference(ref)
    shifts = pcor.estimate(raw)
    for o, m in zip(shifts, truth):
        assert abs(o[0] - m[2]) < 0.35 and abs(o[1] - m[5]) < 0.35, (o, m)
    st = pcor.statuses(shifts)
    acc = S.c_bayer_average(W, H, 8, bpp=16)
    acc.add_registered_batch(raw, st, S.MOTION_TRANSLATION)
    ost = oracle32.bayer_state(W, H)
    for k in range(n):
        f = raw[k].cpu().numpy().view(np.uint16).astype(np.float32) / np.float32(65536)
        t = oracle32.xform(0, [st[k].params[0], st[k].params[1]])
        rmap = oracle32.create_map(t, W, H)
        _, mask = oracle32.base_remap(np.zeros((H, W), np.float32), rmap, interp=1, border=4)
        oracle32.bayer_accumulate(f, ost, 8, rmap=rmap, mask=mask)
    avg, mask = acc.compute()
    oavg, omask = oracle32.bayer_compute(ost)
    np.testing.assert_array_equal(avg, oavg)
    np.testing.assert_array_equal(acc.counts(), ost[2])

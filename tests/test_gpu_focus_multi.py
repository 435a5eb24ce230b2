"""ssg_focus_stack_batch: several focus sequences of different sizes in one batch, each step of
a chunk one launch over all sequences, against the FP32 oracle (oracle/focus.py) through the C
ABI.  Each frame is first remapped by its registration's translation (cv::remap INTER_LINEAR,
BORDER_REFLECT101, the oracle's base remap arithmetic), then added with the lpg-weight
composition: W = GaussianBlur(lpg(k, p), sigma) reduced to every level, acc_i += W_i L_i,
wsum_i += W_i.  Same canonical arithmetic on both sides: compared bit for bit."""
import numpy as np
import pytest

from tests.test_gpu_focus import focus_frames, gpu, planes

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ssg():
    import torch  # noqa: F401

    import serstacker_amd as S

    S._lib.load()
    return S


def oracle_stack(oracle32, frames, shifts, cn, sigma, max_levels):
    from oracle import focus

    fs = focus.FocusStack(oracle32, focus.LPG_WEIGHTS, lpg_k=6.0, lpg_p=2.0, lpg_sigma=sigma, max_levels=max_levels)
    for k, f in enumerate(frames):
        pl = planes(f)
        if shifts is not None:
            h, w = pl[0].shape
            t = oracle32.xform(0, [float(shifts[k][0]), float(shifts[k][1])])
            rmap = oracle32.create_map(t, w, h)
            pl = [oracle32.remap(p, rmap, interp=1, border=4) for p in pl]
        fs.add(pl)
    out = fs.compute()
    return out[0] if cn == 1 else np.stack(out, -1)


# (width, height, channels, frames, smoothing sigma, level cap, shifted)
SEQS = [(97, 71, 3, 5, 5.0, 0, True), (160, 120, 1, 3, 2.0, 6, False), (200, 150, 3, 4, 5.0, 4, True),
        (64, 48, 1, 37, 1.5, 0, True)]


def test_mixed_sequences_bitexact(ssg, oracle32):
    S = ssg
    rng = np.random.default_rng(17)
    stacks, frames, shifts, want = [], [], [], []
    for q, (w, h, cn, n, sigma, ml, shifted) in enumerate(SEQS):
        fr = focus_frames(n, w, h, cn, seed=40 + q)
        sh = rng.uniform(-3, 3, (n, 2)).astype(np.float32) if shifted else None
        stacks.append(S.c_laplacian_pyramid_focus_stacking(w, h, cn, fusing_policy=2, lpg_k=6.0, lpg_p=2.0,
                                                           lpg_sigma=sigma, max_levels=ml))
        frames.append([gpu(f) for f in fr])
        shifts.append(sh)
        want.append(oracle_stack(oracle32, fr, sh, cn, sigma, ml))
        print(f"[focus multi] sequence {q}: {w}x{h}x{cn}, {n} frames, sigma {sigma}, levels {ml}", flush=True)
    S.focus_stack_batch(stacks, frames, shifts)
    for q, st in enumerate(stacks):
        assert st.accumulated_frames() == SEQS[q][3]
        np.testing.assert_array_equal(st.compute(), want[q], err_msg=f"sequence {q}")


def test_batch_continues_single_adds(ssg):
    """A batch after single adds, then another batch: the stack of the same frames added one by
    one (no shifts), bit for bit; an empty sequence in the batch is left alone."""
    S = ssg
    kw = dict(fusing_policy=2, lpg_k=6.0, lpg_p=2.0, lpg_sigma=3.0)
    fa = focus_frames(6, 130, 98, 3, seed=3)
    fb = focus_frames(5, 90, 70, 1, seed=4)
    one_a = S.c_laplacian_pyramid_focus_stacking(130, 98, 3, **kw)
    one_b = S.c_laplacian_pyramid_focus_stacking(90, 70, 1, **kw)
    for f in fa:
        assert one_a.add(gpu(f))
    for f in fb:
        assert one_b.add(gpu(f))
    bat_a = S.c_laplacian_pyramid_focus_stacking(130, 98, 3, **kw)
    bat_b = S.c_laplacian_pyramid_focus_stacking(90, 70, 1, **kw)
    empty = S.c_laplacian_pyramid_focus_stacking(40, 30, 1, **kw)
    assert bat_a.add(gpu(fa[0]))
    S.focus_stack_batch([bat_a, bat_b, empty], [[gpu(f) for f in fa[1:4]], [gpu(f) for f in fb[:2]], []])
    S.focus_stack_batch([bat_b, bat_a], [[gpu(f) for f in fb[2:]], [gpu(f) for f in fa[4:]]])
    assert (bat_a.accumulated_frames(), bat_b.accumulated_frames(), empty.accumulated_frames()) == (6, 5, 0)
    np.testing.assert_array_equal(bat_a.compute(), one_a.compute())
    np.testing.assert_array_equal(bat_b.compute(), one_b.compute())


def test_batch_rejects_other_policies_and_sizes(ssg):
    S = ssg
    a = S.c_laplacian_pyramid_focus_stacking(64, 48, 1, fusing_policy=0)
    b = S.c_laplacian_pyramid_focus_stacking(64, 48, 1, fusing_policy=2)
    f = gpu(np.zeros((48, 64), np.float32))
    with pytest.raises(S.SsgError):
        S.focus_stack_batch([a], [[f]])
    with pytest.raises(S.SsgError):
        S.focus_stack_batch([b], [[gpu(np.zeros((40, 64), np.float32))]])
    with pytest.raises(S.SsgError):
        S.focus_stack_batch([b, b], [[f], [f]])
    assert b.accumulated_frames() == 0

"""The fused warp/accumulate at each tile shape of the SSG_TUNE_ACC_ROWS knob (64x8 and 64x16),
through the C ABI, on a batch that mixes interior tiles (the fast path: per-64-frame
classification, tap weights from the LDS table, positions quantised with one fma) and bordered
tiles (the frame edge, and one frame rotated by 9 degrees).

At every shape: per-pixel counts bit-exact against the FP32 oracle's base_remap masks, the stack
within 1e-5 relative L2 of the FP64 sum and of the reference-faithful FP32 incremental mean; the
two shapes add the same values in the same order, so their stacks are identical."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

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


def affine(ang_deg, sx, sy, scale=1.0, cx=0.0, cy=0.0):
    a = np.deg2rad(ang_deg)
    c, s = np.cos(a) * scale, np.sin(a) * scale
    # source position = M @ (x, y, 1) about (cx, cy)
    return [c, -s, sx + cx - c * cx + s * cy, s, c, sy + cy - s * cx - c * cy]


def frames_and_params(n=40, w=333, h=250, seed=3):
    rng = np.random.default_rng(seed)
    frames = [rng.random((h, w)).astype(np.float32) for _ in range(n)]
    params = []
    for i in range(n):
        if i == 7:  # a large rotation: more tiles take the bordered path
            params.append(affine(9.0, 1.5, -0.7, cx=w / 2, cy=h / 2))
        else:
            params.append(affine(rng.uniform(-0.4, 0.4), rng.uniform(-6, 6), rng.uniform(-6, 6),
                                 scale=1 + rng.uniform(-0.004, 0.004), cx=w / 2, cy=h / 2))
    return frames, params


def statuses(S, params):
    st = (S.RegistrationStatus * len(params))()
    for s, p in zip(st, params):
        s.ok = 1
        s.nparams = 6
        for i, v in enumerate(p):
            s.params[i] = v
    return st


def stack(S, frames, st, weights, rows):
    from serstacker_amd import autotune

    old = autotune.get_knob("acc_rows")
    autotune.set_knob("acc_rows", rows)
    try:
        h, w = frames[0].shape
        acc = S.c_weigthed_average(w, h, 1)
        acc.add_registered_batch([gpu_tensor(f) for f in frames], st, 3, frame_weights=weights)
        avg, mask = acc.compute()
        return avg, mask, acc.counts(), acc.accumulated_frames()
    finally:
        autotune.set_knob("acc_rows", old)


def test_tile_shapes_counts_and_stacks(ssg, oracle32):
    S = ssg
    frames, params = frames_and_params()
    h, w = frames[0].shape
    st = statuses(S, params)
    weights = [float(np.float32(0.5 + 0.01 * i)) for i in range(len(frames))]
    # oracle: base_remap of each frame, the weight remapped as a constant image, * mask / 255
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
        S_sum += np.where(mb, np.float64(wts) * d, 0)
        W_sum += np.where(mb, np.float64(wts), 0)
        n_cnt += mb
    ref_avg = np.where(W_sum > 0, S_sum / np.maximum(W_sum, 1e-300), 0)
    out = {}
    for rows in (2, 4):
        print(f"[accum tiles] rows per thread {rows}", flush=True)
        avg, mask, cnt, nacc = stack(S, frames, st, weights, rows)
        assert nacc == len(frames)
        assert np.array_equal(cnt, n_cnt), rows
        assert np.array_equal(mask > 0, W_sum > 0), rows
        e64, e32 = rel_l2(avg, ref_avg), rel_l2(avg, A32)
        print(f"  rel L2 vs FP64 sum {e64:.3g}, vs FP32 incremental mean {e32:.3g}", flush=True)
        assert e64 <= STACK_REL_L2 and e32 <= STACK_REL_L2
        out[rows] = avg
    assert np.array_equal(out[2], out[4])


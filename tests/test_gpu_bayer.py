"""GPU parity of c_bayer_average (core/average/c_frame_accumulation.cc:988-1334)
against the FP32 oracle, through the C ABI.

The kernels reproduce the reference arithmetic (float side weights, double
products, float accumulators, frames in order), so the average, the counters
and the integer per-cell counts are compared bit for bit.
"""
import ctypes as C

import numpy as np
import pytest

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


def check_same(S, acc, ora, st):
    avg, mask = acc.compute()
    oavg, omask = ora.bayer_compute(st)
    np.testing.assert_array_equal(avg, oavg.astype(np.float32))
    np.testing.assert_array_equal(mask, omask)
    cn = acc.get_acc_counters()
    np.testing.assert_array_equal(cn, st[1] * np.array([1, 0.5, 1], np.float32))
    np.testing.assert_array_equal(acc.counts(), st[2])


@pytest.mark.parametrize("colorid", [8, 9, 10, 11])
@pytest.mark.parametrize("src_type", ["f32", "u16"])
@pytest.mark.parametrize("wmode", ["none", "mask", "weights"])
def test_bayer_add_without_remap(ssg, oracle32, colorid, src_type, wmode):
    S = ssg
    r = np.random.default_rng(80 + colorid)
    W, H, n = 70, 46, 3
    acc = S.c_bayer_average(W, H, colorid, bpp=12)
    st = oracle32.bayer_state(W, H)
    for k in range(n):
        raw = r.integers(0, 4096, (H, W)).astype(np.uint16)
        f = raw.astype(np.float32) / np.float32(4096)
        m = (r.uniform(size=(H, W)) > 0.25).astype(np.uint8) * 255 if wmode == "mask" else None
        w = (r.uniform(0, 2, (H, W)) * (r.uniform(size=(H, W)) > 0.2)).astype(np.float32) if wmode == "weights" else None
        print(f"[bayer add] colorid {colorid} {src_type} {wmode} frame {k}", flush=True)
        assert acc.add(gpu(raw if src_type == "u16" else f), None if m is None and w is None else gpu(m if m is not None else w))
        oracle32.bayer_accumulate(f, st, colorid, mask=m, weights=w)
    assert acc.accumulated_frames() == n
    check_same(S, acc, oracle32, st)


@pytest.mark.parametrize("wmode", ["none", "mask", "weights"])
def test_bayer_add_with_remap(ssg, oracle32, wmode):
    """set_remap(rmap) + add(raw, weights): sub-pixel affine maps, some points
    outside, some with negative coordinates in (-1, 0) (truncated to 0)."""
    S = ssg
    r = np.random.default_rng(90)
    W, H = 96, 64
    acc = S.c_bayer_average(W, H, 8)
    st = oracle32.bayer_state(W, H)
    yy, xx = np.mgrid[0:H, 0:W].astype(np.float64)
    for k in range(4):
        A = [1 + 0.004 * r.normal(), 0.01 * r.normal(), r.uniform(-3, 3), 0.01 * r.normal(),
             1 + 0.004 * r.normal(), r.uniform(-3, 3)]
        rmap = np.stack([A[0] * xx + A[1] * yy + A[2], A[3] * xx + A[4] * yy + A[5]], -1).astype(np.float32)
        f = r.uniform(0, 1, (H, W)).astype(np.float32)
        m = (r.uniform(size=(H, W)) > 0.2).astype(np.uint8) * 255 if wmode == "mask" else None
        w = r.uniform(0, 1.5, (H, W)).astype(np.float32) if wmode == "weights" else None
        print(f"[bayer remap] {wmode} frame {k}", flush=True)
        acc.set_remap(gpu(rmap))
        assert acc.add(gpu(f), None if m is None and w is None else gpu(m if m is not None else w))
        oracle32.bayer_accumulate(f, st, 8, rmap=rmap, mask=m, weights=w)
    check_same(S, acc, oracle32, st)
    # clear() also drops the remap (c_frame_accumulation.cc:1154-1161)
    acc.clear()
    f = r.uniform(0, 1, (H, W)).astype(np.float32)
    acc.add(gpu(f))
    st2 = oracle32.bayer_state(W, H)
    oracle32.bayer_accumulate(f, st2, 8)
    check_same(S, acc, oracle32, st2)


def _statuses(S, params_list, motion):
    st = (S._lib.RegistrationStatus * len(params_list))()
    for i, p in enumerate(params_list):
        st[i].ok = 1
        for k, v in enumerate(p):
            st[i].params[k] = v
    return st


@pytest.mark.parametrize("motion", [0, 3])
@pytest.mark.parametrize("src_type", ["f32", "u16"])
def test_bayer_batch_equals_per_frame_sequence(ssg, oracle32, motion, src_type):
    """The batched kernel = the pipeline's per-frame set_remap(current_remap) +
    add(raw, current_mask) for every frame in order, mask = base_remap's."""
    S = ssg
    r = np.random.default_rng(100 + motion)
    W, H, n = 200, 136, 7
    raws = [r.integers(0, 4096, (H, W)).astype(np.uint16) for _ in range(n)]
    fl = [a.astype(np.float32) / np.float32(4096) for a in raws]
    params = []
    for k in range(n):
        tx, ty = 2 * r.integers(-6, 7) + r.uniform(-0.5, 0.5), 2 * r.integers(-6, 7) + r.uniform(-0.5, 0.5)
        if motion == 0:
            params.append([tx, ty])
        else:
            params.append([1 + 0.003 * r.normal(), 0.004 * r.normal(), tx, 0.004 * r.normal(), 1 + 0.003 * r.normal(), ty])
    sts = _statuses(S, params, motion)
    sts[3].ok = 0  # a rejected frame is skipped
    acc = S.c_bayer_average(W, H, 9, bpp=12)
    frames = [gpu(a if src_type == "u16" else b) for a, b in zip(raws, fl)]
    print(f"[bayer batch] motion {motion} {src_type} {W}x{H}x{n}", flush=True)
    acc.add_registered_batch(frames, sts, motion)
    assert acc.accumulated_frames() == n - 1
    st = oracle32.bayer_state(W, H)
    for k in range(n):
        if k == 3:
            continue
        t = oracle32.xform(motion, [np.float32(v) for v in params[k]])
        rmap = oracle32.create_map(t, W, H)
        _, mask = oracle32.base_remap(np.zeros((H, W), np.float32), rmap, interp=1, border=4)
        oracle32.bayer_accumulate(fl[k], st, 9, rmap=rmap, mask=mask)
    check_same(S, acc, oracle32, st)

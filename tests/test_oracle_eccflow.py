"""Pins of the eccflow oracle (oracle/eccflow.py) on the CPU.

The resize restatements are checked against what the mathematics fixes
(brute-force Keys-kernel and exact-area evaluations in float64, linear-ramp
reproduction, block means); the flow estimator against known answers on
band-limited analytic scenes: a zero flow for identical frames, the update
multiplier's overshoot on a small shift, and recovery of a global shift and of
a smooth displacement field (ecc2.cc:2103-2865)."""
import math

import numpy as np
import pytest

from oracle import eccflow as EF
from tests import synth_np


# ---------------------------------------------------------------- resize INTER_CUBIC
def keys(t, a=-0.75):
    t = np.abs(t)
    return np.where(t <= 1, (a + 2) * t ** 3 - (a + 3) * t ** 2 + 1,
                    np.where(t < 2, a * t ** 3 - 5 * a * t ** 2 + 8 * a * t - 4 * a, 0.0))


def brute_cubic(img, dw, dh):
    h, w = img.shape
    out = np.zeros((dh, dw))
    sx, sy = w / dw, h / dh
    for dy in range(dh):
        fy = (dy + 0.5) * sy - 0.5
        y0 = math.floor(fy)
        for dx in range(dw):
            fx = (dx + 0.5) * sx - 0.5
            x0 = math.floor(fx)
            s = 0.0
            for j in range(y0 - 1, y0 + 3):
                wy = keys(fy - j)
                for i in range(x0 - 1, x0 + 3):
                    s += wy * keys(fx - i) * img[min(max(j, 0), h - 1), min(max(i, 0), w - 1)]
            out[dy, dx] = s
    return out


@pytest.mark.parametrize("src,dst", [((16, 12), (32, 24)), ((16, 8), (8, 4)), ((13, 9), (40, 17)), ((20, 20), (7, 31))])
def test_resize_cubic_matches_keys_kernel(src, dst):
    r = np.random.default_rng(src[0] * 7 + dst[1])
    img = r.uniform(-1, 1, (src[1], src[0])).astype(np.float32)
    got = EF.resize_cubic(img, *dst)
    ref = brute_cubic(img.astype(np.float64), *dst)
    assert got.shape == (dst[1], dst[0])
    np.testing.assert_allclose(got, ref, atol=2e-6, rtol=0)


def test_resize_cubic_interpolates_and_keeps_constants():
    # 3x upscale: output dx = 3i + 1 sits on source pixel i (t = 0), where the kernel is (0, 1, 0, 0)
    r = np.random.default_rng(9)
    img = r.uniform(-1, 1, (10, 14)).astype(np.float32)
    up = EF.resize_cubic(img, 42, 30)
    np.testing.assert_array_equal(up[1::3, 1::3], img)
    const = np.full((7, 9), 0.7, np.float32)
    np.testing.assert_allclose(EF.resize_cubic(const, 23, 16), 0.7, atol=3e-7)


def test_resize_cubic_equal_size_is_a_copy_and_channels_are_independent():
    r = np.random.default_rng(3)
    img = r.uniform(-1, 1, (9, 11, 2)).astype(np.float32)
    np.testing.assert_array_equal(EF.resize_cubic(img, 11, 9), img)
    two = EF.resize_cubic(img, 30, 20)
    for c in range(2):
        np.testing.assert_array_equal(two[..., c], EF.resize_cubic(img[..., c], 30, 20))


# ---------------------------------------------------------------- resize INTER_AREA
def brute_area(img, dw, dh):
    h, w = img.shape

    def overlaps(n, d):
        sc = n / d
        W = np.zeros((d, n))
        for o in range(d):
            a, b = o * sc, (o + 1) * sc
            for i in range(n):
                W[o, i] = max(0.0, min(b, i + 1) - max(a, i))
        return W / W.sum(1, keepdims=True)

    return overlaps(h, dh) @ img.astype(np.float64) @ overlaps(w, dw).T


@pytest.mark.parametrize("src,dst", [((64, 48), (8, 6)), ((48, 40), (16, 5)), ((100, 60), (40, 24)),
                                     ((1080, 34), (68, 17)), ((37, 29), (9, 8))])
def test_resize_area_matches_exact_cell_means(src, dst):
    r = np.random.default_rng(src[0] + dst[0])
    img = r.uniform(0, 1, (src[1], src[0])).astype(np.float32)
    got = EF.resize_area(img, *dst)
    np.testing.assert_allclose(got, brute_area(img, *dst), atol=3e-6, rtol=0)


def test_resize_area_fast_path_sums_in_groups_of_four():
    # a 2x2 cell of (1, 2^-24, 2^-24, 2^-24): ((1 + e) + e) + e rounds each step back to 1,
    # a pairwise sum (1 + e) + (e + e) would reach 1 + 2^-23
    e = np.float32(2.0 ** -24)
    img = np.array([[1, e], [e, e]], np.float32)
    assert EF.area_is_fast(2, 1)[0]
    assert EF.resize_area(img, 1, 1)[0, 0] == np.float32(0.25)


def test_area_tables_are_partitions_of_unity():
    for n, d in [(1080, 68), (135, 9), (37, 9), (768, 576 // 8)]:
        sums = np.zeros(d)
        for di, si, al in EF.area_tab(n, d):
            assert 0 <= si < n
            sums[di] += float(al)
        np.testing.assert_allclose(sums, 1.0, atol=2e-6)


def test_is_area_fast_needs_an_exact_integer_scale():
    assert EF.area_is_fast(1024, 128) == (True, 8)
    assert EF.area_is_fast(1080, 135) == (True, 8)
    assert not EF.area_is_fast(540, 68)[0]


# ---------------------------------------------------------------- masks
def test_remap_nearest_rounds_half_to_even_and_zeroes_outside():
    m = (np.arange(20, dtype=np.uint8).reshape(4, 5) + 1)
    rmap = np.array([[[0.5, 0.0], [1.5, 0.0], [2.49, 1.51], [-0.6, 0.0], [4.4, 3.4], [4.6, 0.0]]], np.float32)
    got = EF.remap_nearest_u8(m, rmap)
    np.testing.assert_array_equal(got[0], [m[0, 0], m[0, 2], m[2, 2], 0, m[3, 4], 0])


def test_resize_nearest_matches_the_c_oracle(oracle32):
    r = np.random.default_rng(5)
    m = r.integers(0, 256, (37, 53)).astype(np.uint8)
    for dw, dh in [(27, 19), (53, 37), (13, 5)]:
        np.testing.assert_array_equal(EF.resize_nearest_u8(m, dw, dh), oracle32.resize_nearest_u8(m, (dw, dh)))


# ---------------------------------------------------------------- pyramid geometry
def test_level_sizes(oracle32):
    f = EF.EccFlow(oracle32, downscale=EF.PYRAMID, max_pyramid_level=4, support_scale=3, min_image_size=-1)
    assert f.level_sizes(1024, 1024)[0] == [(1024, 1024), (512, 512), (256, 256), (128, 128), (64, 64)]
    f = EF.EccFlow(oracle32, downscale=EF.PYRAMID, min_image_size=-1)
    assert f.level_sizes(40, 22)[0] == [(40, 22), (20, 11), (10, 6)]
    f = EF.EccFlow(oracle32, downscale=EF.RECURSIVE_RESIZE, scale_factor=0.75, min_image_size=-1)
    sizes, srcs = f.level_sizes(100, 40)
    # (int)((n+1) 0.75) while max(size) > 4; the smaller side reaching 5 (min_size + 1)
    # switches to level 0 as the source for a big aspect ratio
    assert sizes[:4] == [(100, 40), (75, 30), (57, 23), (43, 18)]
    assert max(sizes[-1]) > 4 and srcs[1:4] == [0, 1, 2]
    for (w, h), s in zip(sizes[1:], srcs[1:]):
        assert s == (0 if min(w, h) <= 5 else s)
    f = EF.EccFlow(oracle32, downscale=EF.FULL_RESIZE, scale_factor=0.5, min_image_size=4)
    sizes, srcs = f.level_sizes(64, 64)
    assert sizes == [(64, 64), (32, 32), (16, 16), (8, 8)] and srcs[1:] == [0, 0, 0]


# ---------------------------------------------------------------- the flow estimator
def scene(w, h, seed):
    # dense texture: the cell solve divides by the local structure-tensor determinant
    return synth_np.blob_scene(w, h, seed, n=int(w * h / 50), smin=3.0, smax=6.0)


def test_identical_frames_give_zero_flow(oracle32):
    w, h = 96, 80
    S = scene(w, h, 1)
    ref = S([1, 0, 0, 0, 1, 0]).astype(np.float32)
    f = EF.EccFlow(oracle32, downscale=EF.PYRAMID, support_scale=2, max_iterations=2, min_image_size=-1)
    f.set_reference_image(ref)
    uv = f.compute_uv(ref)
    assert np.all(uv == 0)
    yy, xx = np.mgrid[0:h, 0:w].astype(np.float32)
    ident = np.stack([xx, yy], -1)
    np.testing.assert_array_equal(f.compute(ref), ident)
    # an identity initial remap is the zero flow exactly
    np.testing.assert_array_equal(f.compute_uv(ref, ident), uv)


def test_one_iteration_overshoots_by_the_update_multiplier(oracle32):
    # for a small shift T the linearised cell solve returns update_multiplier * T
    w, h = 128, 128
    S = scene(w, h, 2)
    T = np.array([0.06, -0.04])
    ref = S([1, 0, 0, 0, 1, 0]).astype(np.float32)
    cur = S([1, 0, -T[0], 0, 1, -T[1]]).astype(np.float32)  # cur(x + T) = ref(x)
    for um in (1.5, 1.0):
        f = EF.EccFlow(oracle32, update_multiplier=um, downscale=EF.PYRAMID, max_pyramid_level=0, support_scale=3,
                       max_iterations=1)
        f.set_reference_image(ref)
        uv = f.compute_uv(cur)
        inner = uv[24:-24, 24:-24]
        np.testing.assert_allclose(np.median(inner[..., 0]), um * T[0], rtol=0.08)
        np.testing.assert_allclose(np.median(inner[..., 1]), um * T[1], rtol=0.08)


def test_global_shift_is_recovered(oracle32):
    w, h = 160, 128
    S = scene(w, h, 3)
    T = np.array([1.7, -1.2])
    ref = S([1, 0, 0, 0, 1, 0]).astype(np.float32)
    cur = S([1, 0, -T[0], 0, 1, -T[1]]).astype(np.float32)
    f = EF.EccFlow(oracle32, downscale=EF.PYRAMID, support_scale=3, max_iterations=3, min_image_size=-1)
    f.set_reference_image(ref)
    uv = f.compute_uv(cur)
    err = np.hypot(uv[..., 0] - T[0], uv[..., 1] - T[1])[16:-16, 16:-16]
    assert np.median(err) < 0.04 and np.percentile(err, 95) < 0.1


def smooth_field(w, h, seed, amp):
    r = np.random.default_rng(seed)
    p = r.uniform(0, 2 * np.pi, 4)

    def e(x, y):
        return (amp * np.sin(2 * np.pi * x / (1.3 * w) + p[0]) * np.cos(2 * np.pi * y / (1.1 * h) + p[1]),
                amp * np.cos(2 * np.pi * x / (0.9 * w) + p[2]) * np.sin(2 * np.pi * y / (1.4 * h) + p[3]))

    return e


def true_flow(e, w, h):
    # cur(p) = S(p - e(p)); the flow solves uv = e(x + uv) (fixed point, |grad e| << 1)
    yy, xx = np.mgrid[0:h, 0:w].astype(np.float64)
    u = np.zeros((h, w))
    v = np.zeros((h, w))
    for _ in range(30):
        u, v = e(xx + u, yy + v)
    return u, v


def test_smooth_displacement_field_is_recovered(oracle32):
    w, h = 192, 160
    S = scene(w, h, 4)
    e = smooth_field(w, h, 4, 1.5)
    yy, xx = np.mgrid[0:h, 0:w].astype(np.float64)
    ex, ey = e(xx, yy)
    ref = S([1, 0, 0, 0, 1, 0]).astype(np.float32)
    cur = S.at(xx - ex, yy - ey).astype(np.float32)
    f = EF.EccFlow(oracle32, downscale=EF.PYRAMID, support_scale=3, max_iterations=3, min_image_size=-1)
    f.set_reference_image(ref)
    uv = f.compute_uv(cur)
    tu, tv = true_flow(e, w, h)
    err = np.hypot(uv[..., 0] - tu, uv[..., 1] - tv)[16:-16, 16:-16]
    assert np.median(err) < 0.05 and np.percentile(err, 95) < 0.2
    # the estimate starting from a global-shift initial remap agrees
    rmap = np.stack([xx + 0.4, yy - 0.3], -1).astype(np.float32)
    uv2 = f.compute_uv(cur, rmap)
    err2 = np.hypot(uv2[..., 0] - tu, uv2[..., 1] - tv)[16:-16, 16:-16]
    assert np.median(err2) < 0.05


def test_masks_gate_the_update(oracle32):
    w, h = 96, 80
    S = scene(w, h, 5)
    ref = S([1, 0, 0, 0, 1, 0]).astype(np.float32)
    cur = S([1, 0, -0.5, 0, 1, 0.3]).astype(np.float32)
    full = np.full((h, w), 255, np.uint8)
    kw = dict(downscale=EF.PYRAMID, support_scale=2, max_iterations=2, min_image_size=-1)
    g = EF.EccFlow(oracle32, **kw)
    g.set_reference_image(ref)
    plain = g.compute_uv(cur)
    # a full reference mask changes nothing
    f = EF.EccFlow(oracle32, **kw)
    f.set_reference_image(ref, full)
    np.testing.assert_array_equal(f.compute_uv(cur), plain)
    # a full input mask is remapped with BORDER_CONSTANT: with the identity map of a
    # single iteration at level 0 every pixel stays inside, so nothing changes either
    one = dict(downscale=EF.PYRAMID, support_scale=2, max_iterations=1, max_pyramid_level=0)
    a = EF.EccFlow(oracle32, **one)
    a.set_reference_image(ref)
    np.testing.assert_array_equal(a.compute_uv(cur, None, full), a.compute_uv(cur))
    # ... while a shifted map loses the columns that leave the image
    yy, xx = np.mgrid[0:h, 0:w].astype(np.float32)
    rmap = np.stack([xx + 3.0, yy], -1)
    assert not np.array_equal(a.compute_uv(cur, rmap, full), a.compute_uv(cur, rmap))
    # an empty input mask zeroes every residual: the flow stays the initial one
    rmap = np.stack([xx + 0.25, yy], -1)
    np.testing.assert_array_equal(g.compute(cur, rmap, np.zeros((h, w), np.uint8)), _expected_init(g, rmap))


def _expected_init(f, rmap):
    """identity + the initial flow carried to level 0 with no updates."""
    h0, w0 = rmap.shape[:2]
    saved = f.iters
    f.iters = 0
    try:
        return f.compute(np.zeros((h0, w0), np.float32), rmap)
    finally:
        f.iters = saved

"""Pins of the CPU oracle (oracle/ssg_oracle.c) against what fixes each function
independently of its code: library routines (numpy, scipy.ndimage) for the
filters, borders, morphology and statistics; brute-force restatements in
float64 written from the documented OpenCV definitions (not from the oracle's
loops) for pyrDown/pyrUp/remap; closed forms (exact shifts, reproduced linear
functions, constant images, incremental-mean identity); and known-answer
registrations of analytically warped scenes.

The reference's own test tree holds no golden vectors for this path, and the
reference cannot be built here (it needs OpenCV, oneTBB and libconfig), so these
checks are what pins the oracle.  Each is chosen so that a plausible slip -- a
transposed kernel, a wrong border rule, an off-by-one tap, a flipped sign in
the ECC update -- fails it.
"""
import numpy as np
import pytest
from scipy import ndimage

from tests import synth_np

# OpenCV border codes
CONSTANT, REPLICATE, REFLECT, WRAP, REFLECT101 = 0, 1, 2, 3, 4
NP_PAD = {REPLICATE: "edge", REFLECT: "symmetric", REFLECT101: "reflect", WRAP: "wrap"}
ND_MODE = {REPLICATE: "nearest", REFLECT: "reflect", REFLECT101: "mirror", WRAP: "grid-wrap", CONSTANT: "constant"}


def rng(seed=0):
    return np.random.default_rng(seed)


# ---- borders ------------------------------------------------------------------------------------
@pytest.mark.parametrize("border", [REPLICATE, REFLECT, REFLECT101, WRAP])
@pytest.mark.parametrize("n", [1, 2, 3, 7])
def test_border_interpolate_matches_numpy_pad(oracle64, border, n):
    pad = 3 * n + 2
    idx = np.pad(np.arange(n), pad, mode=NP_PAD[border])
    for p in range(-pad, n + pad):
        assert oracle64.border_interpolate(p, n, border) == idx[p + pad], (p, n, border)


def test_border_interpolate_constant_is_minus_one(oracle64):
    assert oracle64.border_interpolate(-1, 5, CONSTANT) == -1
    assert oracle64.border_interpolate(5, 5, CONSTANT) == -1
    assert oracle64.border_interpolate(3, 5, CONSTANT) == 3


# ---- gaussian kernel ----------------------------------------------------------------------------
def test_gaussian_kernel_fixed_small_tables(oracle64):
    # getGaussianKernel's fixed tables for sigma <= 0 and n <= 7: binomial rows / 2^(n-1)
    from math import comb
    for n in (1, 3, 5, 7):
        binom = np.array([comb(n - 1, k) for k in range(n)], np.float64)
        want = binom / binom.sum()
        if n == 7:  # OpenCV's 7-tap table is not binomial: {1, 3.5, 7, 9, 7, 3.5, 1} / 32
            want = np.array([1, 3.5, 7, 9, 7, 3.5, 1]) / 32
        np.testing.assert_allclose(oracle64.gaussian_kernel(n, 0.0), want, rtol=0, atol=1e-15)


@pytest.mark.parametrize("n,sigma", [(7, 1.0), (9, 1.5), (13, 2.0), (11, 0.0)])
def test_gaussian_kernel_closed_form(oracle64, n, sigma):
    s = sigma if sigma > 0 else ((n - 1) * 0.5 - 1) * 0.3 + 0.8
    x = np.arange(n) - (n - 1) / 2
    want = np.exp(-x * x / (2 * s * s))
    want /= want.sum()
    k = oracle64.gaussian_kernel(n, sigma)
    np.testing.assert_allclose(k, want, rtol=1e-13)
    assert abs(k.sum() - 1) < 1e-14 and np.allclose(k, k[::-1], rtol=0, atol=1e-16)


# ---- separable filter, differentiation ----------------------------------------------------------
@pytest.mark.parametrize("border", [REPLICATE, REFLECT, REFLECT101, CONSTANT])
@pytest.mark.parametrize("kind", ["sym5", "asym5", "gen3"])
def test_sep_filter_matches_scipy_correlate(oracle64, oracle32, border, kind):
    r = rng(1)
    src = r.normal(size=(29, 37))
    if kind == "sym5":
        kx = np.array([0.1, 0.2, 0.4, 0.2, 0.1]); ky = np.array([0.05, 0.25, 0.4, 0.25, 0.05])
    elif kind == "asym5":
        kx = np.array([1 / 12, -2 / 3, 0, 2 / 3, -1 / 12]); ky = np.array([0.25, 0.5, 0.25, 0, 0])[:3]
    else:
        kx = np.array([0.3, -0.9, 0.45]); ky = np.array([0.7, 0.1, -0.35])
    mode = ND_MODE[border]
    want = ndimage.correlate1d(ndimage.correlate1d(src, kx, axis=1, mode=mode, cval=0.0), ky, axis=0, mode=mode,
                               cval=0.0)
    np.testing.assert_allclose(oracle64.sep_filter(src, kx, ky, border), want, rtol=0, atol=1e-12)
    got32 = oracle32.sep_filter(src.astype(np.float32), kx.astype(np.float32), ky.astype(np.float32), border)
    np.testing.assert_allclose(got32, want, rtol=0, atol=2e-5)


def test_sep_filter_impulse_response_is_outer_product(oracle64):
    src = np.zeros((15, 17)); src[7, 8] = 1.0
    kx = np.array([1.0, 2.0, 5.0, -3.0, 0.5]); ky = np.array([0.25, -1.0, 4.0])
    out = oracle64.sep_filter(src, kx, ky, CONSTANT)
    # correlation: out(y, x) = sum ky[j] kx[i] src(y + j - 1, x + i - 2) -> flipped kernels around the impulse
    want = np.zeros_like(src)
    want[6:9, 6:11] = np.outer(ky[::-1], kx[::-1])
    np.testing.assert_array_equal(out, want)


def test_ecc_differentiate_matches_scipy(oracle64):
    src = rng(2).normal(size=(23, 31))
    d5 = np.array([1 / 12, -2 / 3, 0, 2 / 3, -1 / 12]); s3 = np.array([0.25, 0.5, 0.25])
    gx, gy = oracle64.ecc_differentiate(src)
    wx = ndimage.correlate1d(ndimage.correlate1d(src, d5, axis=1, mode="nearest"), s3, axis=0, mode="nearest")
    wy = ndimage.correlate1d(ndimage.correlate1d(src, s3, axis=1, mode="nearest"), d5, axis=0, mode="nearest")
    # d5 is the float32 table of the reference (ecc2.cc:142-169): 1e-8 absolute on O(1) data
    np.testing.assert_allclose(gx, wx, atol=1e-7)
    np.testing.assert_allclose(gy, wy, atol=1e-7)
    # a linear ramp has the exact derivative away from the border
    yy, xx = np.mgrid[0:23, 0:31].astype(np.float64)
    gx, gy = oracle64.ecc_differentiate(3.0 * xx - 2.0 * yy)
    np.testing.assert_allclose(gx[2:-2, 2:-2], 3.0, atol=1e-6)
    np.testing.assert_allclose(gy[2:-2, 2:-2], -2.0, atol=1e-6)


# ---- pyramids -----------------------------------------------------------------------------------
def _reflect101(i, n):
    return np.pad(np.arange(n), 4 * n, mode="reflect")[i + 4 * n]


def pyr_down_brute(src, dw, dh):
    """dst(y, x) = sum_ij k_i k_j src(2y+i-2, 2x+j-2) / 256, k = [1 4 6 4 1], REFLECT101."""
    h, w = src.shape
    k = np.array([1, 4, 6, 4, 1], np.float64)
    out = np.zeros((dh, dw))
    for y in range(dh):
        rows = [_reflect101(2 * y + i - 2, h) for i in range(5)]
        for x in range(dw):
            cols = [_reflect101(2 * x + j - 2, w) for j in range(5)]
            out[y, x] = (np.outer(k, k) * src[np.ix_(rows, cols)]).sum() / 256.0
    return out


@pytest.mark.parametrize("shape,dsize", [((29, 37), None), ((30, 40), (20, 15)), ((31, 33), (16, 16)),
                                         ((17, 24), (12, 8))])
def test_pyr_down_brute_force(oracle64, shape, dsize):
    src = rng(3).normal(size=shape)
    dw, dh = dsize if dsize else ((shape[1] + 1) // 2, (shape[0] + 1) // 2)
    np.testing.assert_allclose(oracle64.pyr_down(src, (dw, dh)), pyr_down_brute(src, dw, dh), rtol=0, atol=1e-13)


def test_pyr_down_u8_rounds_half_up(oracle64):
    src = rng(4).integers(0, 256, (21, 26)).astype(np.uint8)
    dw, dh = 13, 11
    want = np.floor(pyr_down_brute(src.astype(np.float64), dw, dh) * 256 + 128) / 256
    np.testing.assert_array_equal(oracle64.pyr_down_u8(src, (dw, dh)), np.floor(want).astype(np.uint8))


def pyr_up_brute(src, dw, dh):
    """Zero-insertion upsampling convolved with 4*[1 4 6 4 1]/16 per axis: even outputs
    (s[i-1] + 6 s[i] + s[i+1]) / 8, odd outputs (4 s[i] + 4 s[i+1]) / 8; s[-1] = s[1]
    (REFLECT101 on the left/top), s[n] = s[n-1] (replicated on the right/bottom).  An odd
    extra output column repeats the last column; an odd extra row repeats row 2n-2
    (OpenCV 4.5 pyrUp_)."""
    def axis(a, n_out, extra_from):
        n = a.shape[0]
        ext = np.concatenate([a[1:2] if n > 1 else a[0:1], a, a[-1:]], axis=0)  # s[-1], s[0..n-1], s[n]
        up = np.zeros((2 * n + 4,) + a.shape[1:])
        up[0:2 * n + 3:2] = ext  # s[i] sits at 2i + 2
        k = np.array([1, 4, 6, 4, 1], np.float64) / 8.0
        out = np.stack([sum(k[t] * up[o + t] for t in range(5)) for o in range(2 * n)])
        if n_out > 2 * n:
            out = np.concatenate([out, out[extra_from(n):extra_from(n) + 1]], axis=0)
        return out[:n_out]
    rows = axis(src, dh, lambda n: 2 * n - 2)
    return axis(rows.T, dw, lambda n: 2 * n - 1).T


@pytest.mark.parametrize("shape,dsize", [((9, 13), None), ((10, 7), (14, 20)), ((6, 6), (13, 13))])
def test_pyr_up_brute_force(oracle64, shape, dsize):
    src = rng(5).normal(size=shape)
    dw, dh = dsize if dsize else (2 * shape[1], 2 * shape[0])
    np.testing.assert_allclose(oracle64.pyr_up(src, (dw, dh)), pyr_up_brute(src, dw, dh), rtol=0, atol=1e-13)
    # a constant image stays constant
    np.testing.assert_allclose(oracle64.pyr_up(np.full(shape, 2.5), (dw, dh)), 2.5, atol=1e-14)


def test_resize_nearest_index_rule(oracle64):
    src = rng(6).integers(0, 256, (13, 17)).astype(np.uint8)
    dw, dh = 9, 7
    sy = np.minimum(np.floor(np.arange(dh) * 13 / dh).astype(int), 12)
    sx = np.minimum(np.floor(np.arange(dw) * 17 / dw).astype(int), 16)
    np.testing.assert_array_equal(oracle64.resize_nearest_u8(src, (dw, dh)), src[np.ix_(sy, sx)])


# ---- remap --------------------------------------------------------------------------------------
def _quant(m):
    X = np.rint(m * 32.0).astype(np.int64)  # cvRound: round half to even
    return X >> 5, X & 31


def _tap(src, y, x, border, cval):
    h, w = src.shape
    if border == CONSTANT:
        ok = (x >= 0) & (x < w) & (y >= 0) & (y < h)
        return np.where(ok, src[np.clip(y, 0, h - 1), np.clip(x, 0, w - 1)], cval)
    ypad = np.pad(np.arange(h), 4 * h, mode=NP_PAD[border])
    xpad = np.pad(np.arange(w), 4 * w, mode=NP_PAD[border])
    return src[ypad[y + 4 * h], xpad[x + 4 * w]]


def remap_linear_brute(src, rmap, border, cval):
    ix, fx = _quant(rmap[..., 0])
    iy, fy = _quant(rmap[..., 1])
    tx, ty = fx / 32.0, fy / 32.0
    h, w = src.shape
    v = ((1 - ty) * (1 - tx) * _tap(src, iy, ix, border, cval) + (1 - ty) * tx * _tap(src, iy, ix + 1, border, cval)
         + ty * (1 - tx) * _tap(src, iy + 1, ix, border, cval) + ty * tx * _tap(src, iy + 1, ix + 1, border, cval))
    if border == CONSTANT:
        outside = (ix >= w) | (ix + 1 < 0) | (iy >= h) | (iy + 1 < 0)
        v = np.where(outside, cval, v)
    return v


def keys(t, a=-0.75):
    t = np.abs(t)
    return np.where(t <= 1, (a + 2) * t ** 3 - (a + 3) * t ** 2 + 1,
                    np.where(t < 2, a * t ** 3 - 5 * a * t ** 2 + 8 * a * t - 4 * a, 0.0))


def remap_cubic_brute(src, rmap, border, cval):
    ix, fx = _quant(rmap[..., 0])
    iy, fy = _quant(rmap[..., 1])
    tx, ty = fx / 32.0, fy / 32.0
    v = np.zeros(rmap.shape[:2])
    for j in range(4):
        for i in range(4):
            v += keys(ty - (j - 1)) * keys(tx - (i - 1)) * _tap(src, iy + j - 1, ix + i - 1, border, cval)
    return v


def _random_affine_map(w, h, seed, spread=1.0):
    r = rng(seed)
    A = np.array([[1 + 0.05 * r.normal(), 0.08 * r.normal(), r.uniform(-6, 6) * spread],
                  [0.08 * r.normal(), 1 + 0.05 * r.normal(), r.uniform(-6, 6) * spread]])
    from oracle.oracle import affine_map
    return affine_map(A.ravel(), w, h)


@pytest.mark.parametrize("border", [CONSTANT, REPLICATE, REFLECT101, REFLECT, WRAP])
def test_remap_linear_brute_force(oracle64, border):
    src = rng(7).normal(size=(24, 30))
    rmap = _random_affine_map(34, 27, 8, spread=1.5)
    got = oracle64.remap(src, rmap, interp=1, border=border, cval=0.7)
    np.testing.assert_allclose(got, remap_linear_brute(src, rmap, border, 0.7), rtol=0, atol=1e-12)


@pytest.mark.parametrize("border", [CONSTANT, REPLICATE, REFLECT101])
def test_remap_cubic_brute_force(oracle64, border):
    src = rng(9).normal(size=(24, 30))
    rmap = _random_affine_map(28, 22, 10)
    got = oracle64.remap(src, rmap, interp=2, border=border, cval=0.3)
    # the coefficient table is float32, as in OpenCV's interpolateCubic
    np.testing.assert_allclose(got, remap_cubic_brute(src, rmap, border, 0.3), rtol=0, atol=2e-6)


def test_remap_closed_forms(oracle32):
    r = rng(11)
    src = r.normal(size=(20, 26)).astype(np.float32)
    yy, xx = np.mgrid[0:20, 0:26].astype(np.float32)
    # integer shift: an exact copy (float32, bit for bit)
    rmap = np.stack([xx + 3, yy - 2], -1)
    got = oracle32.remap(src, rmap, interp=1, border=REPLICATE)
    np.testing.assert_array_equal(got[2:, :23], src[:18, 3:])
    got = oracle32.remap(src, rmap, interp=2, border=REPLICATE)
    np.testing.assert_array_equal(got[3:-1, 1:22], src[1:17, 4:25])
    # bilinear interpolation reproduces a linear function at the quantised position
    lin = (0.5 * xx - 0.25 * yy + 3).astype(np.float32)
    rmap = _random_affine_map(20, 16, 12).astype(np.float32) * 0.6 + 4
    ix, fx = _quant(rmap[..., 0].astype(np.float64))
    iy, fy = _quant(rmap[..., 1].astype(np.float64))
    want = 0.5 * (ix + fx / 32) - 0.25 * (iy + fy / 32) + 3
    got = oracle32.remap(lin, rmap, interp=1, border=REPLICATE)
    inside = (ix >= 0) & (ix < 25) & (iy >= 0) & (iy < 19)
    np.testing.assert_allclose(got[inside], want[inside], atol=2e-5)


def bilinear_itab_brute(fx, fy):
    """initInterTab2D for INTER_LINEAR: float weights, cvRound(w * 32768) saturated to short,
    the deficit to 32768 added to tap (1,1) (the only tap of the ksize/2 scan for ksize 2)."""
    tx, ty = np.float32(fx / 32), np.float32(fy / 32)
    wf = [np.float32((1 - ty) * (1 - tx)), np.float32((1 - ty) * tx), np.float32(ty * (1 - tx)), np.float32(ty * tx)]
    wi = [int(np.clip(np.rint(np.float32(v) * np.float32(32768)), -32768, 32767)) for v in wf]
    wi[3] += 32768 - sum(wi)
    return wi


@pytest.mark.parametrize("border", [CONSTANT, REPLICATE])
def test_remap_u8_fixed_point(oracle64, border):
    r = rng(13)
    src = r.integers(0, 256, (18, 23)).astype(np.uint8)
    src[::4] = 255
    rmap = _random_affine_map(25, 19, 14)
    got = oracle64.remap_u8(src, rmap, border=border, cval=0)
    ix, fx = _quant(rmap[..., 0]); iy, fy = _quant(rmap[..., 1])
    want = np.zeros(rmap.shape[:2], np.int64)
    for y in range(rmap.shape[0]):
        for x in range(rmap.shape[1]):
            a, b = ix[y, x], iy[y, x]
            if border == CONSTANT and (a >= 23 or a + 1 < 0 or b >= 18 or b + 1 < 0):
                continue
            wi = bilinear_itab_brute(fx[y, x], fy[y, x])
            taps = [_tap(src.astype(np.int64), np.array(b + j), np.array(a + i), border, 0)
                    for j in (0, 1) for i in (0, 1)]
            s = sum(int(t) * wgt for t, wgt in zip(taps, wi))
            want[y, x] = min(max((s + (1 << 14)) >> 15, 0), 255)
    np.testing.assert_array_equal(got, want)
    # all-255 mask sampled strictly inside stays 255 (the (0,0) table quirk included)
    full = np.full((18, 23), 255, np.uint8)
    yy, xx = np.mgrid[0:10, 0:12].astype(np.float64)
    np.testing.assert_array_equal(oracle64.remap_u8(full, np.stack([xx + 3.3, yy + 2], -1)), 255)


# ---- morphology, statistics, linear algebra ----------------------------------------------------
@pytest.mark.parametrize("border,mode", [(CONSTANT, "constant"), (REPLICATE, "nearest")])
def test_erode_matches_scipy(oracle64, border, mode):
    src = (rng(15).uniform(size=(26, 31)) > 0.15).astype(np.uint8) * 255
    got = oracle64.erode(src, 5, border=border, cval=255)
    np.testing.assert_array_equal(got, ndimage.grey_erosion(src, size=(5, 5), mode=mode, cval=255))


def test_mean_stddev_matches_numpy(oracle64):
    r = rng(16)
    src = r.normal(3, 2, (40, 50))
    mask = (r.uniform(size=src.shape) > 0.3).astype(np.uint8) * 255
    m, s = oracle64.mean_stddev(src, mask)
    sel = src[mask > 0]
    assert abs(m - sel.mean()) < 1e-12 and abs(s - sel.std()) < 1e-10


@pytest.mark.parametrize("n", [2, 3, 4, 6])
def test_cholesky_invert_matches_numpy(oracle64, n):
    r = rng(20 + n)
    B = r.normal(size=(n, n))
    A = B @ B.T + n * np.eye(n)
    ok, inv = oracle64.cholesky_invert(A.copy())
    assert ok
    np.testing.assert_allclose(inv, np.linalg.inv(A), rtol=1e-10, atol=1e-12)


def test_cholesky_invert_failure_rules(oracle64):
    # n <= 3 fails only on a zero determinant; n > 3 fails on a non-positive pivot
    ok, inv = oracle64.cholesky_invert(np.array([[1.0, 2.0], [2.0, 4.0]]))
    assert not ok and not inv.any()
    ok, _ = oracle64.cholesky_invert(np.array([[1.0, 0.0], [0.0, -1.0]]))  # indefinite but invertible
    assert ok
    v = rng(30).normal(size=(6, 1))
    ok, _ = oracle64.cholesky_invert(v @ v.T)  # rank 1
    assert not ok


# ---- transforms -----------------------------------------------------------------------------------
def test_create_map_closed_forms(oracle64):
    w, h = 13, 9
    yy, xx = np.mgrid[0:h, 0:w].astype(np.float64)
    a = [1.01, -0.02, 3.5, 0.015, 0.99, -1.25]
    m = oracle64.create_map(oracle64.xform(3, a), w, h)
    np.testing.assert_allclose(m[..., 0], a[0] * xx + a[1] * yy + a[2], atol=1e-12)
    np.testing.assert_allclose(m[..., 1], a[3] * xx + a[4] * yy + a[5], atol=1e-12)
    m = oracle64.create_map(oracle64.xform(0, [2.5, -1.75]), w, h)
    np.testing.assert_allclose(m[..., 0], xx + 2.5); np.testing.assert_allclose(m[..., 1], yy - 1.75)
    # scaled euclidean about a centre: s R(theta) (x - C) + T
    th, s, T, Cc = 0.03, 1.02, (1.5, -2.0), (6.0, 4.0)
    t = oracle64.xform(2, [T[0], T[1], th, s], center=Cc)
    m = oracle64.create_map(t, w, h)
    u = s * (np.cos(th) * (xx - Cc[0]) - np.sin(th) * (yy - Cc[1])) + T[0]
    v = s * (np.sin(th) * (xx - Cc[0]) + np.cos(th) * (yy - Cc[1])) + T[1]
    np.testing.assert_allclose(m[..., 0], u, atol=1e-12); np.testing.assert_allclose(m[..., 1], v, atol=1e-12)
    M = oracle64.xform_matrix(t).reshape(2, 3)
    np.testing.assert_allclose(M[0, 0] * xx + M[0, 1] * yy + M[0, 2], u, atol=1e-12)
    np.testing.assert_allclose(M[1, 0] * xx + M[1, 1] * yy + M[1, 2], v, atol=1e-12)


def test_xform_eps_closed_forms(oracle64):
    w, h = 640, 480
    assert abs(oracle64.xform_eps(oracle64.xform(0), [3.0, 4.0], w, h) - 5.0) < 1e-12
    dp = [1e-4, -2e-4, 0.3, 3e-4, 1e-4, -0.4]
    want = np.sqrt((w * dp[0]) ** 2 + (h * dp[1]) ** 2 + dp[2] ** 2 + (w * dp[3]) ** 2 + (h * dp[4]) ** 2 + dp[5] ** 2)
    assert abs(oracle64.xform_eps(oracle64.xform(3), dp, w, h) - want) < 1e-12
    # euclidean with a fixed scale: ds is the transform's own scale (1), so eps >= max(W, H)
    e = oracle64.xform_eps(oracle64.xform(1), [0.0, 0.0, 0.0], w, h)
    assert abs(e - max(w, h)) < 1e-9
    dp4 = [0.1, 0.2, 1e-3, 2e-4]
    want = np.sqrt(0.1 ** 2 + 0.2 ** 2 + (w * np.sin(1e-3)) ** 2 + (h * np.sin(1e-3)) ** 2 + (w * 2e-4) ** 2)
    assert abs(oracle64.xform_eps(oracle64.xform(2), dp4, w, h) - want) < 1e-12


def test_xform_scale_rules(oracle64):
    t = oracle64.xform(3, [1.1, 0.1, 4.0, -0.1, 0.9, -6.0])
    oracle64.xform_scale(t, 0.5)
    assert list(t.p) == [1.1, 0.1, 2.0, -0.1, 0.9, -3.0]  # only the translation column scales
    t = oracle64.xform(1, [4.0, -2.0, 0.01], center=(10.0, 8.0))
    oracle64.xform_scale(t, 2.0)
    assert list(t.p)[:3] == [8.0, -4.0, 0.01] and list(t.C) == [20.0, 16.0]


# ---- ECC, correlation: known answers -------------------------------------------------------------
@pytest.mark.parametrize("bits", [64, 32])
def test_ecc_fa_align_recovers_known_translation(oracle64, oracle32, bits):
    ora = oracle64 if bits == 64 else oracle32
    w, h = 160, 128
    render = synth_np.blob_scene(w, h, seed=41)
    T = np.array([[1.0, 0.0, 2.37], [0.0, 1.0, -1.61]])
    t = ora.xform(0)
    ok, res = ora.ecc_fa_align(render([1, 0, 0, 0, 1, 0]), render(T), t, max_iterations=100, max_eps=1e-3)
    assert ok
    # positions are quantised to 1/32 px inside remap: the solution is known to that grain
    assert synth_np.max_displacement(ora.xform_matrix(t), synth_np.invert(T), w, h) < 0.02


@pytest.mark.parametrize("scale", [1.0, 1.5])
def test_ecc_fa_first_step_is_the_scaled_gauss_newton_step(oracle64, scale):
    """One forward-additive iteration from the identity, on a smooth scene rendered under a
    known small translation: to first order the Gauss-Newton step of a smooth image is the
    shift itself, so the first update lands at update_step_scale x the true correction (the
    inverse shift) -- a dropped, doubled or misplaced step scale, or a wrong sign, misses it by
    far more than the linearisation error.  The count is the reference's post-incremented
    `while (_num_iterations++ < _max_iterations)` (ecc2.cc:1295): a budget of one that does not
    converge reports 2."""
    w, h = 160, 128
    render = synth_np.blob_scene(w, h, seed=41)
    tx, ty = 0.23, -0.17
    T = np.array([[1.0, 0.0, tx], [0.0, 1.0, ty]])
    t = oracle64.xform(0)
    ok, res = oracle64.ecc_fa_align(render([1, 0, 0, 0, 1, 0]), render(T), t, max_iterations=1, max_eps=1e-12,
                                    update_step_scale=scale)
    assert ok and res.num_iterations == 2
    step = oracle64.xform_matrix(t).reshape(2, 3)[:, 2]
    want = -scale * np.array([tx, ty])
    assert np.linalg.norm(step - want) < 0.02 * np.linalg.norm(want), (step, want)


def test_ecc_fa_align_identity_converges_at_once(oracle64):
    ref, _, _ = synth_np.frames(1, 128, 96, radius=40.0, seed=42)
    t = oracle64.xform(3)
    ok, res = oracle64.ecc_fa_align(ref, ref.copy(), t, max_iterations=20, max_eps=1e-3)
    assert ok and res.num_iterations == 1
    np.testing.assert_allclose(list(t.p), [1, 0, 0, 0, 1, 0], atol=1e-9)


@pytest.mark.parametrize("motion", [0, 1, 3])
@pytest.mark.parametrize("bits", [64, 32])
def test_registration_recovers_known_warp(oracle64, oracle32, motion, bits):
    """Pyramid registration (translation-first + main pass) of a band-limited scene
    rendered under a known warp T: the registration is T^-1 within 0.04 px."""
    from oracle.oracle import reg_options
    ora = oracle64 if bits == 64 else oracle32
    w, h = 256, 192
    rnd = np.random.default_rng(100 + motion)
    T = synth_np.random_affine(rnd, w, h, 4.0, 0.0 if motion == 0 else 1.0,
                               0.005 if motion == 3 else 0.0, 0.003 if motion == 3 else 0.0)
    render = synth_np.blob_scene(w, h, seed=7 + motion)
    r = ora.registration(reg_options(motion_type=motion, scale=1.0, eps=0.01, ecch_max_level=3, max_iterations=80))
    r.setup_reference_frame(render([1, 0, 0, 0, 1, 0]))
    st = r.register_frame(render(T))
    assert st.ok and st.code == 0
    assert synth_np.max_displacement(st.matrix, synth_np.invert(T), w, h) < 0.04


def test_compute_correlation_closed_forms(oracle64):
    r = rng(50)
    ref = ndimage.gaussian_filter(r.normal(size=(48, 64)), 2)
    from oracle.oracle import affine_map
    ident = affine_map([1, 0, 0, 0, 1, 0], 64, 48)
    assert abs(oracle64.compute_correlation(ref.copy(), ref, ident) - 1) < 1e-12
    assert abs(oracle64.compute_correlation(2.0 * ref + 3.0, ref, ident) - 1) < 1e-12
    assert abs(oracle64.compute_correlation(-ref, ref, ident) + 1) < 1e-12
    other = ndimage.gaussian_filter(r.normal(size=(48, 64)), 2) + 0.5 * ref
    assert abs(oracle64.compute_correlation(other, ref, ident) - np.corrcoef(other.ravel(), ref.ravel())[0, 1]) < 1e-12


# ---- accumulation, weights -------------------------------------------------------------------------
def test_weighted_average_update_is_the_weighted_mean(oracle64, oracle32):
    r = rng(60)
    frames = [r.normal(size=(12, 14)) for _ in range(7)]
    wts = [r.uniform(0.1, 2, size=(12, 14)) * (r.uniform(size=(12, 14)) > 0.2) for _ in range(7)]
    want = sum(w * f for w, f in zip(wts, frames)) / np.maximum(sum(wts), 1e-300)
    for ora, tol in ((oracle64, 1e-12), (oracle32, 2e-5)):
        acc = np.zeros((12, 14), ora.dt); accw = np.zeros((12, 14), ora.dt)
        for f, w in zip(frames, wts):
            ora.weighted_average_update(f, acc, accw, weights=w)
        have = sum(wts) > 0
        np.testing.assert_allclose(acc[have], want[have], atol=tol)
        np.testing.assert_allclose(accw, sum(wts), rtol=tol)


def test_sharpness_mean_squared_laplacian(oracle64):
    src = rng(61).normal(size=(21, 27))
    p = np.pad(src, 1, mode="edge")
    lap = p[:-2, 1:-1] + p[2:, 1:-1] + p[1:-1, :-2] + p[1:-1, 2:] - 4 * src
    assert abs(oracle64.sharpness(src) - (lap ** 2).mean()) < 1e-12


def test_base_remap_mask_composition(oracle64):
    """base_remap's mask = erode5(CONSTANT 255)(remap_u8(mask or 255) >= 255) -- composed
    from the pinned primitives (c_frame_registration.cc:1319-1338)."""
    r = rng(62)
    src = r.normal(size=(40, 52))
    smask = (r.uniform(size=src.shape) > 0.05).astype(np.uint8) * 255
    rmap = _random_affine_map(50, 38, 63)
    dst, dmask = oracle64.base_remap(src, rmap, smask, interp=1, border=REFLECT101)
    np.testing.assert_allclose(dst, remap_linear_brute(src, rmap, REFLECT101, 0.0), atol=1e-12)
    m = oracle64.remap_u8(smask, rmap, border=CONSTANT, cval=0)
    m = np.where(m >= 255, 255, 0).astype(np.uint8)
    np.testing.assert_array_equal(dmask, ndimage.grey_erosion(m, size=(5, 5), mode="constant", cval=255))


# ---- c_bayer_average ------------------------------------------------------------------------------
B, G, R = 0, 1, 2
# the 2x2 cell layouts of generate_bayer_pattern_mask (c_frame_accumulation.cc:1267-1334)
BAYER_CELLS = {8: [[R, G], [G, B]], 9: [[G, R], [B, G]], 10: [[G, B], [R, G]], 11: [[B, G], [G, R]]}


@pytest.mark.parametrize("colorid", [8, 9, 10, 11])
def test_bayer_pattern_tables(oracle32, colorid):
    for y in range(4):
        for x in range(6):
            assert oracle32.bayer_channel(colorid, x, y) == BAYER_CELLS[colorid][y & 1][x & 1]


def _bayer_brute(frames, colorid, rmaps=None, masks=None, weights=None):
    """float64 restatement of _bayer_accumulate (c_frame_accumulation.cc:988-1124), one add per frame."""
    h, w = frames[0].shape
    acc = np.zeros((h, w, 3)); cntr = np.zeros((h, w, 3)); cnt = np.zeros((h, w, 3), np.int64)
    cell = np.array(BAYER_CELLS[colorid])
    for k, f in enumerate(frames):
        for y in range(h):
            for x in range(w):
                if masks is not None and not masks[k][y, x]:
                    continue
                wt = 1.0 if weights is None else float(weights[k][y, x])
                if rmaps is None:
                    c = cell[y & 1, x & 1]
                    acc[y, x, c] += f[y, x] * wt
                    cntr[y, x, c] += wt
                    cnt[y, x, c] += wt != 0
                    continue
                px, py = rmaps[k][y, x]
                sx, sy = int(px), int(py)  # truncation toward zero
                if not (0 <= sx < w - 1 and 0 <= sy < h - 1):
                    continue
                ax, ay, bx, by = sx + 1 - px, sy + 1 - py, px - sx, py - sy
                for (i, j, s) in ((0, 0, ax * ay), (1, 0, bx * ay), (0, 1, ax * by), (1, 1, bx * by)):
                    c = cell[(sy + j) & 1, (sx + i) & 1]
                    acc[y, x, c] += f[sy + j, sx + i] * s * wt
                    cntr[y, x, c] += s * wt
                    cnt[y, x, c] += s * wt != 0
    return acc, cntr, cnt


@pytest.mark.parametrize("colorid", [8, 11])
@pytest.mark.parametrize("wmode", ["none", "mask", "weights"])
def test_bayer_accumulate_without_remap(oracle64, colorid, wmode):
    r = rng(70)
    frames = [r.uniform(0, 1, (10, 14)) for _ in range(4)]
    masks = [(r.uniform(size=(10, 14)) > 0.3).astype(np.uint8) * 255 for _ in range(4)] if wmode == "mask" else None
    wts = [r.uniform(0, 2, (10, 14)) * (r.uniform(size=(10, 14)) > 0.2) for _ in range(4)] if wmode == "weights" else None
    st = oracle64.bayer_state(14, 10)
    for k, f in enumerate(frames):
        oracle64.bayer_accumulate(f, st, colorid, mask=None if masks is None else masks[k],
                                  weights=None if wts is None else wts[k])
    acc, cntr, cnt = _bayer_brute(frames, colorid, masks=masks, weights=wts)
    np.testing.assert_allclose(st[0], acc, atol=1e-12)
    np.testing.assert_allclose(st[1], cntr, atol=1e-12)
    np.testing.assert_array_equal(st[2], cnt)


@pytest.mark.parametrize("colorid", [8, 9, 10, 11])
def test_bayer_accumulate_remap_brute_force(oracle64, colorid):
    r = rng(71)
    frames = [r.uniform(0, 1, (12, 16)) for _ in range(3)]
    yy, xx = np.mgrid[0:12, 0:16].astype(np.float64)
    rmaps = [np.stack([xx + r.uniform(-2.5, 2.5), yy + r.uniform(-2.5, 2.5)], -1) for _ in range(3)]
    rmaps[1][..., 0] += 0.01 * yy  # a sheared one
    st = oracle64.bayer_state(16, 12)
    for f, m in zip(frames, rmaps):
        oracle64.bayer_accumulate(f, st, colorid, rmap=m)
    acc, cntr, cnt = _bayer_brute(frames, colorid, rmaps=rmaps)
    np.testing.assert_allclose(st[0], acc, atol=1e-12)
    np.testing.assert_allclose(st[1], cntr, atol=1e-12)
    np.testing.assert_array_equal(st[2], cnt)
    # the four side weights are a partition of unity wherever a frame landed
    landed = cnt.sum(-1) > 0
    assert np.all(np.abs(st[1].sum(-1)[landed] - np.round(st[1].sum(-1)[landed])) < 1e-12)


def test_bayer_even_shift_keeps_the_phase(oracle32):
    """An even integer shift reads every destination pixel from a source pixel of
    the same Bayer colour: one channel per pixel, exact copies, integer counts."""
    r = rng(72)
    f = r.uniform(0, 1, (12, 16)).astype(np.float32)
    yy, xx = np.mgrid[0:12, 0:16].astype(np.float32)
    st = oracle32.bayer_state(16, 12)
    oracle32.bayer_accumulate(f, st, 8, rmap=np.stack([xx + 2, yy - 2], -1))
    avg, mask = oracle32.bayer_compute(st)
    for y in range(2, 12):
        for x in range(0, 13):
            c = BAYER_CELLS[8][y & 1][x & 1]
            assert st[2][y, x].tolist() == [int(c == k) for k in range(3)]
            assert avg[y, x, c] == f[y - 2, x + 2]
    assert mask[2:, :13].all() and not mask[:2].any()


# ---- FFT phase correlation (oracle/phasecorr.py; the reference has none) --------------------------
def test_phasecorr_circular_shift_is_exact():
    from oracle import phasecorr as pc
    r = rng(90)
    ref = r.normal(size=(64, 128))
    for dx, dy in ((5, -3), (-17, 9), (0, 0), (63, -31)):
        frame = np.roll(np.roll(ref, dy, axis=0), dx, axis=1)  # frame(x) = ref(x - d)
        tx, ty, peak = pc.estimate(pc.spectrum(ref, window=False), frame, window=False)
        # an exact circular shift gives the weighted delta: a symmetric gaussian peak at the
        # shift (no sub-pixel offset) of height mean(weight)
        assert abs(tx - dx) < 1e-9 and abs(ty - dy) < 1e-9
        assert abs(peak - pc.peak_weight(64, 128).mean()) < 1e-12


@pytest.mark.parametrize("binning,tol", [(1, 0.1), (2, 0.3)])
def test_phasecorr_recovers_known_subpixel_shift(binning, tol):
    """ref(x) = frame(x + T) for a textured scene (600 blobs, sigma 1.2-3 px, 1% noise)
    rendered under a known shift; the tolerance is the parabolic-fit bias (binned
    pixels are 2 px)."""
    from oracle import phasecorr as pc
    w, h = 256, 192
    render = synth_np.blob_scene(w, h, seed=91, n=600, smin=1.2, smax=3.0)
    r = rng(92)
    ref = render([1, 0, 0, 0, 1, 0]) + 0.01 * r.normal(size=(h, w))
    spec = pc.spectrum(ref, binning)
    for T in ((3.3, -1.7), (-7.6, 4.25), (0.4, 0.0)):
        frame = render([1, 0, -T[0], 0, 1, -T[1]]) + 0.01 * r.normal(size=(h, w))  # frame(x) = scene(x - T)
        tx, ty, peak = pc.estimate(spec, frame, binning)
        # frame(x + T) = scene(x) = ref(x): the estimate is T
        assert abs(tx - T[0]) < tol and abs(ty - T[1]) < tol, (T, tx, ty)
        assert 0.1 < peak <= 1.0


# ---- focus stacking: lpg, laplacian pyramids, fusion (oracle/focus.py) -----------------------------
def _lpg_brute(g, k, p):
    """float64 loops over lpg.cc:63-127 (interior), then the edge replication and pow."""
    g = g.astype(np.float64)
    h, w = g.shape
    alpha = k / (k + 1) * 25.0 ** 2
    beta = 1 / (k + 1) * (100 / 36) ** 2
    m = np.zeros((h, w))
    for y in range(2, h - 2):
        r = [g[y + d] for d in (-2, -1, 0, 1, 2)]
        for x in range(2, w - 2):
            gx = (sum(c * r[j][x + 2] for j, c in enumerate((1, 2, 4, 2, 1)))
                  - sum(c * r[j][x - 2] for j, c in enumerate((1, 2, 4, 2, 1)))
                  + 2 * ((r[1][x + 1] + 2 * r[2][x + 1] + r[3][x + 1]) - (r[1][x - 1] + 2 * r[2][x - 1] + r[3][x - 1])))
            gy = (sum(c * r[4][x + d] for d, c in zip((-2, -1, 0, 1, 2), (1, 2, 4, 2, 1)))
                  - sum(c * r[0][x + d] for d, c in zip((-2, -1, 0, 1, 2), (1, 2, 4, 2, 1)))
                  + 2 * ((r[3][x - 1] + 2 * r[3][x] + r[3][x + 1]) - (r[1][x - 1] + 2 * r[1][x] + r[1][x + 1])))
            lap = (16 * r[2][x] - 2 * (r[1][x] + r[3][x] + r[2][x - 1] + r[2][x + 1])
                   - (r[1][x - 1] + r[1][x + 1] + r[3][x - 1] + r[3][x + 1]) - (r[0][x] + r[4][x] + r[2][x - 2] + r[2][x + 2]))
            m[y, x] = alpha * lap * lap + beta * (gx * gx + gy * gy) + 1e-9
    m[:, :2] = m[:, 2:3]; m[:, -2:] = m[:, -3:-2]
    m[:2] = m[2]; m[-2:] = m[-3]
    return m ** p


def test_lpg_brute_force_and_closed_forms():
    from oracle import focus
    r = rng(100)
    g = r.uniform(0, 1, (13, 17)).astype(np.float32)
    np.testing.assert_allclose(focus.lpg(g, 6.0, 2.0), _lpg_brute(g, 6.0, 2.0), rtol=2e-5)
    np.testing.assert_allclose(focus.lpg(g, 3.0, 1.0), _lpg_brute(g, 3.0, 1.0), rtol=2e-5)
    # constant image: only eps remains; a ramp a*x: no laplacian, gx = 56 a
    assert np.all(focus.lpg(np.full((9, 9), 0.3, np.float32), 6, 1) == np.float32(1e-9))
    yy, xx = np.mgrid[0:12, 0:15].astype(np.float32)
    m = focus.lpg(np.float32(0.02) * xx, 6.0, 1.0)
    np.testing.assert_allclose(m, (100 / 36) ** 2 / 7 * (56 * 0.02) ** 2, rtol=1e-5)


def test_laplacian_pyramid_reconstructs(oracle32):
    from oracle import focus
    img = rng(101).uniform(0, 1, (70, 90)).astype(np.float32)
    sizes = focus.level_sizes(90, 70)
    assert sizes == [(90, 70), (45, 35), (23, 18), (12, 9)]  # stop when the next smaller side <= 8
    pyr = focus.build_laplacian(oracle32, img, sizes)
    np.testing.assert_allclose(focus.reconstruct(oracle32, pyr), img, atol=2e-6)
    assert focus.level_sizes(90, 70, max_levels=2) == [(90, 70), (45, 35)]


@pytest.mark.parametrize("policy", [0, 1, 2])
def test_focus_of_identical_frames_is_the_frame(oracle32, policy):
    from oracle import focus
    img = rng(102).uniform(0.2, 1, (48, 64)).astype(np.float32)
    fs = focus.FocusStack(oracle32, policy)
    for _ in range(3):
        fs.add([img])
    np.testing.assert_allclose(fs.compute()[0], img, atol=1e-4)


@pytest.mark.parametrize("policy", [0, 1, 2])
def test_focus_picks_the_sharp_halves(oracle32, policy):
    """A is sharp on the left and blurred on the right, B the opposite: the fused
    image is closer to the sharp scene than either input (a min-selection or a
    swapped weight fails this)."""
    from oracle import focus
    r = rng(103)
    sharp = ndimage.gaussian_filter(r.uniform(0, 1, (64, 96)), 0.7).astype(np.float32)
    blur = ndimage.gaussian_filter(sharp, 3.0).astype(np.float32)
    A = sharp.copy(); A[:, 48:] = blur[:, 48:]
    B = sharp.copy(); B[:, :48] = blur[:, :48]
    fs = focus.FocusStack(oracle32, policy, lpg_sigma=2.0)
    fs.add([A]); fs.add([B])
    out = fs.compute()[0]
    err = lambda x: float(np.sqrt(np.mean((x[8:-8, 8:-8] - sharp[8:-8, 8:-8]) ** 2)))
    assert err(out) < 0.6 * min(err(A), err(B)), (err(out), err(A), err(B))


# ---- local-variance sharpness map (oracle/lvm.py) ---------------------------------------------------
def test_lvm_morph_gradient_matches_scipy():
    from oracle import lvm
    m = rng(110).normal(size=(21, 30)).astype(np.float32)
    for r in (1, 2):
        k = 2 * r + 1
        want = ndimage.grey_dilation(m, size=(k, k), mode="nearest") - ndimage.grey_erosion(m, size=(k, k), mode="nearest")
        np.testing.assert_array_equal(lvm.morph_gradient(m, r), want)


def test_lvm_resize_linear_closed_forms():
    from oracle import lvm
    # constants stay constant; a ramp is reproduced at the half-pixel-centre positions
    np.testing.assert_array_equal(lvm.resize_linear(np.full((7, 9), 2.5, np.float32), 17, 13), 2.5)
    src = np.tile(np.arange(9, dtype=np.float32) * 2, (7, 1))
    out = lvm.resize_linear(src, 18, 7)
    xs = (np.arange(18) + 0.5) * 0.5 - 0.5
    want = 2 * np.clip(xs, 0, 8)
    np.testing.assert_allclose(out[3], want, atol=1e-6)


def test_lvm_closed_forms(oracle32):
    from oracle import lvm
    # a flat frame has no gradient: Q = 0 and a zero map
    m, q = lvm.local_variance_map(oracle32, np.full((32, 40), 0.4, np.float32))
    assert q == 0 and not m.any()
    # a vertical step: G is the step height on the 2 columns around the edge (3x3 gradient)
    img = np.zeros((24, 32), np.float32); img[:, 16:] = 0.5
    m, q = lvm.local_variance_map(oracle32, img, kradius=1, dscale=0)
    G = np.zeros_like(img); G[:, 15:17] = 0.5
    W = float(np.abs(G).sum())
    assert abs(q - 8000.0 / W * float((G.astype(np.float64) ** 4).sum())) < 1e-9 * q
    np.testing.assert_allclose(m, G ** 3 * 8000 + np.float32(0.05 * q), rtol=1e-6)

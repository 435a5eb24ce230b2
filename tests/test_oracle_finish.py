"""Pins of the stack-finish oracle (oracle/finish.py): linear-interpolation inpainting
(linear_interpolation_inpaint.cc), unsharp_mask (unsharp_mask.cc) and the frame / mask / remap
upscales of stacking at a larger size by closed forms and by the pinned FP64 pyrUp."""
import math

import numpy as np
import pytest

from oracle import finish as F


def ramp(h, w, cn=1):
    y, x = np.mgrid[0:h, 0:w].astype(np.float64)
    planes = [0.25 + 0.01 * x - 0.02 * y + 0.5 * c for c in range(cn)]
    img = np.stack(planes, -1) if cn > 1 else planes[0]
    return img.astype(np.float32)


@pytest.mark.parametrize("cn", [1, 3])
def test_inpaint_row_formula(cn):
    """A one-row image (no column data): a run of holes [start, end) between the valid
    pixels s = start - 1 and e = end takes sv + (x - s) * (ev - sv) / (end - start), the
    reference's slope (linear_interpolation_inpaint.cc:50-63), which reaches ev at the last
    hole rather than at e."""
    img = ramp(1, 12, cn)
    mask = np.full((1, 12), 255, np.uint8)
    mask[0, 3:7] = 0
    holed = img.copy()
    holed[mask == 0] = 9.0
    out = F.inpaint_linear(holed, mask)
    sv, ev = img[0, 2], img[0, 7]
    kk = (ev - sv) * (np.float32(1) / np.float32(4))
    for x in range(3, 7):
        np.testing.assert_array_equal(out[0, x], sv + np.float32(x - 2) * kk)
    np.testing.assert_array_equal(out[0, 6], ev)


def test_inpaint_edges_copy_and_constant():
    img = np.full((6, 8), 3.5, np.float32)
    mask = np.full((6, 8), 255, np.uint8)
    mask[:, :2] = 0  # holes touching the left edge: the first valid pixel of each row
    mask[0, :] = 0   # a whole row: filled from the columns
    holed = img.copy()
    holed[mask == 0] = -1
    np.testing.assert_array_equal(F.inpaint_linear(holed, mask), img)
    # no data at all: nothing is filled, and the loop ends
    empty = np.zeros((4, 5), np.uint8)
    src = np.arange(20, dtype=np.float32).reshape(4, 5)
    np.testing.assert_array_equal(F.inpaint_linear(src, empty), src)


def test_inpaint_single_hole_closed_form():
    """One hole at (2, 2) with row neighbours a, b and column neighbours c, d: the run slope
    (b - a) / 1 makes the row fill b and the column fill d, both at distance max(1, 1) = 1,
    so the hole takes (1 / 2) * (b + d)."""
    img = np.zeros((5, 5), np.float32)
    img[2, 1], img[2, 3], img[1, 2], img[3, 2] = 1.0, 3.0, 10.0, 20.0
    mask = np.full((5, 5), 255, np.uint8)
    mask[2, 2] = 0
    out = F.inpaint_linear(img, mask)
    assert out[2, 2] == np.float32(0.5) * (np.float32(3.0) + np.float32(20.0))


def test_lpass_plan():
    assert F.lpass_plan(1.5, 100, 100) == (0, 0, 1.5)
    lv, ci, d = F.lpass_plan(3.0, 100, 100)
    assert (lv, ci) == (1, 1) and math.isclose(d, math.sqrt(7) / 2)
    lv, ci, d = F.lpass_plan(10.0, 400, 300)
    assert (lv, ci) == (3, 21) and math.isclose(d, math.sqrt(58) / 8)
    assert F.lpass_plan(10.0, 6, 6)[0] == 2  # limited by the image size (imax = 2)


@pytest.mark.parametrize("sigma", [1.0, 3.0, 8.0])
def test_unsharp_of_a_constant_is_the_constant(oracle32, sigma):
    img = np.full((40, 50), 0.25, np.float32)
    out = F.unsharp_mask(oracle32, img, sigma, 0.6)
    np.testing.assert_allclose(out, img, atol=1e-6)
    np.testing.assert_array_equal(F.unsharp_mask(oracle32, img, 0, 0.6), img)


# ---- average_pyramid_inpaint: closed forms worked from the reference's rules ----------------------

def _holes(h, w, pts):
    m = np.full((h, w), 255, np.uint8)
    for y, x in pts:
        m[y, x] = 0
    return m


def test_average_pyramid_single_holes_closed_forms():
    """Isolated interior holes: the odd-odd sampling of downstrike_even and the odd-odd placement of
    upject_even give an even-even hole the mean of its 4 diagonal neighbours, an odd-odd hole the
    mean of its 8 neighbours, and a mixed hole the mean of its 2 neighbours along the odd axis."""
    from oracle import finish

    r = np.random.default_rng(3)
    img = r.uniform(0, 1, (11, 13)).astype(np.float32)
    pts = [(2, 2), (3, 7), (6, 3), (5, 10), (8, 8)]
    out, om = finish.average_pyramid_inpaint(img, _holes(11, 13, pts))
    I = img.astype(np.float64)
    assert out[2, 2] == np.float32((I[1, 1] + I[1, 3] + I[3, 1] + I[3, 3])) * np.float32(0.25)
    assert out[3, 7] == np.float32(I[2:5, 6:9].sum() - I[3, 7]) * np.float32(0.125)
    assert out[6, 3] == np.float32(I[5, 3] + I[7, 3]) * np.float32(0.5)
    assert out[5, 10] == np.float32(I[5, 9] + I[5, 11]) * np.float32(0.5)
    assert out[8, 8] == np.float32((I[7, 7] + I[7, 9] + I[9, 7] + I[9, 9])) * np.float32(0.25)
    keep = _holes(11, 13, pts) != 0
    np.testing.assert_array_equal(out[keep], img[keep])
    assert (om == 255).all()


def test_average_pyramid_replicate_corners():
    """BORDER_REPLICATE: a hole at (0, 0) takes pixel (1, 1); a hole at the last odd-odd corner
    takes its replicated 3x3 average, (a + 2b + 2c) / 5."""
    from oracle import finish

    r = np.random.default_rng(4)
    img = r.uniform(0, 1, (8, 10)).astype(np.float32)
    out, _ = finish.average_pyramid_inpaint(img, _holes(8, 10, [(0, 0), (7, 9)]))
    I = img.astype(np.float64)
    assert out[0, 0] == img[1, 1]
    s = np.float32(I[6, 8] + 2 * I[6, 9] + 2 * I[7, 8])
    assert out[7, 9] == s * (np.float32(1.0) / np.float32(5.0))


def test_average_pyramid_constant_and_copies():
    from oracle import finish

    # a power-of-two constant survives every scaling k * (1 / k) of the averages, at every level
    img = np.full((40, 52), 4.0, np.float32)
    m = np.full((40, 52), 255, np.uint8)
    m[5:33, 8:47] = 0
    out, om = finish.average_pyramid_inpaint(img, m)
    assert (out == 4.0).all() and (om == 255).all()
    # nothing to fill: a copy; no levels: the masked copy and the mask unchanged
    out, om = finish.average_pyramid_inpaint(img, np.full_like(m, 255))
    assert (out == img).all() and (om == 255).all()
    out, om = finish.average_pyramid_inpaint(img, m, max_levels=0)
    assert (out[m == 0] == 0).all() and (out[m != 0] == 4.0).all()
    np.testing.assert_array_equal(om, m)


def test_upscale_map_closed_forms(oracle32):
    """The identity remap upscaled: x1.5 / x3 reproduce the positions (X + 0.5) / s - 0.5 inside
    (clamped at the borders), x2 (pyrUp) reproduces X / 2 at every interior output pixel."""
    from oracle import finish

    h, w = 12, 16
    yy, xx = np.mgrid[0:h, 0:w].astype(np.float32)
    ident = np.stack([xx, yy], -1)
    for opt, s in ((finish.UPSCALE_X15, 1.5), (finish.UPSCALE_X3, 3.0)):
        up = finish.upscale_map(oracle32, ident, opt)
        dw, dh = finish.upscale_size(opt, w, h)
        assert up.shape == (dh, dw, 2)
        want = np.clip((np.arange(dw) + 0.5) / s - 0.5, 0, w - 1)
        np.testing.assert_allclose(up[dh // 2, :, 0], want, atol=2e-6)
    up = finish.upscale_map(oracle32, ident, finish.UPSCALE_X2)
    assert up.shape == (2 * h, 2 * w, 2)
    np.testing.assert_array_equal(up[4:-4, 2:-2, 0], np.tile(np.arange(2, 2 * w - 2) / 2, (2 * h - 8, 1)))
    np.testing.assert_array_equal(up[2:-2, 4:-4, 1], np.tile((np.arange(2, 2 * h - 2) / 2)[:, None], (1, 2 * w - 8)))
    assert finish.upscale_size(finish.UPSCALE_X15, 33, 21) == (49, 31)


@pytest.mark.parametrize("shape", [(8, 12), (7, 9), (1, 5), (5, 1)])
def test_upscale_u8_constants(shape):
    """Every constant 8-bit image stays that constant through the three 8-bit upscales (the
    rounding terms of the fixed-point passes: + 2 >> 2 after the >> 4 / >> 16 products, + 2^15
    >> 16, + 32 >> 6)."""
    h, w = shape
    for c in range(256):
        m = np.full(shape, c, np.uint8)
        assert np.all(F.resize_linear_u8(m, w * 3 // 2 or 1, h * 3 // 2 or 1) == c)
        assert np.all(F.resize_linear_exact_u8(m, 3 * w, 3 * h) == c)
        assert np.all(F.pyr_up_u8(m) == c)


def test_upscale_u8_linear_tracks_the_float_resize():
    """The fixed-point INTER_LINEAR of random 8-bit data is within 1 of the float32 resize of the
    same data (same taps, weights exact to 2^-11): a swapped weight or a wrong tap is off by far
    more."""
    from oracle import lvm

    r = np.random.default_rng(4)
    for (h, w), (dh, dw) in (((20, 30), (30, 45)), ((13, 17), (19, 25)), ((6, 4), (9, 6))):
        m = r.integers(0, 256, (h, w), dtype=np.uint8)
        got = F.resize_linear_u8(m, dw, dh).astype(np.float64)
        ref = lvm.resize_linear(m.astype(np.float32), dw, dh).astype(np.float64)
        assert np.abs(got - ref).max() <= 1.0


def test_upscale_u8_exact_x3_impulse():
    """INTER_LINEAR_EXACT x3 of a single 255 sample: the 8-bit weights of 1/3 and 2/3 are 85 and
    171, so the row through it reads ... 0, 85, 170, 255, 170, 85, 0 ... rounded half up
    (255 * 171 / 256 = 170.3, 255 * 85 / 256 = 84.7), centred on output 3j + 1."""
    m = np.zeros((5, 7), np.uint8)
    m[2, 3] = 255
    out = F.resize_linear_exact_u8(m, 21, 15)
    np.testing.assert_array_equal(out[7, 7:14], [0, 85, 170, 255, 170, 85, 0])
    np.testing.assert_array_equal(out[4:11, 10], [0, 85, 170, 255, 170, 85, 0])
    # left / top edges replicate the first sample, the right / bottom ones the last
    e = np.zeros((3, 4), np.uint8)
    e[:, 0] = 200
    e[:, 3] = 90
    out = F.resize_linear_exact_u8(e, 12, 9)
    assert np.all(out[:, :2] == 200) and np.all(out[:, -2:] == 90)


def test_upscale_u8_pyr_up_is_the_rounded_float_pyr_up(oracle64):
    """The 8-bit pyrUp is the FP64 pyrUp (pinned by brute force) of the same samples rounded half
    up: the integer sums are exact in both."""
    r = np.random.default_rng(6)
    for shape in ((9, 13), (2, 2), (1, 6), (6, 1), (16, 20)):
        m = r.integers(0, 256, shape, dtype=np.uint8)
        want = np.floor(oracle64.pyr_up(m.astype(np.float64)) + 0.5)
        np.testing.assert_array_equal(F.pyr_up_u8(m), want)


@pytest.mark.parametrize("option", [F.UPSCALE_X15, F.UPSCALE_X2])
def test_upscale_image_binary_mask_support(oracle64, option):
    """A 0 / 255 mask upscaled and compared with 255 keeps exactly the output pixels whose
    taps (nonzero weights) are all valid: the upscale of the indicator of the holes is 0 there
    (weights of x1.5 on even sizes are >= 1/6, so no tap is negligible)."""
    from oracle import lvm

    r = np.random.default_rng(8)
    h, w = 16, 22
    mask = np.where(r.random((h, w)) < 0.08, 0, 255).astype(np.uint8)
    img = r.random((h, w, 3)).astype(np.float32)
    dst, dm = F.upscale_image(oracle64, img, mask, option)
    dw, dh = F.upscale_size(option, w, h)
    assert dst.shape == (dh, dw, 3) and dm.shape == (dh, dw)
    holes = (mask == 0).astype(np.float32)
    spread = oracle64.pyr_up(holes.astype(np.float64)) if option == F.UPSCALE_X2 else \
        lvm.resize_linear(holes, dw, dh)
    np.testing.assert_array_equal(dm == 255, spread == 0)
    assert set(np.unique(dm)) <= {0, 255}
    # no mask: the image alone; none: copies
    d2, m2 = F.upscale_image(oracle64, img, None, option)
    assert m2 is None and np.array_equal(d2, dst)
    d3, m3 = F.upscale_image(oracle64, img, mask, F.UPSCALE_NONE)
    assert np.array_equal(d3, img) and np.array_equal(m3, mask)


def test_ecc_normalize_closed_forms(oracle32):
    """ecc_normalize removes the local mean: a constant image gives exactly 0 (pyrDown and pyrUp
    weights sum to 1 in powers of two), a linear ramp gives 0 away from the borders (both
    filters reproduce linear functions there), and pixels off the mask are 0."""
    from oracle import eccnorm

    c = np.full((48, 64), 0.375, np.float32)
    for lvl in (1, 2, 3):
        assert not eccnorm.ecc_normalize(oracle32, c, lvl=lvl).any()
    yy, xx = np.mgrid[0:48, 0:64].astype(np.float32)
    ramp = (0.01 * xx + 0.02 * yy).astype(np.float32)
    d = eccnorm.ecc_normalize(oracle32, ramp, lvl=2)
    assert np.abs(d[12:-12, 12:-12]).max() < 1e-5
    mask = np.full((48, 64), 255, np.uint8)
    mask[10:20, 5:9] = 0
    mask[30, 40] = 128
    r = np.random.default_rng(2).random((48, 64)).astype(np.float32)
    dm = eccnorm.ecc_normalize(oracle32, r, mask, 1)
    assert not dm[mask != 255].any()
    np.testing.assert_array_equal(dm[mask == 255], eccnorm.ecc_normalize(oracle32, r, None, 1)[mask == 255])

"""Pins of the stack-finish oracle (oracle/finish.py): linear-interpolation inpainting
(linear_interpolation_inpaint.cc) and unsharp_mask (unsharp_mask.cc) by closed forms."""
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

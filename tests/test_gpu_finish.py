"""GPU stack finish against the oracle (oracle/finish.py), through the C ABI:
linear-interpolation inpainting (linear_interpolation_inpaint.cc) and unsharp_mask
(unsharp_mask.cc), bit for bit."""
import numpy as np
import pytest
from scipy import ndimage

from oracle import finish as F

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def S():
    import torch  # noqa: F401

    import serstacker_amd as S

    S._lib.load()
    return S


def holes(h, w, seed, frac=0.08):
    rng = np.random.default_rng(seed)
    m = ndimage.binary_dilation(rng.random((h, w)) < frac * 0.3, iterations=2)
    m[:, :3] |= rng.random((h, 1)) < 0.3  # runs touching the left edge
    m[-2:, :] |= rng.random((1, w)) < 0.3  # and the bottom
    return np.where(m, 0, 255).astype(np.uint8)


@pytest.mark.parametrize("cn", [1, 3])
@pytest.mark.parametrize("shape", [(37, 53), (64, 70), (1, 40), (40, 1)])
def test_inpaint_linear_bitexact(S, cn, shape):
    import torch

    h, w = shape
    rng = np.random.default_rng(h * 100 + w + cn)
    img = rng.random((h, w, cn) if cn > 1 else (h, w)).astype(np.float32)
    mask = holes(h, w, h + w)
    want = F.inpaint_linear(img, mask)
    got = S.inpaint_linear(img, mask)
    np.testing.assert_array_equal(got, want)
    t = torch.from_numpy(img.copy()).cuda()
    S.inpaint_linear(t, torch.from_numpy(mask).cuda())
    np.testing.assert_array_equal(t.cpu().numpy(), want)


def test_inpaint_no_holes_and_no_data(S):
    img = np.random.default_rng(1).random((8, 9)).astype(np.float32)
    np.testing.assert_array_equal(S.inpaint_linear(img, np.full((8, 9), 255, np.uint8)), img)
    np.testing.assert_array_equal(S.inpaint_linear(img, np.zeros((8, 9), np.uint8)), img)


@pytest.mark.parametrize("sigma", [0.8, 2.0, 3.0, 6.5])
@pytest.mark.parametrize("cn", [1, 3])
def test_unsharp_mask_bitexact(S, oracle32, sigma, cn):
    import torch

    rng = np.random.default_rng(int(sigma * 10) + cn)
    img = ndimage.gaussian_filter(rng.random((75, 93, cn) if cn > 1 else (75, 93)), 1.5).astype(np.float32)
    want = F.unsharp_mask(oracle32, img, sigma, 0.7)
    np.testing.assert_array_equal(S.unsharp_mask(img, sigma, 0.7), want)
    t = torch.from_numpy(img).cuda()
    np.testing.assert_array_equal(S.unsharp_mask(t, sigma, 0.7).cpu().numpy(), want)
    np.testing.assert_array_equal(S.unsharp_mask(img, sigma, 0.0), img)


def big_holes(h, w, seed):
    """Holes of every scale: scattered pixels, blobs and a band wide enough to need several levels."""
    rng = np.random.default_rng(seed)
    m = rng.random((h, w)) < 0.05
    m |= ndimage.binary_dilation(rng.random((h, w)) < 0.004, iterations=4)
    m[h // 3:h // 3 + max(1, h // 4), w // 5:w // 5 + max(1, w // 2)] = True
    return np.where(m, 0, 255).astype(np.uint8)


@pytest.mark.parametrize("cn", [1, 3])
@pytest.mark.parametrize("shape", [(37, 53), (64, 70), (120, 97), (2, 40), (41, 1)])
@pytest.mark.parametrize("max_levels", [100, 2])
def test_average_pyramid_inpaint_bitexact(S, cn, shape, max_levels):
    import torch

    h, w = shape
    rng = np.random.default_rng(h * 7 + w + cn + max_levels)
    img = ndimage.gaussian_filter(rng.random((h, w, cn) if cn > 1 else (h, w)), 1.0).astype(np.float32) + 0.5
    mask = big_holes(h, w, h + w)
    want, want_mask = F.average_pyramid_inpaint(img, mask, max_levels)
    got, got_mask = S.average_pyramid_inpaint(img, mask, max_levels)
    np.testing.assert_array_equal(got, want)
    np.testing.assert_array_equal(got_mask, want_mask)
    # device buffers, in place on the image and its mask (the pipeline's call)
    t = torch.from_numpy(img.copy()).cuda()
    tm = torch.from_numpy(mask.copy()).cuda()
    S.average_pyramid_inpaint(t, tm, max_levels, out=t, out_mask=tm)
    np.testing.assert_array_equal(t.cpu().numpy(), want)
    np.testing.assert_array_equal(tm.cpu().numpy(), want_mask)


def test_average_pyramid_inpaint_copies(S):
    img = np.random.default_rng(2).random((16, 18)).astype(np.float32)
    full = np.full((16, 18), 255, np.uint8)
    out, om = S.average_pyramid_inpaint(img, full)
    np.testing.assert_array_equal(out, img)
    np.testing.assert_array_equal(om, full)
    out, om = S.average_pyramid_inpaint(img, None)
    np.testing.assert_array_equal(out, img)
    assert (om == 255).all()
    # no valid pixel at all: every level stays empty and the box sums stay zero
    empty = np.zeros((16, 18), np.uint8)
    want, want_mask = F.average_pyramid_inpaint(img, empty)
    out, om = S.average_pyramid_inpaint(img, empty)
    np.testing.assert_array_equal(out, want)
    np.testing.assert_array_equal(om, want_mask)

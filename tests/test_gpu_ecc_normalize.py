"""GPU ecc_normalize (ecc2.cc:385-397) against the oracle, bit for bit, and registration with
ecc.normalization_scale > 0 (c_frame_registration.cc:624-628, 807-812): the same as
registering the normalized frames against the normalized reference with normalization off."""
import numpy as np
import pytest

from tests import synth_np

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def S():
    import torch  # noqa: F401

    import serstacker_amd as S

    S._lib.load()
    return S


def gpu(a):
    import torch

    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


@pytest.mark.parametrize("shape", [(161, 113), (320, 240), (64, 9)])
@pytest.mark.parametrize("lvl", [1, 2, 3])
@pytest.mark.parametrize("masked", [False, True])
def test_ecc_normalize_bitexact(S, oracle32, shape, lvl, masked):
    import torch

    from oracle import eccnorm

    w, h = shape
    r = np.random.default_rng(w + 10 * lvl + masked)
    img = r.random((h, w)).astype(np.float32)
    mask = None
    if masked:
        mask = np.where(r.random((h, w)) < 0.1, 0, 255).astype(np.uint8)
        mask[0, 0] = 200
    want = eccnorm.ecc_normalize(oracle32, img, mask, lvl)
    np.testing.assert_array_equal(S.ecc_normalize(img, mask, lvl), want)
    out = torch.empty((h, w), dtype=torch.float32, device="cuda")
    S.ecc_normalize(gpu(img), gpu(mask) if masked else None, lvl, out=out)
    np.testing.assert_array_equal(out.cpu().numpy(), want)


@pytest.mark.parametrize("motion", [0, 3])
def test_registration_with_normalization(S, oracle32, motion):
    from oracle import eccnorm

    w, h, n, lvl = 192, 160, 4, 2
    ref, frames, _ = synth_np.frames(n, w, h, radius=0.38 * min(w, h), seed=61, shift=3.0, rot_deg=0.4,
                                     scale_dev=0.002, shear=0.001, blur=(0.8, 1.6), noise=0.002)
    kw = dict(motion_type=motion, scale=1.0, eps=0.02, ecch_max_level=3, max_iterations=50)
    on = S.c_frame_registration(S.c_image_registration_options(normalization_scale=lvl, normalization_noise=0.01,
                                                               **kw))
    assert on.setup_reference_frame(gpu(ref))
    st_on = on.register_batch([gpu(f) for f in frames])
    off = S.c_frame_registration(S.c_image_registration_options(**kw))
    assert off.setup_reference_frame(gpu(eccnorm.ecc_normalize(oracle32, ref, None, lvl)))
    st_off = off.register_batch([gpu(eccnorm.ecc_normalize(oracle32, f, None, lvl)) for f in frames])
    for a, b in zip(st_on, st_off):
        assert (a.ok, a.code, a.iterations) == (b.ok, b.code, b.iterations)
        assert list(a.params) == list(b.params)
    # the plain per-frame form agrees with the batch
    assert on.register_frame(gpu(frames[1])) == bool(st_on[1].ok)
    assert list(on.status().params) == list(st_on[1].params)

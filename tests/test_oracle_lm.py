"""Pins of the oracle's Levenberg-Marquardt and inverse-compositional ECC
(c_ecclm, c_ecc_inverse_compositional, c_ecclm_inverse_compositional,
ecc2.cc:1371-2086) and of invert_and_compose (c_image_transform.h:164-167,
312-318; c_image_transform.cc:736-833).

invert_and_compose is pinned against numpy matrix algebra on homogeneous 3x3
matrices.  The aligners are pinned by closed forms: identical images stop at
once with known iteration counts, and scenes rendered under a known warp are
recovered to the 1/32-px grain of the remap (a wrong sign, a transposed
composition or a dropped factor diverges or stops elsewhere)."""
import numpy as np
import pytest

import synth_np


def _mat_affine(p):
    return np.array([[p[0], p[1], p[2]], [p[3], p[4], p[5]], [0, 0, 1.0]])


def _mat_euclid(tx, ty, a, s, cx, cy):
    ca, sa = np.cos(a), np.sin(a)
    R = np.array([[s * ca, -s * sa], [s * sa, s * ca]])
    t = np.array([tx, ty]) - R @ np.array([cx, cy])
    return np.array([[R[0, 0], R[0, 1], t[0]], [R[1, 0], R[1, 1], t[1]], [0, 0, 1.0]])


def test_invert_and_compose_translation(oracle64, oracle32):
    for ora in (oracle64, oracle32):
        t = ora.xform(0, [3.25, -1.5])
        np.testing.assert_array_equal(ora.invert_and_compose(t, [0.25, 0.5]), [3.0, -2.0])


def test_invert_and_compose_affine_matches_matrix_algebra(oracle64):
    rnd = np.random.default_rng(3)
    for _ in range(5):
        p = np.array([1, 0, 0, 0, 1, 0.0]) + rnd.normal(0, [0.02, 0.02, 4, 0.02, 0.02, 4])
        dp = rnd.normal(0, [0.01, 0.01, 0.5, 0.01, 0.01, 0.5])
        t = oracle64.xform(3, p)
        got = oracle64.invert_and_compose(t, dp)
        want = np.linalg.inv(np.linalg.inv(_mat_affine(p)) + np.vstack([dp.reshape(2, 3), [0, 0, 0]]))
        np.testing.assert_allclose(got, want[:2].ravel(), rtol=0, atol=1e-12)
    # a zero increment returns p
    t = oracle64.xform(3, [1.01, 0.002, 3.0, -0.001, 0.99, -2.0])
    np.testing.assert_allclose(oracle64.invert_and_compose(t, np.zeros(6)), list(t.p), atol=1e-13)


@pytest.mark.parametrize("motion", [1, 2])
def test_invert_and_compose_euclidean_matches_matrix_algebra(oracle64, motion):
    rnd = np.random.default_rng(10 + motion)
    for _ in range(5):
        c = rnd.uniform(0, 100, 2)
        p = [rnd.normal(0, 5), rnd.normal(0, 5), rnd.normal(0, 0.05)] + ([1 + rnd.normal(0, 0.01)] if motion == 2 else [])
        dp = [rnd.normal(0, 0.3), rnd.normal(0, 0.3), rnd.normal(0, 0.01)] + ([rnd.normal(0, 0.005)] if motion == 2 else [])
        t = oracle64.xform(motion, p, center=tuple(c))
        got = oracle64.invert_and_compose(t, dp)
        s = p[3] if motion == 2 else 1.0
        ds = 1 + dp[3] if motion == 2 else 1.0
        want = _mat_euclid(*p[:3], s, *c) @ np.linalg.inv(_mat_euclid(*dp[:3], ds, *c))
        gs = got[3] if motion == 2 else 1.0
        np.testing.assert_allclose(_mat_euclid(*got[:3], gs, *c), want, rtol=0, atol=1e-10)


def _pair(w, h, seed, T, scale=100.0):
    render = synth_np.blob_scene(w, h, seed=seed)
    return scale * render([1, 0, 0, 0, 1, 0]), scale * render(T)


@pytest.mark.parametrize("method", ["lm", "ic", "iclm"])
def test_identical_images_stop_at_once(oracle64, method):
    ref, _ = _pair(128, 96, 5, [1, 0, 0, 0, 1, 0])
    t = oracle64.xform(3)
    if method == "lm":
        # err = |rhs|^2 = 0 < 1: converged before the first step (ecc2.cc:1566-1570)
        ok, res = oracle64.ecc_lm_align(ref, ref.copy(), t, max_eps=1e-3)
        assert ok and res.num_iterations == 0
    else:
        # dp = 0: one iteration, stopped by eps
        ok, res = oracle64.ecc_ic_align(ref, ref.copy(), t, max_eps=1e-3, lm=method == "iclm")
        assert ok and res.num_iterations == 1
    np.testing.assert_allclose(list(t.p), [1, 0, 0, 0, 1, 0], atol=1e-12)


@pytest.mark.parametrize("method", ["lm", "ic", "iclm"])
@pytest.mark.parametrize("motion", [0, 1, 3])
def test_single_level_recovers_known_warp(oracle64, method, motion):
    w, h = 160, 128
    rnd = np.random.default_rng(200 + motion)
    T = synth_np.random_affine(rnd, w, h, 1.5, 0.0 if motion == 0 else 0.4, 0.002 if motion == 3 else 0.0,
                               0.001 if motion == 3 else 0.0)
    ref, cur = _pair(w, h, 41 + motion, T)
    t = oracle64.xform(motion)
    if method == "lm":
        ok, res = oracle64.ecc_lm_align(ref, cur, t, max_iterations=100, max_eps=1e-4)
        assert ok
    else:
        ok, res = oracle64.ecc_ic_align(ref, cur, t, max_iterations=100, max_eps=1e-4, lm=method == "iclm")
    assert res.num_iterations <= 100
    assert synth_np.max_displacement(oracle64.xform_matrix(t), synth_np.invert(T), w, h) < 0.03


@pytest.mark.parametrize("method", [1, 2, 3])
@pytest.mark.parametrize("motion", [0, 1, 3])
def test_registration_methods_recover_known_warp(oracle64, oracle32, method, motion):
    """Pyramid registration with each ECC_ALIGN_METHOD recovers T^-1 (FP32 and FP64 builds)."""
    from oracle.oracle import reg_options
    w, h = 256, 192
    rnd = np.random.default_rng(300 + motion)
    T = synth_np.random_affine(rnd, w, h, 4.0, 0.0 if motion == 0 else 1.0, 0.004 if motion == 3 else 0.0,
                               0.002 if motion == 3 else 0.0)
    render = synth_np.blob_scene(w, h, seed=17 + motion)
    for ora in (oracle64, oracle32):
        r = ora.registration(reg_options(motion_type=motion, scale=1.0, eps=0.01, ecch_max_level=3,
                                         max_iterations=80, ecc_method=method))
        r.setup_reference_frame(100 * render([1, 0, 0, 0, 1, 0]))
        st = r.register_frame(100 * render(T))
        assert st.ok and st.code == 0, (method, motion, st.code)
        assert synth_np.max_displacement(st.matrix, synth_np.invert(T), w, h) < 0.05, (method, motion)


def test_lm_not_converged_level_keeps_its_scale(oracle64):
    """c_ecch::align rescales the transform only after a level whose align() returned true
    (ecc2.cc:1161-1170): c_ecclm returns false when max_iterations ends it."""
    from oracle.oracle import reg_options
    w, h = 128, 96
    T = np.array([[1, 0, 3.3], [0, 1, -2.2]])
    render = synth_np.blob_scene(w, h, seed=3)
    r = oracle64.registration(reg_options(motion_type=0, scale=1.0, eps=1e-9, ecch_max_level=2, max_iterations=1,
                                          ecc_method=2))
    r.setup_reference_frame(100 * render([1, 0, 0, 0, 1, 0]))
    st = r.register_frame(100 * render(T))
    assert st.levels_failed == 2

"""Master-frame selection host rules (c_image_stacking_pipeline_base.cc:312-395,
c_image_stacking_pipeline.cc:1219-1232): the library's C entry points against the
oracle's restatement over a grid of sizes, and the oracle against cases worked by hand
from the reference text.  Host-only calls: no GPU needed."""
import itertools

import numpy as np
import pytest


@pytest.fixture(scope="module")
def S():
    import serstacker_amd as S

    S._lib.load()
    return S


def test_scan_window_hand_cases():
    from oracle import master

    # 1000 frames, scan 200: starts 100 before the middle and stops at the middle (start + 100)
    assert master.scan_window(1000, 200) == (400, 500)
    # fewer frames than the scan size: the whole sequence
    assert master.scan_window(150, 200) == (0, 150)
    assert master.scan_window(200, 200) == (0, 200)
    # the scan size is at least 3
    assert master.scan_window(10, 1) == (4, 5)
    assert master.scan_window(3, 0) == (0, 3)


def test_select_hand_cases():
    from oracle import master

    assert master.select([0.5, 2.0, 2.0, 1.0], 40) == 41  # ties keep the earlier frame
    assert master.select([0.0, 0.0], 7) == 7  # nothing beats 0: the first frame
    assert master.select([-1.0, float("nan"), 3.0], 0) == 2  # NaN never wins a comparison
    assert master.select([], 5) == 5


def test_stack_window_hand_cases():
    from oracle import master

    # window centred on the master frame
    assert master.stack_window(1000, 0, 500, 100) == (450, 550)
    # clipped at the start index
    assert master.stack_window(1000, 20, 30, 100) == (20, 120)
    # clipped at the end: shifted back to keep max_frames_to_stack frames
    assert master.stack_window(1000, 0, 980, 100) == (900, 1000)
    # the sequence is shorter than start + max: everything from the start index
    assert master.stack_window(90, 5, 40, 100) == (5, 90)
    assert master.stack_window(100, -3, 50, 100) == (0, 100)


def test_scan_window_matches_oracle(S):
    from oracle import master

    for size, m in itertools.product(range(0, 260, 7), [0, 1, 2, 3, 4, 5, 50, 99, 100, 101, 200]):
        assert S.master_scan_window(size, m) == master.scan_window(size, m), (size, m)


def test_select_matches_oracle(S):
    from oracle import master

    r = np.random.default_rng(5)
    for n in [0, 1, 2, 17, 200]:
        for start in [0, 13]:
            q = r.integers(0, 6, n).astype(np.float64)  # many ties
            assert S.select_master_frame(q, start) == master.select(list(q), start)
            q = r.normal(size=n)
            assert S.select_master_frame(q, start) == master.select(list(q), start)


def test_stack_window_matches_oracle(S):
    from oracle import master

    for size, sfi, pos, m in itertools.product([2, 50, 101, 1000], [-5, 0, 10, 60], [0, 25, 50, 99, 500, 999],
                                                [2, 3, 10, 49, 50, 51, 100, 2000]):
        assert S.master_stack_window(size, sfi, pos, m) == master.stack_window(size, sfi, pos, m), (size, sfi, pos, m)


def test_host_entry_points_reject_bad_arguments(S):
    """Argument checks of the host-only entry points (no GPU call is made)."""
    import ctypes as C

    lib = S._lib.load()
    a, b = C.c_int32(), C.c_int32()
    assert lib.ssg_master_scan_window(10, 5, None, C.byref(b)) == S._lib.SSG_E_INVALID_ARG
    assert lib.ssg_master_stack_window(10, 0, 5, 4, C.byref(a), None) == S._lib.SSG_E_INVALID_ARG
    assert lib.ssg_select_master_frame(3, None, 0, C.byref(a)) == S._lib.SSG_E_INVALID_ARG
    assert lib.ssg_upscale_size(7, 10, 10, C.byref(a), C.byref(b)) == S._lib.SSG_E_INVALID_ARG
    assert lib.ssg_upscale_size(-1, 10, 10, C.byref(a), C.byref(b)) == S._lib.SSG_E_INVALID_ARG
    assert S.upscale_size(S._lib.UPSCALE_X15, 1920, 1080) == (2880, 1620)
    assert S.upscale_size(S._lib.UPSCALE_X2, 33, 21) == (66, 42)
    assert S.upscale_size(S._lib.UPSCALE_X3, 33, 21) == (99, 63)
    assert S.upscale_size(S._lib.UPSCALE_NONE, 33, 21) == (33, 21)

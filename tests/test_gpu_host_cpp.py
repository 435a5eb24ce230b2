"""The plain C++ host interface (include/ssg.hpp) on the GPU: the same seeded
frames registered and stacked from C++ and from the Python binding give the
same parameters, iteration counts and stack (same kernels, same order)."""
import json
import os
import subprocess

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DEMO = os.path.join(ROOT, "tests", "cpp", "host_api_demo")


def test_cpp_host_api_matches_python_binding():
    if not os.path.exists(DEMO):
        pytest.fail("tests/cpp/host_api_demo is not built (run __graft_entry__.build())")
    W, H, N = 320, 240, 6
    print(f"[cpp host demo] {W}x{H} x {N}", flush=True)
    r = subprocess.run([DEMO, str(W), str(H), str(N)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    out = json.loads(r.stdout.strip().splitlines()[-1])
    assert out["plain_equals_batch"] == 1
    assert out["map_err"] < 1e-3
    assert out["accumulated"] == N

    import torch

    import serstacker_amd as S
    p = S.synth_params(width=W, height=H, disk_radius=0.38 * min(W, H), max_shift=6.0, max_rotation_deg=0.5,
                       max_scale_dev=0.003, max_shear=0.002, blur_sigma_min=0.8, blur_sigma_max=2.0, seed=4242)
    frames = [torch.empty((H, W), dtype=torch.float32, device="cuda") for _ in range(N)]
    ref = torch.empty((H, W), dtype=torch.float32, device="cuda")
    S.synth_render(p, 0, frames)
    S.synth_render(p, -1, [ref])
    reg = S.c_frame_registration(S.c_image_registration_options(motion_type=3, scale=1.0, eps=0.02,
                                                                ecch_max_level=3, max_iterations=50))
    assert reg.setup_reference_frame(ref)
    st = reg.register_batch(frames)
    # %.9g round-trips float32 exactly
    assert np.array_equal(np.float32(list(st[0].params)), np.float32(out["params0"]))
    assert [s.iterations for s in st] == out["iterations"]
    wts = S.frame_sharpness(frames)
    acc = S.c_weigthed_average(W, H, 1)
    acc.add_registered_batch(frames, st, S.MOTION_AFFINE, frame_weights=wts)
    avg, mask = acc.compute()
    avg = np.asarray(avg, np.float64)
    assert abs(avg.sum() - out["avg_sum"]) <= 1e-12 * max(1.0, abs(out["avg_sum"]))
    assert int((np.asarray(mask) != 0).sum()) == out["covered"]
    # eccflow refinement + dense-field stack
    maps = [torch.zeros((H, W, 2), dtype=torch.float32, device="cuda") for _ in range(N)]
    reg.register_batch_remaps(frames, maps)
    flow = S.c_eccflow(downscale=2, support_scale=3, max_iterations=2, min_image_size=-1, max_pyramid_level=3)
    flow.set_reference_image(ref)
    assert flow.num_levels() == out["flow_levels"]
    flow.compute_batch(frames, maps, rmaps=maps)
    acc2 = S.c_weigthed_average(W, H, 1)
    acc2.add_remapped_batch(frames, maps, frame_weights=wts)
    favg, fmask = acc2.compute()
    favg = np.asarray(favg, np.float64)
    assert abs(favg.sum() - out["flow_avg_sum"]) <= 1e-12 * max(1.0, abs(out["flow_avg_sum"]))
    assert int((np.asarray(fmask) != 0).sum()) == out["flow_covered"]
    # x1.5 upscale of the refined remaps + stack, pyramid inpaint of that stack, master-frame scan
    dw, dh = S.upscale_size(S._lib.UPSCALE_X15, W, H)
    up = [torch.empty((dh, dw, 2), dtype=torch.float32, device="cuda") for _ in range(N)]
    S.upscale_remaps(S._lib.UPSCALE_X15, maps, out=up)
    acc3 = S.c_weigthed_average(dw, dh, 1)
    acc3.add_remapped_batch(frames, up, frame_weights=wts)
    uavg, umask = acc3.compute()
    assert abs(np.asarray(uavg, np.float64).sum() - out["up_avg_sum"]) <= 1e-12 * max(1.0, abs(out["up_avg_sum"]))
    assert int((np.asarray(umask) != 0).sum()) == out["up_covered"]
    fin, _ = S.average_pyramid_inpaint(uavg, umask, 100)
    assert abs(np.asarray(fin, np.float64).sum() - out["apr_sum"]) <= 1e-12 * max(1.0, abs(out["apr_sum"]))
    q = S.master_frame_quality(frames, 0, 1, 1)
    np.testing.assert_array_equal(q, np.asarray(out["quality"]))
    assert S.select_master_frame(q, 0) == out["master"]
    assert list(S.master_scan_window(N, 3)) == out["scan"]

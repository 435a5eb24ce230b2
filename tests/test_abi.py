"""The C-ABI library loads and exports every entry point include/ssg.h declares
(no compute calls: this runs without a GPU)."""
import ctypes as C
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "ssg.h")
LIB = os.path.join(ROOT, "serstacker_amd", "libssg.so")


def declared():
    src = open(HEADER).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(ssg_[a-z0-9_]+)\s*\(", src)))


@pytest.fixture(scope="module")
def lib():
    if not os.path.exists(LIB):
        pytest.fail("libssg.so is not built (run __graft_entry__.build())")
    return C.CDLL(LIB)


def test_header_declares_entry_points():
    names = declared()
    assert "ssg_registration_register_batch" in names
    assert len(names) >= 20


def test_every_declared_symbol_is_exported(lib):
    missing = [n for n in declared() if not hasattr(lib, n)]
    assert not missing, missing


def test_binding_covers_header():
    from serstacker_amd import _lib

    assert sorted(_lib.ENTRY_POINTS) == declared()


def test_abi_version_and_defaults(lib):
    from serstacker_amd import _lib

    L = _lib.load()
    assert L.ssg_abi_version() == 1
    o = _lib.ImageRegistrationOptions()
    L.ssg_default_image_registration_options(C.byref(o))
    # c_frame_registration.h:47-64 and :119-136
    assert o.motion_type == _lib.MOTION_AFFINE
    assert o.interpolation == _lib.INTER_LINEAR and o.border_mode == _lib.BORDER_REFLECT101
    assert (o.ecc.scale, o.ecc.eps, o.ecc.min_rho) == (0.5, 0.2, 0.8)
    assert (o.ecc.input_smooth_sigma, o.ecc.reference_smooth_sigma, o.ecc.update_step_scale) == (1.0, 1.0, 1.5)
    assert (o.ecc.max_iterations, o.ecc.ecch_max_level, o.ecc.ecch_minimum_image_size) == (50, 0, 16)
    assert o.ecc.ecc_method == _lib.ECC_ALIGN_LM and o.ecc.ecch_estimate_translation_first == 1
    assert (o.eccflow.support_scale, o.eccflow.max_iterations, o.eccflow.scale_factor) == (4, 3, 0.75)
    assert o.enable_feature_registration == 1 and o.enable_ecc_registration == 0
    assert o.ecc_registration_channel == 104  # color_channel_gray (core/proc/extract_channel.h:31)


def test_struct_layout_matches_header(tmp_path):
    """sizeof/offsetof of every ABI struct, compiled from include/ssg.h with gcc,
    equal the ctypes binding's layout."""
    import subprocess

    from serstacker_amd import _lib

    structs = {
        "ssg_image": _lib.Image,
        "ssg_ecc_registration_options": _lib.EccRegistrationOptions,
        "ssg_eccflow_registration_options": _lib.EccflowRegistrationOptions,
        "ssg_image_registration_options": _lib.ImageRegistrationOptions,
        "ssg_registration_status": _lib.RegistrationStatus,
        "ssg_profile": _lib.Profile,
        "ssg_synth_params": _lib.SynthParams,
        "ssg_ser_header": _lib.SerHeader,
    }
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "ssg.h"', "int main(void) {"]
    for cname, cls in structs.items():
        lines.append(f'printf("{cname} size %zu\\n", sizeof({cname}));')
        for fname, _ in cls._fields_:
            lines.append(f'printf("{cname} {fname} %zu\\n", offsetof({cname}, {fname}));')
    lines.append("return 0; }")
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split("\n")
    got = {}
    for ln in out:
        if ln:
            s, f, v = ln.split()
            got[(s, f)] = int(v)
    for cname, cls in structs.items():
        assert got[(cname, "size")] == C.sizeof(cls), cname
        for fname, _ in cls._fields_:
            assert got[(cname, fname)] == getattr(cls, fname).offset, (cname, fname)


def test_invalid_arguments_fail_without_gpu_work(lib):
    from serstacker_amd import _lib

    L = _lib.load()
    out = C.c_void_p()
    o = _lib.ImageRegistrationOptions()
    L.ssg_default_image_registration_options(C.byref(o))
    # the reference defaults select feature registration, which is not on this path
    assert L.ssg_registration_create(C.byref(o), C.byref(out)) == -4
    assert b"feature" in L.ssg_last_error() or b"ECC" in L.ssg_last_error()
    assert L.ssg_accumulation_create(99, 8, 8, 1, C.byref(out)) == -4


def test_cpp_host_header_compiles():
    """include/ssg.hpp is plain C++17 over the C ABI (no HIP or OpenCV headers)."""
    import shutil
    import subprocess
    import tempfile

    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("g++ not available")
    src = "#include \"ssg.hpp\"\nint main() { ssg::c_frame_registration r; (void)r.options(); return 0; }\n"
    with tempfile.TemporaryDirectory() as d:
        f = os.path.join(d, "t.cpp")
        with open(f, "w") as fh:
            fh.write(src)
        r = subprocess.run([gxx, "-std=c++17", "-Wall", "-Werror", "-fsyntax-only", f"-I{os.path.join(ROOT, 'include')}",
                            f], capture_output=True, text=True)
        assert r.returncode == 0, r.stderr


def test_integration_shim_compiles():
    """The reference-side shim printed in INTEGRATION.md compiles against include/ and the
    restated reference interfaces under tests/cpp/refshim (the registration class's virtual
    operations and protected state, the accumulator base with its pure virtuals, the option
    structs); every override is marked override, so a signature that drifts from the
    reference's fails here."""
    import shutil
    import subprocess
    import tempfile

    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("g++ not available")
    text = open(os.path.join(ROOT, "INTEGRATION.md")).read()
    m = re.search(r"<!-- integration-shim -->\s*```cpp\n(.*?)\n```", text, re.S)
    assert m, "shim block missing"
    code = m.group(1)
    for cls in ("c_frame_registration_mi355x", "c_weigthed_average_mi355x", "c_bayer_average_mi355x",
                "c_laplacian_pyramid_focus_stacking_mi355x"):
        assert f"class {cls}" in code
    # instantiate every class, so a pure virtual left unimplemented is an error
    main = ("\nint main() {\n  c_image_registration_options o;\n  c_frame_registration_mi355x r(o);\n"
            "  c_weigthed_average_mi355x a;\n  c_bayer_average_mi355x b;\n"
            "  c_laplacian_pyramid_focus_stacking_mi355x f{c_laplacian_pyramid_focus_stacking::options()};\n"
            "  (void)r; (void)a; (void)b; (void)f;\n  return 0;\n}\n")
    with tempfile.TemporaryDirectory() as d:
        f = os.path.join(d, "shim.cpp")
        with open(f, "w") as fh:
            fh.write(code.replace("#pragma once\n", "") + main)
        r = subprocess.run([gxx, "-std=c++17", "-Wall", "-Wextra", "-Werror", "-Wsuggest-override", "-fsyntax-only",
                            f"-I{os.path.join(ROOT, 'include')}", f"-I{os.path.join(ROOT, 'tests', 'cpp', 'refshim')}",
                            f], capture_output=True, text=True)
        assert r.returncode == 0, r.stderr


def test_host_views_keep_their_memory():
    """image_view of a row-strided array views it in place; of an array whose rows are
    not contiguous it views a copy that the view itself keeps alive; an output array
    that cannot be viewed in place is refused."""
    import gc

    import numpy as np

    import serstacker_amd as S

    a = np.arange(40 * 30, dtype=np.float32).reshape(30, 40)
    crop = a[5:20, 3:33]
    v = S.image_view(crop)
    assert v.data == crop.ctypes.data and v.step == a.strides[0]
    assert (v.width, v.height) == (30, 15)

    t = a.T  # columns are rows now: not viewable in place
    v = S.image_view(t)
    assert v.data != t.ctypes.data and v.step == t.shape[1] * 4
    gc.collect()
    junk = [np.full(t.shape, -1.0, np.float32) for _ in range(8)]  # reuse freed blocks, if any
    got = np.ctypeslib.as_array(C.cast(v.data, C.POINTER(C.c_float)), shape=(t.shape[0], t.shape[1]))
    assert np.array_equal(got, t), "the copied input was freed while its view is alive"
    del junk

    views = S.image_views([a[:, ::2], a.T])
    for i, ref in enumerate((a[:, ::2], a.T)):
        got = np.ctypeslib.as_array(C.cast(views[i].data, C.POINTER(C.c_float)),
                                    shape=(views[i].height, views[i].width))
        assert np.array_equal(got, ref)

    rgb = np.zeros((10, 12, 3), np.float32)
    v = S.image_view(rgb[2:8, 1:9])
    assert v.data == rgb[2:8, 1:9].ctypes.data and v.step == 12 * 3 * 4
    with pytest.raises(ValueError):
        S.image_view(a.T, out=True)
    with pytest.raises(ValueError):
        S.image_view(rgb[:, :, :2], out=True)


def test_tuning_knobs_and_autotuner_cache(tmp_path):
    """The tuning knobs keep their ranges; the autotuner's coordinate descent keeps the fastest
    candidate of each choice, caches it per device / workload / size class and reuses it (a stub
    workload whose time depends on the choice; no GPU work)."""
    from serstacker_amd import _lib, autotune

    L = _lib.load()
    v = C.c_int32()
    assert L.ssg_tune_get(_lib.TUNE_ACC_ROWS, C.byref(v)) == 0 and v.value == 4
    assert L.ssg_tune_set(_lib.TUNE_ACC_ROWS, 3) == -1
    assert L.ssg_tune_set(_lib.TUNE_ECC_GRID, 99) == -1
    assert L.ssg_tune_set(7, 1) == -1
    assert autotune.size_class(1920, 1080, 2000) == "2048kpxx2048"
    assert autotune.device_key("AMD Instinct MI355X (gfx950:sramecc+:xnack-, 256 CUs)") == \
        "gfx950:sramecc+:xnack-, 256 CUs"
    import json
    shipped = json.load(open(autotune.SHIPPED))
    assert shipped and all(k.split("|")[0].startswith("gfx950") for k in shipped)
    calls = []

    def run(choice):
        calls.append(dict(choice))
        import time as _t
        _t.sleep(0.001 * (abs(choice["acc_blocks"] - 2048) / 1024 + (choice["frames_per_launch"] != 500)))

    space = {"frames_per_launch": [1000, 500], "acc_blocks": [4096, 2048]}
    t = autotune.Autotuner(str(tmp_path / "tune.json"))
    choice, times, cached = t.tune("stub", "k", space, run, repeats=1)
    assert not cached and choice == {"frames_per_launch": 500, "acc_blocks": 2048}
    assert autotune.get_knob("acc_blocks") == 2048
    n = len(calls)
    again, _, cached = autotune.Autotuner(str(tmp_path / "tune.json")).tune("stub", "k", space, run)
    assert cached and again == choice and len(calls) == n
    autotune.set_knob("acc_blocks", 4096)

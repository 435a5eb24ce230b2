"""ctypes binding of libssg.so (the C ABI declared in include/ssg.h).

The library must have been built (serstacker_amd/csrc/build.py); there is no
fallback path: every call goes to the HIP implementation, and a missing
library raises at import of the symbols.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libssg.so")

SSG_OK = 0
SSG_E_INVALID_ARG, SSG_E_HIP, SSG_E_NOT_READY, SSG_E_UNSUPPORTED, SSG_E_NO_MEMORY = -1, -2, -3, -4, -5
SSG_8UC1, SSG_16UC1, SSG_32SC1, SSG_32FC1, SSG_32FC2, SSG_32SC3, SSG_32FC3 = 0, 2, 4, 5, 13, 20, 21
SSG_64FC1, SSG_8UC3, SSG_16UC3, SSG_64FC3 = 6, 16, 18, 22
UPSCALE_NONE, UPSCALE_X2, UPSCALE_X15, UPSCALE_X3 = 0, 1, 2, 3
COLORID_MONO, COLORID_RGB, COLORID_BGR = 0, 100, 101
COLORID_BAYER_RGGB, COLORID_BAYER_GRBG, COLORID_BAYER_GBRG, COLORID_BAYER_BGGR = 8, 9, 10, 11
COLORID_BAYER_CYYM, COLORID_BAYER_YCMY, COLORID_BAYER_YMCY, COLORID_BAYER_MYYC = 16, 17, 18, 19
(DEBAYER_DISABLE, DEBAYER_NN, DEBAYER_NN2, DEBAYER_VNG, DEBAYER_EA, DEBAYER_SP, DEBAYER_MATRIX, DEBAYER_PLANE_0,
 DEBAYER_PLANE_1, DEBAYER_PLANE_2, DEBAYER_PLANE_3, DEBAYER_AVGBP) = range(12)
DEBAYER_DEFAULT = -1
FOCUS_SELECT_MAX_ENERGY, FOCUS_WEIGHTED_AVERAGE, FOCUS_LPG_WEIGHTS = 0, 1, 2
MOTION_TRANSLATION, MOTION_EUCLIDEAN, MOTION_SCALED_EUCLIDEAN, MOTION_AFFINE = 0, 1, 2, 3
ECC_ALIGN_FORWARD_ADDITIVE, ECC_ALIGN_INVERSE_COMPOSITIONAL, ECC_ALIGN_LM, ECC_ALIGN_INVERSE_COMPOSITIONAL_LM = 0, 1, 2, 3
INTER_NEAREST, INTER_LINEAR, INTER_CUBIC, INTER_AREA = 0, 1, 2, 3
BORDER_CONSTANT, BORDER_REPLICATE, BORDER_REFLECT, BORDER_WRAP, BORDER_REFLECT101, BORDER_TRANSPARENT = range(6)
ACC_AVERAGE, ACC_WEIGHTED_AVERAGE, ACC_FOCUS_STACK, ACC_FFT, ACC_BAYER_AVERAGE = range(5)
REG_OK, REG_NOT_RUN, REG_LEVEL_FAILED, REG_RHO_TRANSLATION, REG_RHO = range(5)


class SsgError(RuntimeError):
    pass


class SerHeader(C.Structure):
    """ssg_ser_header: c_ser_file::file_header (core/io/c_ser_file.h:38-53)."""
    _pack_ = 1
    _fields_ = [("file_id", C.c_char * 14), ("luid", C.c_int32), ("color_id", C.c_int32),
                ("is_little_endian", C.c_int32), ("image_width", C.c_int32), ("image_height", C.c_int32),
                ("bits_per_plane", C.c_int32), ("frames_count", C.c_int32), ("observer", C.c_char * 40),
                ("instrument", C.c_char * 40), ("telescope", C.c_char * 40), ("date_time", C.c_uint64),
                ("date_time_utc", C.c_uint64)]


class Image(C.Structure):
    _fields_ = [
        ("data", C.c_void_p),
        ("width", C.c_int32),
        ("height", C.c_int32),
        ("step", C.c_int64),
        ("type", C.c_int32),
        ("on_device", C.c_int32),
    ]


class EccRegistrationOptions(C.Structure):
    _fields_ = [
        ("scale", C.c_double),
        ("eps", C.c_double),
        ("min_rho", C.c_double),
        ("input_smooth_sigma", C.c_double),
        ("reference_smooth_sigma", C.c_double),
        ("update_step_scale", C.c_double),
        ("se_radius", C.c_int32),
        ("ecc_method", C.c_int32),
        ("max_iterations", C.c_int32),
        ("ecch_max_level", C.c_int32),
        ("ecch_minimum_image_size", C.c_int32),
        ("normalization_noise", C.c_double),
        ("normalization_scale", C.c_int32),
        ("ecch_estimate_translation_first", C.c_int32),
        ("replace_planetary_disk_with_mask", C.c_int32),
    ]


class EccflowRegistrationOptions(C.Structure):
    _fields_ = [
        ("update_multiplier", C.c_double),
        ("input_smooth_sigma", C.c_double),
        ("reference_smooth_sigma", C.c_double),
        ("noise_level", C.c_double),
        ("scale_factor", C.c_double),
        ("max_iterations", C.c_int32),
        ("support_scale", C.c_int32),
        ("min_image_size", C.c_int32),
        ("max_pyramid_level", C.c_int32),
        ("downscale_method", C.c_int32),
    ]


class ImageRegistrationOptions(C.Structure):
    _fields_ = [
        ("motion_type", C.c_int32),
        ("ecc_registration_channel", C.c_int32),
        ("interpolation", C.c_int32),
        ("border_mode", C.c_int32),
        ("border_value", C.c_double * 4),
        ("ecc", EccRegistrationOptions),
        ("eccflow", EccflowRegistrationOptions),
        ("enable_feature_registration", C.c_int32),
        ("enable_ecc_registration", C.c_int32),
        ("enable_eccflow_registration", C.c_int32),
        ("accumulate_and_compensate_turbulent_flow", C.c_int32),
    ]


class RegistrationStatus(C.Structure):
    _fields_ = [
        ("ok", C.c_int32),
        ("code", C.c_int32),
        ("rho_translation", C.c_double),
        ("rho", C.c_double),
        ("eps", C.c_double),
        ("iterations_translation", C.c_int32),
        ("iterations", C.c_int32),
        ("levels_failed", C.c_int32),
        ("nparams", C.c_int32),
        ("params", C.c_float * 6),
        ("matrix", C.c_double * 6),
        ("level_iterations", C.c_int32 * 8),
    ]


PROF_ECC_TRANSLATION, PROF_ECC_MAIN, PROF_CORRELATION, PROF_PYRAMID, PROF_ACCUMULATE = 0, 1, 2, 3, 4
PROF_ECC_SOLVE, PROF_FFT, PROF_ECCFLOW = 5, 6, 7
PROF_FOCUS_PREP, PROF_FOCUS_BLUR, PROF_FOCUS_DOWN, PROF_FOCUS_FUSE, PROF_GRAPH = 8, 9, 10, 11, 12
PROF_COUNT = 13
TUNE_ACC_ROWS, TUNE_ACC_BLOCKS, TUNE_ECC_GRID, TUNE_FOCUS_CHUNK, TUNE_PYR_STRIP = 0, 1, 2, 3, 4
ECCFLOW_DOWNSCALE_RECURSIVE_RESIZE, ECCFLOW_DOWNSCALE_FULL_RESIZE, ECCFLOW_DOWNSCALE_PYRAMID = 0, 1, 2


class EccflowOptions(C.Structure):
    _fields_ = [
        ("input_smooth_sigma", C.c_double),
        ("reference_smooth_sigma", C.c_double),
        ("update_multiplier", C.c_double),
        ("scale_factor", C.c_double),
        ("noise_level", C.c_double),
        ("max_iterations", C.c_int32),
        ("support_scale", C.c_int32),
        ("min_image_size", C.c_int32),
        ("max_pyramid_level", C.c_int32),
        ("downscale", C.c_int32),
    ]


class FocusStackingOptions(C.Structure):
    _fields_ = [("fusing_policy", C.c_int32), ("inpaint_mask_holes", C.c_int32), ("avgchannel", C.c_int32),
                ("kradius", C.c_int32), ("ksigma", C.c_double), ("lpg_k", C.c_double), ("lpg_p", C.c_double),
                ("lpg_sigma", C.c_double), ("max_levels", C.c_int32)]


class Profile(C.Structure):
    _fields_ = [("launches", C.c_int64 * PROF_COUNT), ("ms", C.c_double * PROF_COUNT),
                ("bytes", C.c_double * PROF_COUNT)]


class SynthParams(C.Structure):
    _fields_ = [
        ("width", C.c_int32),
        ("height", C.c_int32),
        ("disk_radius", C.c_double),
        ("texture_sigma", C.c_double),
        ("texture_amplitude", C.c_double),
        ("limb_darkening", C.c_double),
        ("max_shift", C.c_double),
        ("max_rotation_deg", C.c_double),
        ("max_scale_dev", C.c_double),
        ("max_shear", C.c_double),
        ("blur_sigma_min", C.c_double),
        ("blur_sigma_max", C.c_double),
        ("peak_electrons", C.c_double),
        ("read_noise", C.c_double),
        ("output_scale", C.c_double),
        ("seed", C.c_uint64),
    ]


# (name, restype, argtypes) of every entry point of include/ssg.h
_P = C.POINTER
_SIGS = [
    ("ssg_abi_version", C.c_int, []),
    ("ssg_last_error", C.c_char_p, []),
    ("ssg_device_info", C.c_int, [_P(C.c_int), C.c_char_p, C.c_int]),
    ("ssg_default_image_registration_options", None, [_P(ImageRegistrationOptions)]),
    ("ssg_registration_create", C.c_int, [_P(ImageRegistrationOptions), _P(C.c_void_p)]),
    ("ssg_registration_destroy", None, [C.c_void_p]),
    ("ssg_registration_setup_reference_frame", C.c_int, [C.c_void_p, _P(Image), _P(Image)]),
    ("ssg_registration_register_frame", C.c_int, [C.c_void_p, _P(Image), _P(Image), _P(RegistrationStatus)]),
    ("ssg_registration_register_batch", C.c_int, [C.c_void_p, C.c_int, _P(Image), _P(Image), _P(RegistrationStatus)]),
    ("ssg_registration_create_remap", C.c_int, [C.c_void_p, _P(RegistrationStatus), _P(Image)]),
    ("ssg_registration_current_remap", C.c_int, [C.c_void_p, _P(Image)]),
    ("ssg_registration_remap", C.c_int, [C.c_void_p, _P(Image), _P(Image), _P(Image), _P(Image), C.c_int32,
                                         C.c_int32, C.c_double]),
    ("ssg_registration_register_batch_remaps", C.c_int, [C.c_void_p, C.c_int, _P(Image), _P(Image),
                                                         _P(RegistrationStatus), _P(Image)]),
    ("ssg_remap", C.c_int, [_P(Image), _P(Image), _P(Image), _P(C.c_double), C.c_int32, _P(C.c_float), _P(Image),
                            _P(Image), C.c_int32, C.c_int32, C.c_double]),
    ("ssg_accumulation_create", C.c_int, [C.c_int32, C.c_int32, C.c_int32, C.c_int32, _P(C.c_void_p)]),
    ("ssg_accumulation_destroy", None, [C.c_void_p]),
    ("ssg_accumulation_add", C.c_int, [C.c_void_p, _P(Image), _P(Image)]),
    ("ssg_accumulation_add_frames", C.c_int, [C.c_void_p, C.c_int, _P(Image)]),
    ("ssg_accumulation_add_registered_batch", C.c_int, [C.c_void_p, C.c_int, _P(Image), _P(RegistrationStatus),
                                                        C.c_int32, _P(C.c_float), _P(Image), C.c_int32, C.c_int32,
                                                        C.c_double]),
    ("ssg_accumulation_add_remapped_batch", C.c_int, [C.c_void_p, C.c_int, _P(Image), _P(Image), _P(C.c_float),
                                                      _P(Image), C.c_int32, C.c_int32, C.c_double]),
    ("ssg_accumulation_compute", C.c_int, [C.c_void_p, _P(Image), _P(Image), C.c_double]),
    ("ssg_accumulation_get_acc_counters", C.c_int, [C.c_void_p, _P(Image)]),
    ("ssg_accumulation_get_counts", C.c_int, [C.c_void_p, _P(Image)]),
    ("ssg_accumulation_accumulated_frames", C.c_int, [C.c_void_p, _P(C.c_int32)]),
    ("ssg_accumulation_clear", C.c_int, [C.c_void_p]),
    ("ssg_phasecorr_create", C.c_int, [C.c_int32, C.c_int32, C.c_int32, _P(C.c_void_p)]),
    ("ssg_phasecorr_destroy", None, [C.c_void_p]),
    ("ssg_phasecorr_set_reference", C.c_int, [C.c_void_p, _P(Image)]),
    ("ssg_phasecorr_estimate", C.c_int, [C.c_void_p, C.c_int, _P(Image), _P(C.c_double)]),
    ("ssg_phasecorr_set_profiling", C.c_int, [C.c_void_p, C.c_int]),
    ("ssg_phasecorr_get_profile", C.c_int, [C.c_void_p, _P(Profile), C.c_int]),
    ("ssg_default_focus_stacking_options", None, [_P(FocusStackingOptions)]),
    ("ssg_accumulation_set_focus_options", C.c_int, [C.c_void_p, _P(FocusStackingOptions)]),
    ("ssg_lpg", C.c_int, [_P(Image), _P(Image), C.c_double, C.c_double, _P(C.c_double)]),
    ("ssg_ser_open", C.c_int, [C.c_char_p, _P(C.c_void_p)]),
    ("ssg_ser_close", None, [C.c_void_p]),
    ("ssg_ser_get_header", C.c_int, [C.c_void_p, _P(SerHeader), _P(C.c_int32)]),
    ("ssg_ser_timestamps", C.c_int, [C.c_void_p, _P(C.c_uint64), C.c_int32, _P(C.c_int32)]),
    ("ssg_ser_seek", C.c_int, [C.c_void_p, C.c_int32]),
    ("ssg_ser_curpos", C.c_int, [C.c_void_p, _P(C.c_int32)]),
    ("ssg_ser_read", C.c_int, [C.c_void_p, _P(Image)]),
    ("ssg_ser_read_batch", C.c_int, [C.c_void_p, C.c_int32, _P(Image)]),
    ("ssg_debayer", C.c_int, [_P(Image), C.c_int32, C.c_int32, _P(Image)]),
    ("ssg_inpaint_linear", C.c_int, [_P(Image), _P(Image)]),
    ("ssg_unsharp_mask", C.c_int, [_P(Image), _P(Image), C.c_double, C.c_double]),
    ("ssg_average_pyramid_inpaint", C.c_int, [_P(Image), _P(Image), _P(Image), _P(Image), C.c_int32]),
    ("ssg_upscale_size", C.c_int, [C.c_int32, C.c_int32, C.c_int32, _P(C.c_int32), _P(C.c_int32)]),
    ("ssg_upscale_remaps", C.c_int, [C.c_int32, C.c_int32, _P(Image), _P(Image)]),
    ("ssg_upscale_images", C.c_int, [C.c_int32, C.c_int32, _P(Image), _P(Image), _P(Image), _P(Image)]),
    ("ssg_read_input_frames", C.c_int, [C.c_int32, _P(Image), C.c_int32, C.c_int32, C.c_int32, _P(Image), _P(Image),
                                        _P(Image)]),
    ("ssg_local_variance_map", C.c_int, [_P(Image), C.c_int32, C.c_int32, _P(Image), _P(C.c_double)]),
    ("ssg_ecc_normalize", C.c_int, [_P(Image), _P(Image), _P(Image), C.c_int32]),
    ("ssg_master_frame_quality", C.c_int, [C.c_int32, _P(Image), C.c_int32, C.c_int32, C.c_int32, _P(C.c_double)]),
    ("ssg_master_scan_window", C.c_int, [C.c_int32, C.c_int32, _P(C.c_int32), _P(C.c_int32)]),
    ("ssg_select_master_frame", C.c_int, [C.c_int32, _P(C.c_double), C.c_int32, _P(C.c_int32)]),
    ("ssg_master_stack_window", C.c_int, [C.c_int32, C.c_int32, C.c_int32, C.c_int32, _P(C.c_int32), _P(C.c_int32)]),
    ("ssg_accumulation_set_bayer_pattern", C.c_int, [C.c_void_p, C.c_int32]),
    ("ssg_accumulation_set_remap", C.c_int, [C.c_void_p, _P(Image)]),
    ("ssg_accumulation_set_input_bpp", C.c_int, [C.c_void_p, C.c_int32]),
    ("ssg_frame_sharpness", C.c_int, [C.c_int, _P(Image), _P(C.c_float)]),
    ("ssg_tune_set", C.c_int, [C.c_int32, C.c_int32]),
    ("ssg_tune_get", C.c_int, [C.c_int32, _P(C.c_int32)]),
    ("ssg_stack_batch", C.c_int, [C.c_void_p, C.c_void_p, C.c_int32, _P(Image), C.c_int32, C.c_int32, C.c_double, C.c_int32, _P(RegistrationStatus)]),
    ("ssg_focus_stack_batch", C.c_int, [C.c_int32, _P(C.c_void_p), _P(C.c_int32), _P(_P(Image)), _P(_P(C.c_float))]),
    ("ssg_accumulation_reinitialize", C.c_int, [C.c_void_p, _P(Image), _P(Image)]),
    ("ssg_accumulation_state_size", C.c_int, [C.c_void_p, _P(C.c_uint64)]),
    ("ssg_accumulation_state_save", C.c_int, [C.c_void_p, C.c_void_p, C.c_uint64]),
    ("ssg_accumulation_state_load", C.c_int, [C.c_void_p, C.c_void_p, C.c_uint64]),
    ("ssg_frame_sharpness_noalloc", C.c_int, [C.c_int, _P(Image), _P(C.c_float)]),
    ("ssg_registration_set_profiling", C.c_int, [C.c_void_p, C.c_int]),
    ("ssg_registration_get_profile", C.c_int, [C.c_void_p, _P(Profile), C.c_int]),
    ("ssg_accumulation_set_profiling", C.c_int, [C.c_void_p, C.c_int]),
    ("ssg_accumulation_get_profile", C.c_int, [C.c_void_p, _P(Profile), C.c_int]),
    ("ssg_default_eccflow_options", None, [_P(EccflowOptions)]),
    ("ssg_eccflow_create", C.c_int, [_P(EccflowOptions), _P(C.c_void_p)]),
    ("ssg_eccflow_destroy", None, [C.c_void_p]),
    ("ssg_eccflow_set_reference_image", C.c_int, [C.c_void_p, _P(Image), _P(Image)]),
    ("ssg_eccflow_num_levels", C.c_int, [C.c_void_p, _P(C.c_int32)]),
    ("ssg_eccflow_level_size", C.c_int, [C.c_void_p, C.c_int32, _P(C.c_int32), _P(C.c_int32)]),
    ("ssg_eccflow_reference_plane", C.c_int, [C.c_void_p, C.c_int32, C.c_int32, _P(Image)]),
    ("ssg_eccflow_compute", C.c_int, [C.c_void_p, _P(Image), _P(Image), _P(Image), _P(Image)]),
    ("ssg_eccflow_compute_batch", C.c_int, [C.c_void_p, C.c_int, _P(Image), _P(Image), _P(Image), _P(Image)]),
    ("ssg_eccflow_set_profiling", C.c_int, [C.c_void_p, C.c_int]),
    ("ssg_eccflow_get_profile", C.c_int, [C.c_void_p, _P(Profile), C.c_int]),
    ("ssg_synth_render", C.c_int, [_P(SynthParams), C.c_int, C.c_int, _P(Image), _P(C.c_double)]),
]

ENTRY_POINTS = [s[0] for s in _SIGS]

_lib = None


def load(path: str = LIB_PATH):
    """Loads libssg.so; raises when it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise SsgError(f"{path} not built: run serstacker_amd/csrc/build.py (or __graft_entry__.build())")
    lib = C.CDLL(path)
    for name, res, args in _SIGS:
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def check(rc: int, what: str = ""):
    if rc != SSG_OK:
        msg = load().ssg_last_error().decode(errors="replace")
        raise SsgError(f"{what}: error {rc}: {msg}")


# ---- image views -------------------------------------------------------------------------------
_NP_TYPE = {(np.dtype(np.uint8), 1): SSG_8UC1, (np.dtype(np.uint16), 1): SSG_16UC1,
            (np.dtype(np.int32), 1): SSG_32SC1, (np.dtype(np.int32), 3): SSG_32SC3,
            (np.dtype(np.float32), 1): SSG_32FC1,
            (np.dtype(np.float32), 2): SSG_32FC2, (np.dtype(np.float32), 3): SSG_32FC3,
            (np.dtype(np.float64), 1): SSG_64FC1, (np.dtype(np.float64), 3): SSG_64FC3,
            (np.dtype(np.uint8), 3): SSG_8UC3, (np.dtype(np.uint16), 3): SSG_16UC3}
TYPE_NP = {v: k for k, v in _NP_TYPE.items()}


def rows_contiguous(a: np.ndarray) -> bool:
    """True when each row of a (H, W) or (H, W, C) array is one dense run of memory
    (the pixels and their channels), rows at a positive step: an ssg_image view."""
    if a.ndim not in (2, 3) or a.strides[0] <= 0:
        return False
    cn = 1 if a.ndim == 2 else a.shape[2]
    if a.ndim == 3 and a.strides[2] != a.itemsize and cn > 1:
        return False
    return a.shape[1] <= 1 or a.strides[1] == cn * a.itemsize


def host_image(a: np.ndarray) -> Image:
    """View of a host array (H, W) or (H, W, C) whose rows are contiguous; the step
    is the array's row stride.  The caller keeps `a` alive while the view is used."""
    if not rows_contiguous(a):
        raise ValueError("array rows must be contiguous")
    cn = 1 if a.ndim == 2 else a.shape[2]
    t = _NP_TYPE[(a.dtype, cn)]
    return Image(a.ctypes.data, a.shape[1], a.shape[0], a.strides[0], t, 0)


def device_image(t) -> Image:
    """View of a contiguous torch tensor on the GPU, shape (H, W) or (H, W, C)."""
    import torch

    if not t.is_cuda:
        raise ValueError("tensor must be on the GPU")
    cn = 1 if t.dim() == 2 else t.shape[2]
    dts = {torch.uint8: np.uint8, torch.int16: np.uint16, torch.int32: np.int32, torch.float32: np.float32,
           torch.float64: np.float64}
    if hasattr(torch, "uint16"):
        dts[torch.uint16] = np.uint16
    dt = dts[t.dtype]
    if t.dtype == torch.int16:
        dt = np.uint16
    ty = _NP_TYPE[(np.dtype(dt), cn)]
    if t.stride(-1) != (1 if cn == 1 else 1) or (cn > 1 and t.stride(1) != cn):
        raise ValueError("tensor rows must be contiguous")
    return Image(t.data_ptr(), t.shape[1], t.shape[0], t.stride(0) * t.element_size(), ty, 1)


def images(views) -> C.Array:
    """ctypes array of views; it keeps the views (and the arrays they own) alive."""
    arr = (Image * len(views))()
    for i, v in enumerate(views):
        arr[i] = v
    arr._keep = list(views)
    return arr

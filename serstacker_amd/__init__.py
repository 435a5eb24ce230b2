"""serstacker_amd -- MI355X (gfx950) align+stack hot path of SerStacker.

Python mirror of the reference's registration and accumulation interfaces,
calling the HIP implementation through the C ABI of include/ssg.h:

    c_image_registration_options   c_frame_registration.h:119-136
    c_frame_registration           c_frame_registration.h:210-354
    c_weigthed_average             core/average/c_frame_accumulation.h:40-70

Images are numpy arrays (host) or torch tensors on the GPU (device); like the
reference, operations return False and keep an error message on failure.
"""
from __future__ import annotations

import ctypes as C
import os
import logging

import numpy as np

from . import _lib
from ._lib import (  # noqa: F401
    ACC_AVERAGE, ACC_WEIGHTED_AVERAGE, BORDER_CONSTANT, BORDER_REFLECT101, BORDER_REPLICATE,
    ECC_ALIGN_FORWARD_ADDITIVE, INTER_CUBIC, INTER_LINEAR, MOTION_AFFINE, MOTION_EUCLIDEAN,
    MOTION_SCALED_EUCLIDEAN, MOTION_TRANSLATION, RegistrationStatus, SsgError,
)

log = logging.getLogger("serstacker_amd")

__all__ = ["c_image_registration_options", "c_frame_registration", "c_weigthed_average", "image_view",
           "c_eccflow", "synth_params", "synth_render", "SsgError", "c_ser_reader", "debayer", "read_input_frames",
           "inpaint_linear", "average_pyramid_inpaint", "unsharp_mask", "local_variance_map", "upscale_size",
           "upscale_remaps", "upscale_images",
           "master_frame_quality", "master_scan_window", "select_master_frame", "master_stack_window",
           "ecc_normalize", "focus_stack_batch"]


def _is_torch(x) -> bool:
    return type(x).__module__.startswith("torch")


def image_view(x, out: bool = False) -> _lib.Image:
    """ssg_image view of a numpy array or a torch GPU tensor.

    Arrays whose rows are contiguous (an ROI crop, a row slice) are viewed in place
    with their row step.  Other host arrays are copied for inputs; the copy is kept
    alive by the view (`view._owner`).  An output (`out=True`) must be viewable in
    place, else ValueError: a copy would never reach the caller's array."""
    if x is None:
        return None
    if _is_torch(x):
        return _lib.device_image(x)
    a = np.asarray(x)
    if not _lib.rows_contiguous(a):
        if out:
            raise ValueError("output array rows must be contiguous (last axis, and channels, dense)")
        a = np.ascontiguousarray(a)
    v = _lib.host_image(a)
    v._owner = a
    return v


def image_views(frames, out: bool = False):
    """ctypes array of ssg_image views of a list of frames; a ctypes array of
    views passes through unchanged (build it once for repeated batch calls)."""
    if isinstance(frames, C.Array):
        return frames
    return _lib.images([image_view(f, out) for f in frames])


def c_image_registration_options(**ecc) -> _lib.ImageRegistrationOptions:
    """The reference defaults (c_frame_registration.h:47-136), ECC enabled with the
    forward-additive method; keyword arguments override fields of .ecc."""
    lib = _lib.load()
    o = _lib.ImageRegistrationOptions()
    lib.ssg_default_image_registration_options(C.byref(o))
    o.enable_feature_registration = 0
    o.enable_ecc_registration = 1
    o.ecc.ecc_method = _lib.ECC_ALIGN_FORWARD_ADDITIVE
    for k, v in ecc.items():
        if hasattr(o.ecc, k):
            setattr(o.ecc, k, v)
        else:
            setattr(o, k, v)
    return o


class c_frame_registration:
    def __init__(self, options: _lib.ImageRegistrationOptions | None = None):
        self._lib = _lib.load()
        self._opts = options if options is not None else c_image_registration_options()
        h = C.c_void_p()
        _lib.check(self._lib.ssg_registration_create(C.byref(self._opts), C.byref(h)), "ssg_registration_create")
        self._h = h
        self._status = _lib.RegistrationStatus()
        self._ref_size = None
        self.last_error = ""

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            self._lib.ssg_registration_destroy(h)
            self._h = None

    def options(self):
        return self._opts

    def _fail(self, rc, what):
        self.last_error = self._lib.ssg_last_error().decode(errors="replace")
        log.error("%s failed (%d): %s", what, rc, self.last_error)
        return False

    def setup_reference_frame(self, image, mask=None) -> bool:
        iv = image_view(image)
        mv = image_view(mask)
        rc = self._lib.ssg_registration_setup_reference_frame(self._h, C.byref(iv), C.byref(mv) if mv else None)
        if rc:
            return self._fail(rc, "setup_reference_frame")
        self._ref_size = (iv.width, iv.height)
        return True

    def register_frame(self, src, srcmask=None) -> bool:
        """Plain per-frame form (host sync per ECC iteration).  True when the
        frame passed the rho gates; status() holds the details."""
        iv = image_view(src)
        mv = image_view(srcmask)
        rc = self._lib.ssg_registration_register_frame(self._h, C.byref(iv), C.byref(mv) if mv else None,
                                                      C.byref(self._status))
        if rc:
            return self._fail(rc, "register_frame")
        return bool(self._status.ok)

    def register_batch(self, frames, masks=None):
        """Batched registration of device frames; returns a ctypes array of statuses."""
        views = image_views(frames)
        mviews = _lib.images([image_view(m) for m in masks]) if masks is not None else None
        st = (_lib.RegistrationStatus * len(frames))()
        rc = self._lib.ssg_registration_register_batch(self._h, len(frames), views, mviews, st)
        if rc:
            self._fail(rc, "register_batch")
            raise SsgError(self.last_error)
        return st

    def stack_batch(self, acc, frames, interpolation=_lib.INTER_LINEAR, border_mode=_lib.BORDER_REFLECT101,
                    border_value=0.0, use_graph=True):
        """ssg_stack_batch: weights, registration and the fused warp + weighted accumulate of a
        batch of device frames into `acc` (a c_weigthed_average), every per-frame decision on the
        device; use_graph replays a captured hipGraph of the launch sequence.  Returns the statuses."""
        views = image_views(frames)
        st = (_lib.RegistrationStatus * len(views))()
        rc = self._lib.ssg_stack_batch(self._h, acc._h, len(views), views, interpolation, border_mode, border_value,
                                       int(bool(use_graph)), st)
        if rc:
            self._fail(rc, "stack_batch")
            raise SsgError(self.last_error)
        return st

    def register_batch_remaps(self, frames, remaps, masks=None):
        """Batched registration that also writes each frame's remap (device tensors
        (H, W, 2)), refined by the dense ECC flow when eccflow registration is enabled."""
        views = image_views(frames)
        mviews = _lib.images([image_view(m) for m in masks]) if masks is not None else None
        rviews = image_views(remaps)
        st = (_lib.RegistrationStatus * len(frames))()
        rc = self._lib.ssg_registration_register_batch_remaps(self._h, len(frames), views, mviews, st, rviews)
        if rc:
            self._fail(rc, "register_batch_remaps")
            raise SsgError(self.last_error)
        return st

    def status(self) -> _lib.RegistrationStatus:
        return self._status

    def current_remap(self, status=None) -> np.ndarray:
        """The remap of a status's transform; without a status, c_frame_registration::
        current_remap() -- the last registered frame's field, eccflow-refined if enabled."""
        w, h = self._ref_size
        m = np.zeros((h, w, 2), np.float32)
        mv = _lib.host_image(m)
        if status is None:
            _lib.check(self._lib.ssg_registration_current_remap(self._h, C.byref(mv)), "current_remap")
        else:
            _lib.check(self._lib.ssg_registration_create_remap(self._h, C.byref(status), C.byref(mv)),
                       "create_remap")
        return m

    def remap(self, src, src_mask=None, interpolation=-1, border_mode=-1, border_value=0.0, want_mask=True):
        """c_frame_registration::remap (c_frame_registration.cc:1405-1417): warps src with the
        current remap of the last registered frame.  Returns (dst, dst_mask) numpy arrays."""
        w, h = self._ref_size
        cn = 1 if np.ndim(src) == 2 else src.shape[2]
        dst = np.zeros((h, w) if cn == 1 else (h, w, cn), np.float32)
        dm = np.zeros((h, w), np.uint8) if want_mask else None
        sv, smv = image_view(src), image_view(src_mask)
        dv = _lib.host_image(dst)
        dmv = _lib.host_image(dm) if want_mask else None
        _lib.check(self._lib.ssg_registration_remap(self._h, C.byref(sv), C.byref(smv) if smv else None, C.byref(dv),
                                                    C.byref(dmv) if dmv else None, interpolation, border_mode,
                                                    border_value), "remap")
        return dst, dm

    def custom_remap(self, src, src_mask=None, interpolation=None, border_mode=None, border_value=0.0,
                     status=None, rmap=None, want_mask=True):
        """c_frame_registration::custom_remap (c_frame_registration.cc:1265-1404).
        Returns (dst, dst_mask) as numpy arrays."""
        st = status if status is not None else self._status
        interp = self._opts.interpolation if interpolation is None else interpolation
        border = self._opts.border_mode if border_mode is None else border_mode
        src_np = src if not _is_torch(src) else None
        h, w = (self._ref_size[1], self._ref_size[0])
        cn = 1 if np.ndim(src) == 2 else src.shape[2]
        dst = np.zeros((h, w) if cn == 1 else (h, w, cn), np.float32)
        dm = np.zeros((h, w), np.uint8) if want_mask else None
        sv = image_view(src if src_np is None else np.ascontiguousarray(src_np, np.float32))
        smv = image_view(src_mask)
        dv = _lib.host_image(dst)
        dmv = _lib.host_image(dm) if want_mask else None
        mapv = _lib.host_image(np.ascontiguousarray(rmap, np.float32)) if rmap is not None else None
        params = (C.c_float * 6)(*st.params)
        rc = self._lib.ssg_remap(C.byref(sv), C.byref(smv) if smv else None, C.byref(mapv) if mapv else None, None,
                                 self._opts.motion_type, params, C.byref(dv), C.byref(dmv) if dmv else None, interp,
                                 border, border_value)
        _lib.check(rc, "ssg_remap")
        return dst, dm


class _accumulator_state:
    """Checkpoint / resume of an accumulator handle (ssg_accumulation_state_*)."""

    def save_state(self) -> bytes:
        n = C.c_uint64()
        _lib.check(self._lib.ssg_accumulation_state_size(self._h, C.byref(n)), "state_size")
        buf = C.create_string_buffer(n.value)
        _lib.check(self._lib.ssg_accumulation_state_save(self._h, buf, n.value), "state_save")
        return buf.raw

    def load_state(self, state: bytes):
        _lib.check(self._lib.ssg_accumulation_state_load(self._h, state, len(state)), "state_load")

    def clear(self):
        """c_frame_accumulation::clear (c_frame_accumulation.h:30)."""
        _lib.check(self._lib.ssg_accumulation_clear(self._h), "clear")

    def reinitialize(self, src, accw) -> bool:
        """c_frame_accumulation::reinitialize(src, accw) (c_frame_accumulation.h:25): a weighted
        average restarts from the mean `src` with weights `accw` (c_frame_accumulation.cc:171-180);
        Bayer and focus accumulators return False, as the reference's do."""
        sv, wv = image_view(src), image_view(accw)
        rc = self._lib.ssg_accumulation_reinitialize(self._h, C.byref(sv), C.byref(wv))
        if rc:
            log.error("reinitialize: %s", self._lib.ssg_last_error().decode())
            return False
        return True


class c_weigthed_average(_accumulator_state):
    """c_weigthed_average (core/average/c_frame_accumulation.h:40-70), sum form."""

    def __init__(self, width: int, height: int, channels: int = 1, method: int = _lib.ACC_WEIGHTED_AVERAGE):
        self._lib = _lib.load()
        h = C.c_void_p()
        _lib.check(self._lib.ssg_accumulation_create(method, width, height, channels, C.byref(h)),
                   "ssg_accumulation_create")
        self._h = h
        self.w, self.h, self.cn = width, height, channels

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            self._lib.ssg_accumulation_destroy(h)
            self._h = None

    def add(self, src, weights=None) -> bool:
        sv = image_view(src)
        wv = image_view(weights)
        rc = self._lib.ssg_accumulation_add(self._h, C.byref(sv), C.byref(wv) if wv else None)
        if rc:
            log.error("add: %s", self._lib.ssg_last_error().decode())
            return False
        return True

    def add_registered_batch(self, frames, statuses, motion_type, frame_weights=None,
                             interpolation=_lib.INTER_LINEAR, border_mode=_lib.BORDER_REFLECT101, border_value=0.0,
                             weight_maps=None):
        views = image_views(frames)
        wts = (C.c_float * len(frames))(*frame_weights) if frame_weights is not None else None
        wmv = image_views(weight_maps) if weight_maps is not None else None
        rc = self._lib.ssg_accumulation_add_registered_batch(self._h, len(frames), views, statuses, motion_type, wts,
                                                             wmv, interpolation, border_mode, border_value)
        _lib.check(rc, "add_registered_batch")

    def add_remapped_batch(self, frames, rmaps, frame_weights=None, interpolation=_lib.INTER_LINEAR,
                           border_mode=_lib.BORDER_REFLECT101, border_value=0.0, weight_maps=None):
        """The fused warp + accumulate with a dense remap field per frame (device (H, W, 2))."""
        views = image_views(frames)
        mv = image_views(rmaps)
        wts = (C.c_float * len(frames))(*frame_weights) if frame_weights is not None else None
        wmv = image_views(weight_maps) if weight_maps is not None else None
        rc = self._lib.ssg_accumulation_add_remapped_batch(self._h, len(frames), views, mv, wts, wmv, interpolation,
                                                           border_mode, border_value)
        _lib.check(rc, "add_remapped_batch")

    def compute(self, dscale: float = 1.0):
        avg = np.zeros((self.h, self.w) if self.cn == 1 else (self.h, self.w, self.cn), np.float32)
        mask = np.zeros((self.h, self.w), np.uint8)
        av, mv = _lib.host_image(avg), _lib.host_image(mask)
        _lib.check(self._lib.ssg_accumulation_compute(self._h, C.byref(av), C.byref(mv), dscale), "compute")
        return avg, mask

    def get_acc_counters(self) -> np.ndarray:
        w = np.zeros((self.h, self.w), np.float32)
        v = _lib.host_image(w)
        _lib.check(self._lib.ssg_accumulation_get_acc_counters(self._h, C.byref(v)), "get_acc_counters")
        return w

    def counts(self) -> np.ndarray:
        c = np.zeros((self.h, self.w), np.int32)
        v = _lib.host_image(c)
        _lib.check(self._lib.ssg_accumulation_get_counts(self._h, C.byref(v)), "get_counts")
        return c

    def accumulated_frames(self) -> int:
        n = C.c_int32()
        _lib.check(self._lib.ssg_accumulation_accumulated_frames(self._h, C.byref(n)), "accumulated_frames")
        return n.value

    def clear(self):
        _lib.check(self._lib.ssg_accumulation_clear(self._h), "clear")


class c_bayer_average(_accumulator_state):
    """c_bayer_average (core/average/c_frame_accumulation.h:222-262): raw Bayer frames
    accumulated per channel (BGR) through an optional remap."""

    def __init__(self, width: int, height: int, colorid: int = _lib.COLORID_BAYER_RGGB, bpp: int = 16):
        self._lib = _lib.load()
        h = C.c_void_p()
        _lib.check(self._lib.ssg_accumulation_create(_lib.ACC_BAYER_AVERAGE, width, height, 1, C.byref(h)),
                   "ssg_accumulation_create")
        self._h = h
        self.w, self.h = width, height
        self.set_bayer_pattern(colorid)
        _lib.check(self._lib.ssg_accumulation_set_input_bpp(self._h, bpp), "set_input_bpp")

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            self._lib.ssg_accumulation_destroy(h)
            self._h = None

    def set_bayer_pattern(self, colorid: int):
        _lib.check(self._lib.ssg_accumulation_set_bayer_pattern(self._h, colorid), "set_bayer_pattern")

    def set_remap(self, rmap=None):
        v = image_view(rmap) if rmap is not None else None
        _lib.check(self._lib.ssg_accumulation_set_remap(self._h, C.byref(v) if v else None), "set_remap")

    def add(self, src, weights=None) -> bool:
        sv = image_view(src)
        wv = image_view(weights)
        rc = self._lib.ssg_accumulation_add(self._h, C.byref(sv), C.byref(wv) if wv else None)
        if rc:
            log.error("add: %s", self._lib.ssg_last_error().decode())
            return False
        return True

    def add_registered_batch(self, frames, statuses, motion_type):
        views = image_views(frames)
        rc = self._lib.ssg_accumulation_add_registered_batch(self._h, len(views), views, statuses, motion_type, None,
                                                             None, _lib.INTER_LINEAR, _lib.BORDER_REFLECT101, 0.0)
        _lib.check(rc, "add_registered_batch")

    def compute(self):
        avg = np.zeros((self.h, self.w, 3), np.float32)
        mask = np.zeros((self.h, self.w), np.uint8)
        av, mv = _lib.host_image(avg), _lib.host_image(mask)
        _lib.check(self._lib.ssg_accumulation_compute(self._h, C.byref(av), C.byref(mv), 1.0), "compute")
        return avg, mask

    def get_acc_counters(self) -> np.ndarray:
        w = np.zeros((self.h, self.w, 3), np.float32)
        v = _lib.host_image(w)
        _lib.check(self._lib.ssg_accumulation_get_acc_counters(self._h, C.byref(v)), "get_acc_counters")
        return w

    def counts(self) -> np.ndarray:
        c = np.zeros((self.h, self.w, 3), np.int32)
        v = _lib.host_image(c)
        _lib.check(self._lib.ssg_accumulation_get_counts(self._h, C.byref(v)), "get_counts")
        return c

    def accumulated_frames(self) -> int:
        n = C.c_int32()
        _lib.check(self._lib.ssg_accumulation_accumulated_frames(self._h, C.byref(n)), "accumulated_frames")
        return n.value

    def clear(self):
        _lib.check(self._lib.ssg_accumulation_clear(self._h), "clear")


class phase_correlation:
    """FFT phase-correlation translation estimate (the pre-alignment of the turbulence
    and Bayer configurations; the reference has none).  estimate() returns, per frame,
    (Tx, Ty, peak) with ref(x) = frame(x + T)."""

    def __init__(self, width: int, height: int, binning: int = 1):
        self._lib = _lib.load()
        h = C.c_void_p()
        _lib.check(self._lib.ssg_phasecorr_create(width, height, binning, C.byref(h)), "ssg_phasecorr_create")
        self._h = h
        self.w, self.h, self.binning = width, height, binning

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            self._lib.ssg_phasecorr_destroy(h)
            self._h = None

    def set_reference(self, ref):
        v = image_view(ref)
        _lib.check(self._lib.ssg_phasecorr_set_reference(self._h, C.byref(v)), "set_reference")

    def estimate(self, frames) -> np.ndarray:
        views = image_views(frames)
        out = np.zeros((len(views), 3), np.float64)
        _lib.check(self._lib.ssg_phasecorr_estimate(self._h, len(views), views,
                                                    out.ctypes.data_as(C.POINTER(C.c_double))), "estimate")
        return out

    def statuses(self, shifts) -> C.Array:
        """Translation registration statuses (ok, params = T) for the accumulators."""
        st = (_lib.RegistrationStatus * len(shifts))()
        for i, (tx, ty, _) in enumerate(shifts):
            st[i].ok = 1
            st[i].nparams = 2
            st[i].params[0], st[i].params[1] = float(tx), float(ty)
            st[i].matrix[:] = [1.0, 0.0, float(np.float32(tx)), 0.0, 1.0, float(np.float32(ty))]
        return st


class c_eccflow:
    """c_eccflow (core/proc/image_registration/ecc2.h:548-662): dense coarse-to-fine ECC flow.
    Keyword options are the fields of c_eccflow_options (ecc2.h:515-527); compute() returns
    the refined remap identity + flow (H, W, 2)."""

    def __init__(self, **opts):
        self._lib = _lib.load()
        o = _lib.EccflowOptions()
        self._lib.ssg_default_eccflow_options(C.byref(o))
        for k, v in opts.items():
            if not hasattr(o, k):
                raise AttributeError(f"c_eccflow_options has no field {k!r}")
            setattr(o, k, v)
        self.options = o
        h = C.c_void_p()
        _lib.check(self._lib.ssg_eccflow_create(C.byref(o), C.byref(h)), "ssg_eccflow_create")
        self._h = h

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            self._lib.ssg_eccflow_destroy(h)
            self._h = None

    def set_reference_image(self, image, mask=None):
        v = image_view(image)
        m = image_view(mask)
        _lib.check(self._lib.ssg_eccflow_set_reference_image(self._h, C.byref(v), C.byref(m) if m else None),
                   "set_reference_image")

    def num_levels(self) -> int:
        n = C.c_int32()
        _lib.check(self._lib.ssg_eccflow_num_levels(self._h, C.byref(n)), "num_levels")
        return n.value

    def level_size(self, level: int):
        w, h = C.c_int32(), C.c_int32()
        _lib.check(self._lib.ssg_eccflow_level_size(self._h, level, C.byref(w), C.byref(h)), "level_size")
        return w.value, h.value

    PLANES = ("image", "Ix", "Iy", "a00", "a01", "a11", "idet")

    def reference_plane(self, level: int, which: str) -> np.ndarray:
        """A plane of the reference pyramid entry: image, Ix, Iy (level size) or a00, a01, a11,
        idet (the D cells)."""
        k = self.PLANES.index(which)
        w, h = self.level_size(level)
        if k >= 3:
            for _ in range(self.options.support_scale):
                w, h = (w + 1) // 2, (h + 1) // 2
        out = np.zeros((h, w), np.float32)
        v = _lib.host_image(out)
        _lib.check(self._lib.ssg_eccflow_reference_plane(self._h, level, k, C.byref(v)), "reference_plane")
        return out

    def compute(self, image, rmap=None, mask=None, out=None):
        """c_eccflow::compute(input, rmap, mask): the refined remap.  out: a device tensor
        (H, W, 2) to write into, else a host array is returned."""
        v = image_view(image)
        m = image_view(mask)
        r = image_view(rmap)
        if out is None:
            res = np.zeros((v.height, v.width, 2), np.float32)
            o = _lib.host_image(res)
        else:
            res, o = out, image_view(out, out=True)
        _lib.check(self._lib.ssg_eccflow_compute(self._h, C.byref(v), C.byref(m) if m else None,
                                                 C.byref(r) if r else None, C.byref(o)), "compute")
        return res

    def compute_batch(self, frames, outs, rmaps=None, masks=None):
        """Batched compute over device frames into device maps outs[i] (H, W, 2)."""
        fv = image_views(frames)
        ov = image_views(outs, out=True)
        rv = image_views(rmaps) if rmaps is not None else None
        mv = image_views(masks) if masks is not None else None
        _lib.check(self._lib.ssg_eccflow_compute_batch(self._h, len(fv), fv, mv, rv, ov), "compute_batch")

    def set_profiling(self, enable: bool = True):
        _lib.check(self._lib.ssg_eccflow_set_profiling(self._h, int(enable)), "set_profiling")

    def profile(self, reset: bool = False) -> _lib.Profile:
        p = _lib.Profile()
        _lib.check(self._lib.ssg_eccflow_get_profile(self._h, C.byref(p), int(reset)), "get_profile")
        return p


class c_laplacian_pyramid_focus_stacking(_accumulator_state):
    """c_laplacian_pyramid_focus_stacking (core/average/c_frame_accumulation.h:141-184);
    keyword options are the fields of ssg_focus_stacking_options."""

    def __init__(self, width: int, height: int, channels: int = 1, **opts):
        self._lib = _lib.load()
        h = C.c_void_p()
        _lib.check(self._lib.ssg_accumulation_create(_lib.ACC_FOCUS_STACK, width, height, channels, C.byref(h)),
                   "ssg_accumulation_create")
        self._h = h
        self.w, self.h, self.cn = width, height, channels
        o = _lib.FocusStackingOptions()
        self._lib.ssg_default_focus_stacking_options(C.byref(o))
        for k, v in opts.items():
            setattr(o, k, v)
        _lib.check(self._lib.ssg_accumulation_set_focus_options(self._h, C.byref(o)), "set_focus_options")

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            self._lib.ssg_accumulation_destroy(h)
            self._h = None

    def add(self, src) -> bool:
        sv = image_view(src)
        rc = self._lib.ssg_accumulation_add(self._h, C.byref(sv), None)
        if rc:
            log.error("add: %s", self._lib.ssg_last_error().decode())
            return False
        return True

    def add_frames(self, frames):
        """n successive add() calls, batched: pyramids built together, levels fused per pixel."""
        views = image_views(frames)
        _lib.check(self._lib.ssg_accumulation_add_frames(self._h, len(views), views), "add_frames")

    def compute(self, out=None):
        """The fused image; into `out` (host array or device tensor of the frame shape) when given."""
        avg = out if out is not None else np.zeros((self.h, self.w) if self.cn == 1 else (self.h, self.w, self.cn),
                                                   np.float32)
        av = image_view(avg, out=True)
        _lib.check(self._lib.ssg_accumulation_compute(self._h, C.byref(av), None, 1.0), "compute")
        return avg

    def accumulated_frames(self) -> int:
        n = C.c_int32()
        _lib.check(self._lib.ssg_accumulation_accumulated_frames(self._h, C.byref(n)), "accumulated_frames")
        return n.value


def focus_stack_batch(stacks, frames, shifts=None):
    """ssg_focus_stack_batch: several focus sequences at once.  stacks: list of
    c_laplacian_pyramid_focus_stacking (lpg-weight policy); frames[q]: list of device tensors
    of stack q's shape; shifts[q] (optional): (n, 2) translations (tx, ty) of the frames."""
    n = len(stacks)
    lib = _lib.load()
    handles = (C.c_void_p * n)(*[st._h.value for st in stacks])
    counts = (C.c_int32 * n)(*[len(f) for f in frames])
    views = [image_views(f) for f in frames]
    vptr = (C.POINTER(_lib.Image) * n)(*[C.cast(v, C.POINTER(_lib.Image)) for v in views])
    sh_arrays, sptr = [], None
    if shifts is not None:
        for sh in shifts:
            sh_arrays.append(None if sh is None else np.ascontiguousarray(sh, np.float32).reshape(-1))
        sptr = (C.POINTER(C.c_float) * n)(*[
            a.ctypes.data_as(C.POINTER(C.c_float)) if a is not None else C.POINTER(C.c_float)() for a in sh_arrays])
    _lib.check(lib.ssg_focus_stack_batch(n, handles, counts, vptr, sptr), "ssg_focus_stack_batch")


def lpg(src, k: float = 6.0, p: float = 2.0):
    """lpg() (core/proc/lpg.cc:277-344), dscale = uscale = 0: (map, mean sharpness)."""
    lib = _lib.load()
    h, w = src.shape[:2]
    out = np.zeros((h, w), np.float32)
    sv, ov = image_view(src), _lib.host_image(out)
    sh = C.c_double()
    _lib.check(lib.ssg_lpg(C.byref(sv), C.byref(ov), k, p, C.byref(sh)), "ssg_lpg")
    return out, sh.value


def local_variance_map(image, kradius: int = 1, dscale: int = 1):
    """compute_local_variance_map (c_local_variance_sharpness_measure.cc:193-249):
    (per-pixel weight map at the image size, quality Q)."""
    lib = _lib.load()
    h, w = image.shape[:2]
    out = np.zeros((h, w), np.float32)
    iv, ov = image_view(image), _lib.host_image(out)
    q = C.c_double()
    _lib.check(lib.ssg_local_variance_map(C.byref(iv), kradius, dscale, C.byref(ov), C.byref(q)),
               "ssg_local_variance_map")
    return out, q.value


def ecc_normalize(src, mask=None, lvl: int = 1, out=None):
    """ecc_normalize(src, mask, dst, lvl) (ecc2.cc:385-397) of an f32 image."""
    if out is None:
        out = np.empty(tuple(src.shape), np.float32)
    sv, ov = image_view(src), image_view(out, out=True)
    mv = image_view(mask) if mask is not None else None
    _lib.check(_lib.load().ssg_ecc_normalize(C.byref(sv), C.byref(mv) if mv is not None else None, C.byref(ov), lvl),
               "ssg_ecc_normalize")
    return out


def master_frame_quality(frames, colorid: int = 0, kradius: int = 1, dscale: int = 1) -> np.ndarray:
    """The per-frame metric of select_master_frame's best-of scan
    (c_image_stacking_pipeline_base.cc:350-359): average_bayer_planes of Bayer frames, then
    compute_local_variance_map's Q.  frames: raw 8/16-bit or f32 arrays / device tensors."""
    lib = _lib.load()
    views = image_views(frames)
    q = np.zeros(len(frames), np.float64)
    _lib.check(lib.ssg_master_frame_quality(len(frames), views, colorid, kradius, dscale,
                                            q.ctypes.data_as(C.POINTER(C.c_double))), "ssg_master_frame_quality")
    return q


def master_scan_window(size: int, max_frames_to_scan: int = 200):
    """(start_pos, end_pos) of master_frame_best_of_100_in_middle
    (c_image_stacking_pipeline_base.cc:312-326)."""
    a, b = C.c_int32(), C.c_int32()
    _lib.check(_lib.load().ssg_master_scan_window(size, max_frames_to_scan, C.byref(a), C.byref(b)),
               "ssg_master_scan_window")
    return a.value, b.value


def select_master_frame(quality, start_pos: int = 0) -> int:
    """The best-of rule (c_image_stacking_pipeline_base.cc:369-395): the global index of the first
    frame whose quality exceeds all earlier ones."""
    q = np.ascontiguousarray(quality, dtype=np.float64)
    idx = C.c_int32()
    _lib.check(_lib.load().ssg_select_master_frame(len(q), q.ctypes.data_as(C.POINTER(C.c_double)), start_pos,
                                                   C.byref(idx)), "ssg_select_master_frame")
    return idx.value


def master_stack_window(size: int, start_frame_index: int, master_frame_pos: int, max_frames_to_stack: int):
    """(startpos, endpos) of the frames create_reference_frame stacks around the master frame
    (c_image_stacking_pipeline.cc:1219-1232)."""
    a, b = C.c_int32(), C.c_int32()
    _lib.check(_lib.load().ssg_master_stack_window(size, start_frame_index, master_frame_pos, max_frames_to_stack,
                                                   C.byref(a), C.byref(b)), "ssg_master_stack_window")
    return a.value, b.value


def frame_sharpness(frames) -> np.ndarray:
    """Per-frame weight (mean squared 3x3 laplacian) of device frames."""
    lib = _lib.load()
    views = image_views(frames)
    out = np.zeros(len(frames), np.float32)
    _lib.check(lib.ssg_frame_sharpness(len(frames), views, out.ctypes.data_as(C.POINTER(C.c_float))),
               "ssg_frame_sharpness")
    return out


def synth_params(**kw) -> _lib.SynthParams:
    """Planetary-disk generator parameters (defaults: 1920x1080 affine frames, disk
    radius 420, texture sigma 3 at 10%, limb darkening 0.6)."""
    p = _lib.SynthParams()
    p.width, p.height = 1920, 1080
    p.disk_radius = 420.0
    p.texture_sigma = 3.0
    p.texture_amplitude = 0.1
    p.limb_darkening = 0.6
    p.max_shift = 20.0
    p.max_rotation_deg = 1.0
    p.max_scale_dev = 0.005
    p.max_shear = 0.002
    p.blur_sigma_min, p.blur_sigma_max = 0.8, 2.5
    p.peak_electrons = 2000.0
    p.read_noise = 4.0
    p.output_scale = 1.0 / 65536.0
    p.seed = 20261202
    for k, v in kw.items():
        setattr(p, k, v)
    return p


def synth_render(p: _lib.SynthParams, first: int, frames) -> np.ndarray:
    """Renders frames[i] = frame first+i (first = -1: the reference) into device
    tensors; returns the (n, 6) true ref->frame maps."""
    lib = _lib.load()
    views = image_views(frames)
    truth = np.zeros((len(frames), 6), np.float64)
    rc = lib.ssg_synth_render(C.byref(p), first, len(frames), views, truth.ctypes.data_as(C.POINTER(C.c_double)))
    _lib.check(rc, "ssg_synth_render")
    return truth


class c_ser_reader:
    """c_ser_reader (core/io/c_ser_file.h:136-180): SER frames read into host arrays or
    device tensors, in the file's pixel type, without debayer."""

    def __init__(self, filename=None):
        self._lib = _lib.load()
        self._h = None
        if filename is not None:
            self.open(filename)

    def open(self, filename) -> bool:
        self.close()
        h = C.c_void_p()
        rc = self._lib.ssg_ser_open(os.fsencode(filename), C.byref(h))
        if rc:
            log.error("open: %s", self._lib.ssg_last_error().decode())
            return False
        self._h = h
        self._hdr = _lib.SerHeader()
        t = C.c_int32()
        _lib.check(self._lib.ssg_ser_get_header(self._h, C.byref(self._hdr), C.byref(t)), "header")
        self.frame_type = t.value
        return True

    def is_open(self) -> bool:
        return bool(self._h)

    def close(self):
        if getattr(self, "_h", None):
            self._lib.ssg_ser_close(self._h)
        self._h = None

    __del__ = close

    def header(self) -> _lib.SerHeader:
        return self._hdr

    def image_width(self) -> int:
        return self._hdr.image_width

    def image_height(self) -> int:
        return self._hdr.image_height

    def color_id(self) -> int:
        return self._hdr.color_id

    def bits_per_plane(self) -> int:
        return self._hdr.bits_per_plane

    def num_frames(self) -> int:
        return self._hdr.frames_count

    def timestamps(self):
        n = C.c_int32()
        _lib.check(self._lib.ssg_ser_timestamps(self._h, None, 0, C.byref(n)), "timestamps")
        ts = (C.c_uint64 * max(1, n.value))()
        _lib.check(self._lib.ssg_ser_timestamps(self._h, ts, n.value, C.byref(n)), "timestamps")
        return np.array(ts[:n.value], np.uint64)

    def curpos(self) -> int:
        p = C.c_int32()
        _lib.check(self._lib.ssg_ser_curpos(self._h, C.byref(p)), "curpos")
        return p.value

    def seek(self, frame_index: int) -> bool:
        rc = self._lib.ssg_ser_seek(self._h, frame_index)
        if rc:
            log.error("seek: %s", self._lib.ssg_last_error().decode())
        return rc == 0

    def _shape_dtype(self):
        dt, cn = _lib.TYPE_NP[self.frame_type]
        shape = (self._hdr.image_height, self._hdr.image_width) + ((cn,) if cn > 1 else ())
        return shape, dt

    def read(self, out=None):
        """The next frame into `out` (host array or device tensor) or a new host array."""
        if out is None:
            shape, dt = self._shape_dtype()
            out = np.empty(shape, dt)
        v = image_view(out, out=True)
        _lib.check(self._lib.ssg_ser_read(self._h, C.byref(v)), "read")
        return out

    def read_batch(self, outs):
        """len(outs) successive frames into device tensors."""
        views = image_views(outs, out=True)
        _lib.check(self._lib.ssg_ser_read_batch(self._h, len(views), views), "read_batch")
        return outs


def debayer(src, colorid: int, algorithm: int = -1, out=None):
    """debayer() (core/io/debayer.cc:1197-1355) on the GPU; `out` of the result's shape and
    the source dtype (host array or device tensor), allocated on the host when omitted."""
    if out is None:
        h, w = src.shape[:2]
        half = algorithm in (_lib.DEBAYER_SP, _lib.DEBAYER_AVGBP) or \
            _lib.DEBAYER_PLANE_0 <= algorithm <= _lib.DEBAYER_PLANE_3
        cn = 3 if algorithm in (-1, _lib.DEBAYER_NN, _lib.DEBAYER_NN2, _lib.DEBAYER_VNG, _lib.DEBAYER_EA,
                                _lib.DEBAYER_SP, _lib.DEBAYER_MATRIX) else 1
        shape = ((h // 2, w // 2) if half else (h, w)) + ((3,) if cn == 3 else ())
        dt = src.cpu().numpy().dtype if _is_torch(src) else np.asarray(src).dtype
        out = np.empty(shape, dt)
    sv, ov = image_view(src), image_view(out, out=True)
    _lib.check(_lib.load().ssg_debayer(C.byref(sv), colorid, algorithm, C.byref(ov)), "debayer")
    return out


def read_input_frames(raws, colorid: int, bpp: int, algorithm: int = -1, dark=None, flat=None, outs=None):
    """The raw-frame conversion of read_input_frame (c_image_stacking_pipeline_base.cc:124-279)
    for a list of frames; outs: f32 host arrays or device tensors of the debayered shape."""
    rv = image_views(raws)
    ov = image_views(outs, out=True)
    dv, fv = image_view(dark), image_view(flat)
    _lib.check(_lib.load().ssg_read_input_frames(len(rv), rv, colorid, bpp, algorithm,
                                                 C.byref(dv) if dv else None, C.byref(fv) if fv else None, ov),
               "read_input_frames")
    return outs


def inpaint_linear(image, mask):
    """linear_interpolation_inpaint(image, mask) (core/proc/inpaint/linear_interpolation_inpaint.cc:
    335-361): a device tensor is filled in place; a host array is returned filled (a copy)."""
    img = image if _is_torch(image) else np.array(image, dtype=np.float32, copy=True)
    iv, mv = image_view(img), image_view(mask)
    _lib.check(_lib.load().ssg_inpaint_linear(C.byref(iv), C.byref(mv)), "inpaint_linear")
    return img


def average_pyramid_inpaint(src, mask, max_levels: int = 100, out=None, out_mask=None):
    """average_pyramid_inpaint(src, mask, dst, dstmask, max_levels)
    (core/proc/inpaint/average_pyramid_inpaint.cc:96-126): returns (image, mask); host arrays are
    allocated for the outputs that are not given."""
    h, w = src.shape[:2]
    if out is None:
        out = np.empty(tuple(src.shape), np.float32)
    if out_mask is None:
        out_mask = np.empty((h, w), np.uint8)
    sv, ov, dmv = image_view(src), image_view(out, out=True), image_view(out_mask, out=True)
    mv = image_view(mask) if mask is not None else None
    _lib.check(_lib.load().ssg_average_pyramid_inpaint(C.byref(sv), C.byref(mv) if mv is not None else None,
                                                       C.byref(ov), C.byref(dmv), max_levels),
               "average_pyramid_inpaint")
    return out, out_mask


def upscale_size(option: int, w: int, h: int):
    """(dw, dh) of a frame_upscale_option (c_image_stacking_pipeline.cc:1869-1905)."""
    a, b = C.c_int32(), C.c_int32()
    _lib.check(_lib.load().ssg_upscale_size(option, w, h, C.byref(a), C.byref(b)), "ssg_upscale_size")
    return a.value, b.value


def upscale_remaps(option: int, maps, out=None):
    """upscale_remap (c_image_stacking_pipeline.cc:1869-1905) of (H, W, 2) f32 remaps: host
    arrays or device tensors; `out` a list of the upscaled size, host arrays when omitted."""
    h, w = maps[0].shape[:2]
    dw, dh = upscale_size(option, w, h)
    if out is None:
        out = [np.empty((dh, dw, 2), np.float32) for _ in maps]
    sv, dv = image_views(maps), image_views(out, out=True)
    _lib.check(_lib.load().ssg_upscale_remaps(option, len(maps), sv, dv), "ssg_upscale_remaps")
    return out


def upscale_images(option: int, frames, masks=None, out=None, out_masks=None):
    """upscale_image (c_image_stacking_pipeline.cc:1949-2000) of f32 frames ((H, W) or (H, W, 3),
    host arrays or device tensors) and their optional u8 masks (a list; None entries for frames
    without one).  Returns (frames, masks) at the upscaled size: host arrays when `out` /
    `out_masks` are omitted, masks None where a frame had none."""
    h, w = frames[0].shape[:2]
    dw, dh = upscale_size(option, w, h)
    if out is None:
        out = [np.empty((dh, dw) + tuple(frames[0].shape[2:]), np.float32) for _ in frames]
    sv, dv = image_views(frames), image_views(out, out=True)
    mv = dmv = None
    if masks is not None:
        if out_masks is None:
            out_masks = [None if m is None else np.empty((dh, dw), np.uint8) for m in masks]
        empty = _lib.Image()  # data NULL: this frame has no mask
        mv = _lib.images([empty if m is None else image_view(m) for m in masks])
        dmv = _lib.images([empty if m is None else image_view(om, out=True) for m, om in zip(masks, out_masks)])
    _lib.check(_lib.load().ssg_upscale_images(option, len(frames), sv, mv, dv, dmv), "ssg_upscale_images")
    return out, (out_masks if masks is not None else None)


def unsharp_mask(src, sigma: float, alpha: float, out=None):
    """unsharp_mask(src, dst, sigma, alpha) (core/proc/unsharp_mask.cc:72-115) of an f32 image."""
    if out is None:
        out = np.empty(tuple(src.shape), np.float32) if not _is_torch(src) else src.clone()
    sv, ov = image_view(src), image_view(out, out=True)
    _lib.check(_lib.load().ssg_unsharp_mask(C.byref(sv), C.byref(ov), sigma, alpha), "unsharp_mask")
    return out


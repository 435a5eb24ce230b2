"""ctypes front-end of the CPU oracle (test infrastructure only).

Only tests/, ``__graft_entry__.smoke()`` and bench.py's ``cpu_baseline`` leg may
import this module.  It is the checker, never the thing measured or shipped:
the product package ``serstacker_amd`` does not import it.

``Oracle(64)`` loads the FP64 build, ``Oracle(32)`` the FP32 reference-faithful
build (same C source, see ssg_oracle.h).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))

MOTION_TRANSLATION, MOTION_EUCLIDEAN, MOTION_SCALED_EUCLIDEAN, MOTION_AFFINE = 0, 1, 2, 3
ECC_FORWARD_ADDITIVE, ECC_INVERSE_COMPOSITIONAL, ECC_LM, ECC_INVERSE_COMPOSITIONAL_LM = 0, 1, 2, 3
BORDER_CONSTANT, BORDER_REPLICATE, BORDER_REFLECT, BORDER_WRAP, BORDER_REFLECT101, BORDER_TRANSPARENT = range(6)
INTER_NEAREST, INTER_LINEAR, INTER_CUBIC, INTER_AREA = 0, 1, 2, 3


class RegOptions(C.Structure):
    _fields_ = [
        ("motion_type", C.c_int32),
        ("interpolation", C.c_int32),
        ("border_mode", C.c_int32),
        ("border_value", C.c_double),
        ("scale", C.c_double),
        ("eps", C.c_double),
        ("min_rho", C.c_double),
        ("input_smooth_sigma", C.c_double),
        ("reference_smooth_sigma", C.c_double),
        ("update_step_scale", C.c_double),
        ("max_iterations", C.c_int32),
        ("ecch_max_level", C.c_int32),
        ("ecch_minimum_image_size", C.c_int32),
        ("ecch_estimate_translation_first", C.c_int32),
        ("ecc_method", C.c_int32),
    ]


class RegStatus(C.Structure):
    _fields_ = [
        ("ok", C.c_int32),
        ("code", C.c_int32),
        ("rho_translation", C.c_double),
        ("rho", C.c_double),
        ("iterations_translation", C.c_int32),
        ("iterations", C.c_int32),
        ("levels_failed", C.c_int32),
        ("nparams", C.c_int32),
        ("params", C.c_double * 6),
        ("matrix", C.c_double * 6),
        ("level_iterations", C.c_int32 * 16),
    ]


class LevelResult(C.Structure):
    _fields_ = [("num_iterations", C.c_int32), ("failed", C.c_int32), ("eps", C.c_double)]


def reg_options(**kw) -> RegOptions:
    """Reference defaults (c_frame_registration.h:47-64, :119-136) with the
    forward-additive method the hot path pins; keyword overrides."""
    o = RegOptions()
    o.motion_type = MOTION_AFFINE
    o.interpolation = INTER_LINEAR
    o.border_mode = BORDER_REFLECT101
    o.border_value = 0.0
    o.scale = 0.5
    o.eps = 0.2
    o.min_rho = 0.8
    o.input_smooth_sigma = 1.0
    o.reference_smooth_sigma = 1.0
    o.update_step_scale = 1.5
    o.max_iterations = 50
    o.ecch_max_level = 0
    o.ecch_minimum_image_size = 16
    o.ecch_estimate_translation_first = 1
    o.ecc_method = ECC_FORWARD_ADDITIVE
    for k, v in kw.items():
        setattr(o, k, v)
    return o


class XForm(C.Structure):
    pass


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


_libs: dict = {}


def _load(bits: int):
    name = os.path.join(HERE, f"liboracle_f{bits}.so")
    if not os.path.exists(name):
        build()
    lib = C.CDLL(name, mode=C.RTLD_LOCAL)
    return lib


def build_native(out_dir: str) -> dict:
    """The host-tuned FP32 builds (-march=native, with and without -ffast-math) into out_dir, a
    directory of the machine that runs them; returns {variant: library path}."""
    subprocess.run(["make", "-s", "-C", HERE, "native", f"NATIVE_DIR={out_dir}"], check=True)
    return {"native": os.path.join(out_dir, "liboracle_f32_native.so"),
            "native_fastmath": os.path.join(out_dir, "liboracle_f32_native_fastmath.so")}


class Oracle:
    def __init__(self, bits: int = 64, library: str | None = None):
        """bits: 32 (reference-faithful FP32) or 64; library: an FP32 build of the same source to
        load instead (a host-tuned variant from build_native, for timing)."""
        if bits not in (32, 64):
            raise ValueError("bits must be 32 or 64")
        key = library or bits
        if key not in _libs:
            _libs[key] = C.CDLL(library, mode=C.RTLD_LOCAL) if library else _load(bits)
        self.lib = _libs[key]
        self.bits = bits
        self.dt = np.float64 if bits == 64 else np.float32
        rp = C.POINTER(C.c_double if bits == 64 else C.c_float)
        self._rp = rp
        self.rctype = C.c_double if bits == 64 else C.c_float

        class _XForm(C.Structure):
            _fields_ = [
                ("motion", C.c_int32),
                ("nparams", C.c_int32),
                ("p", self.rctype * 6),
                ("C", self.rctype * 2),
                ("fixed_scale", self.rctype),
            ]

        self.XForm = _XForm
        L = self.lib
        u8p = C.POINTER(C.c_uint8)
        L.ora_real_size.restype = C.c_int
        assert L.ora_real_size() == (8 if bits == 64 else 4)
        L.ora_set_threads.argtypes = [C.c_int]
        L.ora_border_interpolate.argtypes = [C.c_int, C.c_int, C.c_int]
        L.ora_border_interpolate.restype = C.c_int
        L.ora_gaussian_kernel.argtypes = [C.c_int, C.c_double, rp]
        L.ora_sep_filter.argtypes = [rp, C.c_int, C.c_int, rp, rp, C.c_int, rp, C.c_int, C.c_int, self.rctype]
        L.ora_pyr_down.argtypes = [rp, C.c_int, C.c_int, rp, C.c_int, C.c_int, C.c_int]
        L.ora_pyr_down_u8.argtypes = [u8p, C.c_int, C.c_int, u8p, C.c_int, C.c_int]
        L.ora_pyr_up.argtypes = [rp, C.c_int, C.c_int, rp, C.c_int, C.c_int]
        L.ora_resize_nearest_u8.argtypes = [u8p, C.c_int, C.c_int, u8p, C.c_int, C.c_int]
        L.ora_remap.argtypes = [rp, C.c_int, C.c_int, rp, C.c_int, C.c_int, rp, C.c_int, C.c_int, self.rctype]
        L.ora_remap_u8.argtypes = [u8p, C.c_int, C.c_int, rp, C.c_int, C.c_int, u8p, C.c_int, C.c_uint8]
        L.ora_erode_rect_u8.argtypes = [u8p, C.c_int, C.c_int, C.c_int, C.c_int, u8p, C.c_int, C.c_uint8]
        L.ora_mean_stddev.argtypes = [rp, u8p, C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_double)]
        L.ora_cholesky_invert.argtypes = [rp, C.c_int]
        L.ora_cholesky_invert.restype = C.c_int
        L.ora_ecc_differentiate.argtypes = [rp, C.c_int, C.c_int, rp, rp]
        xp = C.POINTER(_XForm)
        L.ora_xform_reset.argtypes = [xp, C.c_int]
        L.ora_xform_create_map.argtypes = [xp, C.c_int, C.c_int, rp]
        L.ora_xform_scale.argtypes = [xp, C.c_double]
        L.ora_xform_eps.argtypes = [xp, rp, C.c_int, C.c_int]
        L.ora_xform_eps.restype = self.rctype
        L.ora_xform_matrix.argtypes = [xp, C.POINTER(C.c_double)]
        L.ora_ecc_fa_align.argtypes = [rp, u8p, C.c_int, C.c_int, rp, u8p, C.c_int, C.c_int, xp,
                                       C.c_int, C.c_int, C.c_double, C.c_double, C.POINTER(LevelResult)]
        L.ora_ecc_fa_align.restype = C.c_int
        L.ora_ecc_lm_align.argtypes = [rp, u8p, C.c_int, C.c_int, rp, u8p, C.c_int, C.c_int, xp,
                                       C.c_int, C.c_double, C.c_double, C.POINTER(LevelResult)]
        L.ora_ecc_lm_align.restype = C.c_int
        for fn in (L.ora_ecc_ic_align, L.ora_ecc_iclm_align):
            fn.argtypes = [rp, u8p, C.c_int, C.c_int, rp, u8p, C.c_int, C.c_int, xp, xp,
                           C.c_int, C.c_double, C.c_double, C.POINTER(LevelResult)]
            fn.restype = C.c_int
        L.ora_xform_invert_and_compose.argtypes = [xp, rp, rp]
        L.ora_compute_correlation.argtypes = [rp, u8p, C.c_int, C.c_int, rp, u8p, C.c_int, C.c_int, rp]
        L.ora_compute_correlation.restype = C.c_double
        L.ora_reg_create.argtypes = [C.POINTER(RegOptions)]
        L.ora_reg_create.restype = C.c_void_p
        L.ora_reg_destroy.argtypes = [C.c_void_p]
        L.ora_reg_setup_reference.argtypes = [C.c_void_p, rp, u8p, C.c_int, C.c_int]
        L.ora_reg_register.argtypes = [C.c_void_p, rp, u8p, C.c_int, C.c_int, C.POINTER(RegStatus)]
        L.ora_reg_num_levels.argtypes = [C.c_void_p]
        L.ora_reg_level_size.argtypes = [C.c_void_p, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int)]
        L.ora_reg_get_ref_level.argtypes = [C.c_void_p, C.c_int, rp, u8p]
        L.ora_base_remap.argtypes = [rp, C.c_int, C.c_int, C.c_int, u8p, rp, C.c_int, C.c_int, rp, u8p,
                                     C.c_int, C.c_int, self.rctype]
        L.ora_weighted_average_update.argtypes = [rp, C.c_int, C.c_int, C.c_int, u8p, rp, rp, rp]
        L.ora_sharpness_mean_sq_laplacian.argtypes = [rp, C.c_int, C.c_int]
        L.ora_sharpness_mean_sq_laplacian.restype = C.c_double
        L.ora_bayer_channel.argtypes = [C.c_int, C.c_int, C.c_int]
        L.ora_bayer_channel.restype = C.c_int
        L.ora_bayer_accumulate.argtypes = [rp, C.c_int, C.c_int, rp, C.c_int, u8p, rp, rp, rp,
                                           C.POINTER(C.c_int32)]
        L.ora_bayer_compute.argtypes = [rp, rp, C.c_int, C.c_int, rp, u8p]

    # -- helpers -------------------------------------------------------------
    def _r(self, a):
        a = np.ascontiguousarray(a, dtype=self.dt)
        return a, a.ctypes.data_as(self._rp)

    @staticmethod
    def _u8(a):
        if a is None:
            return None, None
        a = np.ascontiguousarray(a, dtype=np.uint8)
        return a, a.ctypes.data_as(C.POINTER(C.c_uint8))

    def set_threads(self, n: int):
        self.lib.ora_set_threads(int(n))

    # -- primitives ----------------------------------------------------------
    def border_interpolate(self, p, n, border):
        return self.lib.ora_border_interpolate(p, n, border)

    def gaussian_kernel(self, n, sigma):
        out = np.zeros(n, self.dt)
        self.lib.ora_gaussian_kernel(n, sigma, out.ctypes.data_as(self._rp))
        return out

    def sep_filter(self, src, kx, ky, border, delta=0.0):
        src, sp = self._r(src)
        h, w = src.shape
        kx, kxp = self._r(kx)
        ky, kyp = self._r(ky)
        dst = np.zeros_like(src)
        self.lib.ora_sep_filter(sp, w, h, dst.ctypes.data_as(self._rp), kxp, len(kx), kyp, len(ky), border, delta)
        return dst

    def pyr_down(self, src, dsize=None, border=BORDER_REFLECT101):
        src, sp = self._r(src)
        h, w = src.shape
        dw, dh = dsize if dsize else ((w + 1) // 2, (h + 1) // 2)
        dst = np.zeros((dh, dw), self.dt)
        self.lib.ora_pyr_down(sp, w, h, dst.ctypes.data_as(self._rp), dw, dh, border)
        return dst

    def pyr_down_u8(self, src, dsize=None):
        src, sp = self._u8(src)
        h, w = src.shape
        dw, dh = dsize if dsize else ((w + 1) // 2, (h + 1) // 2)
        dst = np.zeros((dh, dw), np.uint8)
        self.lib.ora_pyr_down_u8(sp, w, h, dst.ctypes.data_as(C.POINTER(C.c_uint8)), dw, dh)
        return dst

    def pyr_up(self, src, dsize=None):
        src, sp = self._r(src)
        h, w = src.shape
        dw, dh = dsize if dsize else (2 * w, 2 * h)
        dst = np.zeros((dh, dw), self.dt)
        self.lib.ora_pyr_up(sp, w, h, dst.ctypes.data_as(self._rp), dw, dh)
        return dst

    def resize_nearest_u8(self, src, dsize):
        src, sp = self._u8(src)
        h, w = src.shape
        dw, dh = dsize
        dst = np.zeros((dh, dw), np.uint8)
        self.lib.ora_resize_nearest_u8(sp, w, h, dst.ctypes.data_as(C.POINTER(C.c_uint8)), dw, dh)
        return dst

    def remap(self, src, rmap, interp=INTER_LINEAR, border=BORDER_CONSTANT, cval=0.0):
        src, sp = self._r(src)
        h, w = src.shape
        rmap, mp = self._r(rmap)
        dh, dw = rmap.shape[:2]
        dst = np.zeros((dh, dw), self.dt)
        self.lib.ora_remap(sp, w, h, mp, dw, dh, dst.ctypes.data_as(self._rp), interp, border, cval)
        return dst

    def remap_u8(self, src, rmap, border=BORDER_CONSTANT, cval=0):
        src, sp = self._u8(src)
        h, w = src.shape
        rmap, mp = self._r(rmap)
        dh, dw = rmap.shape[:2]
        dst = np.zeros((dh, dw), np.uint8)
        self.lib.ora_remap_u8(sp, w, h, mp, dw, dh, dst.ctypes.data_as(C.POINTER(C.c_uint8)), border, cval)
        return dst

    def erode(self, src, k=5, border=BORDER_REPLICATE, cval=255):
        src, sp = self._u8(src)
        h, w = src.shape
        dst = np.zeros_like(src)
        self.lib.ora_erode_rect_u8(sp, w, h, k, k, dst.ctypes.data_as(C.POINTER(C.c_uint8)), border, cval)
        return dst

    def mean_stddev(self, src, mask=None):
        src, sp = self._r(src)
        mask, mp = self._u8(mask)
        m, s = C.c_double(), C.c_double()
        self.lib.ora_mean_stddev(sp, mp, src.size, C.byref(m), C.byref(s))
        return m.value, s.value

    def cholesky_invert(self, A):
        A = np.array(A, dtype=self.dt, order="C")
        ok = self.lib.ora_cholesky_invert(A.ctypes.data_as(self._rp), A.shape[0])
        return bool(ok), A

    def ecc_differentiate(self, src):
        src, sp = self._r(src)
        h, w = src.shape
        gx, gy = np.zeros_like(src), np.zeros_like(src)
        self.lib.ora_ecc_differentiate(sp, w, h, gx.ctypes.data_as(self._rp), gy.ctypes.data_as(self._rp))
        return gx, gy

    # -- transforms ------------------------------------------------------------
    def xform(self, motion, params=None, center=(0.0, 0.0)):
        t = self.XForm()
        self.lib.ora_xform_reset(C.byref(t), motion)
        if params is not None:
            for i, v in enumerate(params):
                t.p[i] = v
        t.C[0], t.C[1] = center
        return t

    def create_map(self, t, w, h):
        m = np.zeros((h, w, 2), self.dt)
        self.lib.ora_xform_create_map(C.byref(t), w, h, m.ctypes.data_as(self._rp))
        return m

    def xform_scale(self, t, f):
        self.lib.ora_xform_scale(C.byref(t), f)

    def xform_eps(self, t, dp, w, h):
        dp, dpp = self._r(np.asarray(dp))
        return self.lib.ora_xform_eps(C.byref(t), dpp, w, h)

    def xform_matrix(self, t):
        m = (C.c_double * 6)()
        self.lib.ora_xform_matrix(C.byref(t), m)
        return np.array(m[:])

    def ecc_fa_align(self, ref, cur, t, ref_mask=None, cur_mask=None, interp=INTER_LINEAR,
                     max_iterations=50, max_eps=0.2, update_step_scale=1.5):
        ref, rp_ = self._r(ref)
        cur, cp = self._r(cur)
        rm, rmp = self._u8(ref_mask)
        cm, cmp_ = self._u8(cur_mask)
        res = LevelResult()
        ok = self.lib.ora_ecc_fa_align(rp_, rmp, ref.shape[1], ref.shape[0], cp, cmp_, cur.shape[1], cur.shape[0],
                                       C.byref(t), interp, max_iterations, max_eps, update_step_scale, C.byref(res))
        return bool(ok), res

    def ecc_lm_align(self, ref, cur, t, ref_mask=None, cur_mask=None, max_iterations=50, max_eps=0.2,
                     update_step_scale=1.5):
        """c_ecclm::align on one level (ref_mask: the eroded reference mask)."""
        ref, rp_ = self._r(ref)
        cur, cp = self._r(cur)
        rm, rmp = self._u8(ref_mask)
        cm, cmp_ = self._u8(cur_mask)
        res = LevelResult()
        ok = self.lib.ora_ecc_lm_align(rp_, rmp, ref.shape[1], ref.shape[0], cp, cmp_, cur.shape[1], cur.shape[0],
                                       C.byref(t), max_iterations, max_eps, update_step_scale, C.byref(res))
        return bool(ok), res

    def ecc_ic_align(self, ref, cur, t, ref_mask=None, cur_mask=None, max_iterations=50, max_eps=0.2,
                     update_step_scale=1.5, lm=False, jx=None):
        """c_ecc_inverse_compositional::align, or c_ecclm_inverse_compositional::align
        with lm=True; J at the parameters of jx (default: t as passed)."""
        ref, rp_ = self._r(ref)
        cur, cp = self._r(cur)
        rm, rmp = self._u8(ref_mask)
        cm, cmp_ = self._u8(cur_mask)
        if jx is None:
            jx = self.XForm()
            C.pointer(jx)[0] = t
        res = LevelResult()
        fn = self.lib.ora_ecc_iclm_align if lm else self.lib.ora_ecc_ic_align
        ok = fn(rp_, rmp, ref.shape[1], ref.shape[0], cp, cmp_, cur.shape[1], cur.shape[0], C.byref(t),
                C.byref(jx), max_iterations, max_eps, update_step_scale, C.byref(res))
        return bool(ok), res

    def invert_and_compose(self, t, dp):
        dp, dpp = self._r(np.asarray(dp, dtype=self.dt))
        out = np.zeros(6, self.dt)
        out[:t.nparams] = np.asarray(t.p[:t.nparams])
        self.lib.ora_xform_invert_and_compose(C.byref(t), dpp, out.ctypes.data_as(self._rp))
        return out[:t.nparams]

    def compute_correlation(self, cur, ref, rmap, cur_mask=None, ref_mask=None):
        cur, cp = self._r(cur)
        ref, rp_ = self._r(ref)
        rmap, mp = self._r(rmap)
        cm, cmp_ = self._u8(cur_mask)
        rm, rmp = self._u8(ref_mask)
        return self.lib.ora_compute_correlation(cp, cmp_, cur.shape[1], cur.shape[0], rp_, rmp, ref.shape[1],
                                                ref.shape[0], mp)

    def base_remap(self, src, rmap, src_mask=None, interp=INTER_LINEAR, border=BORDER_REFLECT101, cval=0.0,
                   want_mask=True):
        src, sp = self._r(src)
        cn = 1 if src.ndim == 2 else src.shape[2]
        h, w = src.shape[:2]
        rmap, mp = self._r(rmap)
        dh, dw = rmap.shape[:2]
        dst = np.zeros((dh, dw) if cn == 1 else (dh, dw, cn), self.dt)
        sm, smp = self._u8(src_mask)
        dm = np.zeros((dh, dw), np.uint8) if want_mask else None
        self.lib.ora_base_remap(sp, w, h, cn, smp, mp, dw, dh, dst.ctypes.data_as(self._rp),
                                dm.ctypes.data_as(C.POINTER(C.c_uint8)) if want_mask else None, interp, border, cval)
        return dst, dm

    def weighted_average_update(self, src, acc, accw, mask=None, weights=None):
        src, sp = self._r(src)
        cn = 1 if src.ndim == 2 else src.shape[2]
        h, w = src.shape[:2]
        assert acc.dtype == self.dt and accw.dtype == self.dt and acc.flags.c_contiguous
        m, mp = self._u8(mask)
        wts, wp = (None, None) if weights is None else self._r(weights)
        self.lib.ora_weighted_average_update(sp, w, h, cn, mp, wp, acc.ctypes.data_as(self._rp),
                                             accw.ctypes.data_as(self._rp))

    def sharpness(self, src):
        src, sp = self._r(src)
        return self.lib.ora_sharpness_mean_sq_laplacian(sp, src.shape[1], src.shape[0])

    # -- bayer average -------------------------------------------------------------
    def bayer_channel(self, colorid, x, y):
        return self.lib.ora_bayer_channel(colorid, x, y)

    def bayer_state(self, w, h):
        """(acc, cntr, counts) planes of an empty c_bayer_average, (h, w, 3)."""
        return np.zeros((h, w, 3), self.dt), np.zeros((h, w, 3), self.dt), np.zeros((h, w, 3), np.int32)

    def bayer_accumulate(self, src, state, colorid, rmap=None, mask=None, weights=None):
        src, sp = self._r(src)
        h, w = src.shape
        acc, cntr, counts = state
        assert acc.dtype == self.dt and cntr.dtype == self.dt and counts.dtype == np.int32
        rm, rmp = (None, None) if rmap is None else self._r(rmap)
        m, mp = self._u8(mask)
        wts, wp = (None, None) if weights is None else self._r(weights)
        self.lib.ora_bayer_accumulate(sp, w, h, rmp, colorid, mp, wp, acc.ctypes.data_as(self._rp),
                                      cntr.ctypes.data_as(self._rp), counts.ctypes.data_as(C.POINTER(C.c_int32)))

    def bayer_compute(self, state):
        acc, cntr, _ = state
        h, w = acc.shape[:2]
        avg = np.zeros((h, w, 3), self.dt)
        mask = np.zeros((h, w), np.uint8)
        self.lib.ora_bayer_compute(acc.ctypes.data_as(self._rp), cntr.ctypes.data_as(self._rp), w, h,
                                   avg.ctypes.data_as(self._rp), mask.ctypes.data_as(C.POINTER(C.c_uint8)))
        return avg, mask

    # -- registration ------------------------------------------------------------
    def registration(self, opts: RegOptions):
        return OracleRegistration(self, opts)


class OracleRegistration:
    """c_frame_registration restated (ECC-only registration path)."""

    def __init__(self, ora: Oracle, opts: RegOptions):
        self.ora = ora
        self.opts = opts
        self.h = ora.lib.ora_reg_create(C.byref(opts))

    def __del__(self):
        if getattr(self, "h", None):
            self.ora.lib.ora_reg_destroy(self.h)
            self.h = None

    def setup_reference_frame(self, img, mask=None):
        img, ip = self.ora._r(img)
        m, mp = self.ora._u8(mask)
        self.ora.lib.ora_reg_setup_reference(self.h, ip, mp, img.shape[1], img.shape[0])

    def num_levels(self):
        return self.ora.lib.ora_reg_num_levels(self.h)

    def level_size(self, lvl):
        w, h = C.c_int(), C.c_int()
        self.ora.lib.ora_reg_level_size(self.h, lvl, C.byref(w), C.byref(h))
        return w.value, h.value

    def reference_level(self, lvl):
        w, h = self.level_size(lvl)
        img = np.zeros((h, w), self.ora.dt)
        m = np.zeros((h, w), np.uint8)
        self.ora.lib.ora_reg_get_ref_level(self.h, lvl, img.ctypes.data_as(self.ora._rp),
                                           m.ctypes.data_as(C.POINTER(C.c_uint8)))
        return img, m

    def register_frame(self, img, mask=None) -> RegStatus:
        img, ip = self.ora._r(img)
        m, mp = self.ora._u8(mask)
        st = RegStatus()
        self.ora.lib.ora_reg_register(self.h, ip, mp, img.shape[1], img.shape[0], C.byref(st))
        return st


def affine_map(matrix, w, h, dtype=np.float64):
    """Remap field x' = m0 x + m1 y + m2, y' = m3 x + m4 y + m5 (float64 math)."""
    m = np.asarray(matrix, np.float64)
    yy, xx = np.mgrid[0:h, 0:w].astype(np.float64)
    out = np.empty((h, w, 2), np.float64)
    out[..., 0] = m[0] * xx + m[1] * yy + m[2]
    out[..., 1] = m[3] * xx + m[4] * yy + m[5]
    return out.astype(dtype)

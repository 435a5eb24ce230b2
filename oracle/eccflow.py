"""Dense ECC flow (c_eccflow) -- CPU oracle, test infrastructure only.

Only tests/, the smoke entry point and bench.py's CPU baseline load this file;
the product path never does.

Restates c_eccflow (core/proc/image_registration/ecc2.cc:2103-2865) on the FP32
oracle's pyrDown, ecc_differentiate and remap, with numpy float32 arithmetic in the
order the reference source writes it (the reference builds with -ffast-math, so
its own loops have no fixed order; this is the canonical one the HIP kernels use):

  * reference pyramid (:2489-2672): level 0 = the image as float; then by
    downscale method -- RecursiveResize / FullResize: INTER_AREA resize to
    ((int)((w+1) s), (int)((h+1) s)) from the previous level (or level 0), stop when
    the size repeats or max(size) <= max(4, min_image_size); Pyramid: pyrDown to
    ((w+1)/2, (h+1)/2), stop when min(size) <= max(4, min_image_size); masks follow
    with INTER_NEAREST.  Per level Ix, Iy = ecc_differentiate; Ixx, Ixy, Iyy =
    avgdown(products); D = (a00, a01, a11, update_multiplier / (|a00 a11 - a01^2| +
    reg)), reg = (1e-5 noise_level / 2^level)^4 (noise_level < 0 means 1e-3);
  * avgdown (:2400-2418): cv::resize INTER_AREA to the size halved support_scale
    times, then sepFilter2D with [1/4 1/2 1/4] x [1/4 1/2 1/4], BORDER_REPLICATE;
  * compute_uv (:2770-2835): initial flow = rmap - identity resized INTER_CUBIC to
    the coarsest level, then scaled by the size ratio; coarse to fine: scale by the
    size ratio, resize INTER_CUBIC to the level, then max_iterations times
    (:2237-2398): W = remap(current, identity + uv, LINEAR, REPLICATE), mask =
    remap(current mask, NEAREST, CONSTANT) & reference mask, It = ref - W,
    b = avgdown(It Ix, It Iy) (zero off the mask), per cell
    uv = idet (a11 b0 - a01 b1, a00 b1 - a01 b0), resized INTER_CUBIC to the level
    and added to uv; compute() returns identity + uv (:2856-2865).

OpenCV 4.5 resize semantics restated here (imgproc/src/resize.cpp):
  * INTER_CUBIC: half-pixel centres fx = (float)((dx+0.5) scale - 0.5),
    interpolateCubic (A = -0.75) float coefficients, taps clamped to the image,
    horizontal pass then vertical pass, products summed left to right (the scalar
    loops of HResizeCubic / VResizeCubic);
  * INTER_AREA, whole factors (resizeAreaFast_): the cell read row-major, summed in
    groups of four, sum += ((a+b)+c)+d, then the tail, times 1.f/area;
  * INTER_AREA, other factors (ResizeArea_Invoker): computeResizeAreaTab weights,
    per source row buf = sum_k S[si_k] alpha_k, per cell sum += beta buf;
  * equal sizes: a copy.
"""
from __future__ import annotations

import math

import numpy as np

F32 = np.float32
DBL_EPSILON = 2.220446049250313e-16

RECURSIVE_RESIZE, FULL_RESIZE, PYRAMID = 0, 1, 2  # ECCFlowDownscaleMethod (ecc2.h:507-511)


# --------------------------------------------------------------------------------------------
# cv::resize INTER_CUBIC
# --------------------------------------------------------------------------------------------
def _dt(a):
    """Working type of an image: float64 for the FP64 oracle, float32 otherwise."""
    return np.float64 if np.asarray(a).dtype == np.float64 else F32


def cubic_coeffs(x, dt=F32):
    """interpolateCubic (imgproc resize.cpp) in float (or double), vectorised over x."""
    x = np.asarray(x, dt)
    A = dt(-0.75)
    one = dt(1)
    x1 = x + one
    c0 = ((A * x1 - dt(5) * A) * x1 + dt(8) * A) * x1 - dt(4) * A
    c1 = ((A + dt(2)) * x - (A + dt(3))) * x * x + one
    xr = one - x
    c2 = ((A + dt(2)) * xr - (A + dt(3))) * xr * xr + one
    c3 = one - c0 - c1 - c2
    return c0, c1, c2, c3


def cubic_axis(ssize, dsize, dt=F32):
    """Source taps (clamped) and coefficients of one axis."""
    scale = 1.0 / (dsize / ssize)
    f = np.array([(d + 0.5) * scale - 0.5 for d in range(dsize)], np.float64).astype(dt)
    s = np.floor(f).astype(np.int64)
    f = (f - s.astype(dt)).astype(dt)
    taps = np.clip(s[:, None] + np.arange(-1, 3)[None, :], 0, ssize - 1)
    return taps, np.stack(cubic_coeffs(f, dt), 1)


def _cubic_1ch(img, dw, dh):
    h, w = img.shape
    dt = img.dtype.type
    xt, a = cubic_axis(w, dw, dt)
    yt, b = cubic_axis(h, dh, dt)
    rows = ((img[:, xt[:, 0]] * a[:, 0] + img[:, xt[:, 1]] * a[:, 1]) + img[:, xt[:, 2]] * a[:, 2]) \
        + img[:, xt[:, 3]] * a[:, 3]
    out = ((rows[yt[:, 0]] * b[:, 0:1] + rows[yt[:, 1]] * b[:, 1:2]) + rows[yt[:, 2]] * b[:, 2:3]) \
        + rows[yt[:, 3]] * b[:, 3:4]
    return out.astype(dt)


def resize_cubic(img, dw, dh):
    """cv::resize(img, (dw, dh), INTER_CUBIC) of a float image with 1 or more channels."""
    img = np.asarray(img, _dt(img))
    if img.shape[1] == dw and img.shape[0] == dh:
        return img.copy()
    if img.ndim == 2:
        return _cubic_1ch(img, dw, dh)
    return np.stack([_cubic_1ch(img[..., c], dw, dh) for c in range(img.shape[2])], -1)


# --------------------------------------------------------------------------------------------
# cv::resize INTER_AREA (downscaling)
# --------------------------------------------------------------------------------------------
def area_is_fast(ssize, dsize):
    """is_area_fast for one axis: the scale is an integer to within DBL_EPSILON."""
    scale = 1.0 / (dsize / ssize)
    iscale = int(np.rint(scale))
    return abs(scale - iscale) < DBL_EPSILON, iscale


def area_tab(ssize, dsize):
    """computeResizeAreaTab: list of (di, si, alpha) in the reference order (alpha in
    double; the float path rounds it to float as the reference table does)."""
    scale = 1.0 / (dsize / ssize)
    tab = []
    for dx in range(dsize):
        fsx1 = dx * scale
        fsx2 = fsx1 + scale
        cell = min(scale, ssize - fsx1)
        sx1, sx2 = math.ceil(fsx1), math.floor(fsx2)
        sx2 = min(sx2, ssize - 1)
        sx1 = min(sx1, sx2)
        if sx1 - fsx1 > 1e-3:
            tab.append((dx, sx1 - 1, (sx1 - fsx1) / cell))
        for sx in range(sx1, sx2):
            tab.append((dx, sx, 1.0 / cell))
        if fsx2 - sx2 > 1e-3:
            tab.append((dx, sx2, min(min(fsx2 - sx2, 1.0), cell) / cell))
    return tab


def _padded(tab, dsize):
    """(dsize, m) source indices and weights, padded at the end with (0, 0)."""
    per = [[] for _ in range(dsize)]
    for di, si, al in tab:
        per[di].append((si, al))
    m = max(len(p) for p in per)
    idx = np.zeros((dsize, m), np.int64)
    alp = np.zeros((dsize, m), np.float64)
    for d, p in enumerate(per):
        for j, (si, al) in enumerate(p):
            idx[d, j] = si
            alp[d, j] = al
    return idx, alp


def _area_1ch(img, dw, dh):
    h, w = img.shape
    dt = img.dtype.type
    fx, isx = area_is_fast(w, dw)
    fy, isy = area_is_fast(h, dh)
    if fx and fy:
        area = isx * isy
        cells = img[:dh * isy, :dw * isx].reshape(dh, isy, dw, isx).transpose(0, 2, 1, 3).reshape(dh, dw, area)
        s = np.zeros((dh, dw), dt)
        k = 0
        while k <= area - 4:
            s = s + (((cells[..., k] + cells[..., k + 1]) + cells[..., k + 2]) + cells[..., k + 3])
            k += 4
        while k < area:
            s = s + cells[..., k]
            k += 1
        return (s * (dt(1) / dt(area))).astype(dt)
    xi, xa = _padded(area_tab(w, dw), dw)
    yi, ya = _padded(area_tab(h, dh), dh)
    xa, ya = xa.astype(dt), ya.astype(dt)
    buf = np.zeros((h, dw), dt)
    for j in range(xi.shape[1]):
        buf = buf + img[:, xi[:, j]] * xa[:, j]
    s = buf[yi[:, 0]] * ya[:, 0:1]
    for j in range(1, yi.shape[1]):
        s = s + buf[yi[:, j]] * ya[:, j:j + 1]
    return s.astype(dt)


def resize_area(img, dw, dh):
    """cv::resize(img, (dw, dh), INTER_AREA) for a shrinking (or equal) size."""
    img = np.asarray(img, _dt(img))
    h, w = img.shape[:2]
    if w == dw and h == dh:
        return img.copy()
    assert dw <= w and dh <= h, "INTER_AREA upscaling is not restated"
    if img.ndim == 2:
        return _area_1ch(img, dw, dh)
    return np.stack([_area_1ch(img[..., c], dw, dh) for c in range(img.shape[2])], -1)


# --------------------------------------------------------------------------------------------
# masks
# --------------------------------------------------------------------------------------------
def resize_nearest_u8(m, dw, dh):
    """cv::resize INTER_NEAREST: sx = min(floor(dx * sw/dw), sw-1)."""
    h, w = m.shape
    sx = np.minimum(np.floor(np.arange(dw) * (w / dw)).astype(np.int64), w - 1)
    sy = np.minimum(np.floor(np.arange(dh) * (h / dh)).astype(np.int64), h - 1)
    return m[sy[:, None], sx[None, :]].copy()


def remap_nearest_u8(m, rmap):
    """cv::remap(m, rmap(float), INTER_NEAREST, BORDER_CONSTANT 0): the map is rounded
    to nearest (ties to even) and saturated to short."""
    h, w = m.shape
    X = np.clip(np.rint(rmap[..., 0].astype(np.float64)), -32768, 32767).astype(np.int64)
    Y = np.clip(np.rint(rmap[..., 1].astype(np.float64)), -32768, 32767).astype(np.int64)
    inside = (X >= 0) & (X < w) & (Y >= 0) & (Y < h)
    out = np.zeros(rmap.shape[:2], np.uint8)
    out[inside] = m[Y[inside], X[inside]]
    return out


# --------------------------------------------------------------------------------------------
# c_eccflow
# --------------------------------------------------------------------------------------------
def scale_flow(uv, rx, ry):
    """cv::multiply(uv, Scalar(rx, ry), uv): with a Scalar operand the product is taken in
    double and rounded to float."""
    return (uv.astype(np.float64) * np.array([rx, ry], np.float64)).astype(uv.dtype)


class EccFlow:
    """c_eccflow with c_eccflow_options (ecc2.h:515-527); defaults are those options'.
    With the FP64 oracle every plane and flow is float64 (the warp-accuracy reference)."""

    def __init__(self, ora, update_multiplier=1.5, noise_level=-1.0, max_iterations=1, support_scale=5,
                 min_image_size=4, max_pyramid_level=-1, downscale=RECURSIVE_RESIZE, scale_factor=0.5):
        self.o = ora
        self.dt = np.float64 if ora.bits == 64 else F32
        self.um, self.noise = update_multiplier, noise_level
        self.iters, self.ss = max_iterations, support_scale
        self.min_size, self.max_level = min_image_size, max_pyramid_level
        self.method, self.sf = downscale, scale_factor
        self.G3 = np.array([0.25, 0.5, 0.25], self.dt)

    # avgdown (ecc2.cc:2400-2418)
    def cell_size(self, w, h):
        for _ in range(self.ss):
            w, h = (w + 1) // 2, (h + 1) // 2
        return w, h

    def avgdown(self, x):
        h, w = x.shape[:2]
        cw, ch = self.cell_size(w, h)
        a = resize_area(x, cw, ch)
        if a.ndim == 2:
            return self.o.sep_filter(a, self.G3, self.G3, 1)
        return np.stack([self.o.sep_filter(a[..., c], self.G3, self.G3, 1) for c in range(a.shape[2])], -1)

    def _downscale(self, img, mask, size):
        w, h = size
        if self.method == PYRAMID:
            d = self.o.pyr_down(img, (w, h))
        else:
            d = resize_area(img, w, h)
        return d, (None if mask is None else resize_nearest_u8(mask, w, h))

    def level_sizes(self, w0, h0):
        """The reference pyramid's level sizes (:2530-2595) and the source level of each."""
        min_size = max(4, self.min_size)
        big_aspect = max(w0, h0) // min(w0, h0) >= 2
        sizes, srcs = [(w0, h0)], [None]
        level = 0
        while True:
            if self.max_level >= 0 and level >= self.max_level:
                break
            pw, ph = sizes[-1]
            if self.method == PYRAMID:
                nw, nh = max(self.min_size, (pw + 1) // 2), max(self.min_size, (ph + 1) // 2)
                if (nw, nh) == (pw, ph) or min(nw, nh) <= min_size:
                    break
                src = level
            else:
                nw = max(self.min_size, int((pw + 1) * self.sf))
                nh = max(self.min_size, int((ph + 1) * self.sf))
                if (nw, nh) == (pw, ph) or max(nw, nh) <= min_size:
                    break
                if self.method == FULL_RESIZE or (big_aspect and min(nw, nh) <= min_size + 1):
                    src = 0
                else:
                    src = level
            sizes.append((nw, nh))
            srcs.append(src)
            level += 1
        return sizes, srcs

    def _pyramid(self, img, mask):
        h0, w0 = img.shape
        imgs = [np.asarray(img, self.dt).copy()]
        masks = [None if mask is None else np.asarray(mask, np.uint8).copy()]
        for size, src in zip(self.sizes[1:], self.srcs[1:]):
            d, m = self._downscale(imgs[src], masks[src], size)
            imgs.append(d)
            masks.append(m)
        return imgs, masks

    def set_reference_image(self, ref, mask=None):
        h0, w0 = np.asarray(ref).shape
        self.sizes, self.srcs = self.level_sizes(w0, h0)
        noise = self.noise if self.noise >= 0 else 1e-3
        imgs, masks = self._pyramid(ref, mask)
        self.levels = []
        for level, (r, m) in enumerate(zip(imgs, masks)):
            Ix, Iy = self.o.ecc_differentiate(r)
            a00, a01, a11 = self.avgdown(Ix * Ix), self.avgdown(Ix * Iy), self.avgdown(Iy * Iy)
            dt = self.dt
            reg = dt((1e-5 * noise / (1 << level)) ** 4) if noise > 0 else dt(0)
            det = np.abs(a00 * a11 - a01 * a01)
            idet = (dt(self.um) / (det + reg)).astype(dt)
            self.levels.append(dict(ref=r, mask=m, Ix=Ix, Iy=Iy, a00=a00, a01=a01, a11=a11, idet=idet))

    def iterate(self, e, cur, cmask, uv):
        """One compute_uv(pyramid_entry) step (:2237-2398): returns the cell update
        resized to the level."""
        h, w = e["ref"].shape
        yy, xx = np.mgrid[0:h, 0:w].astype(self.dt)
        rmap = np.stack([uv[..., 0] + xx, uv[..., 1] + yy], -1)
        W = self.o.remap(cur, rmap, interp=1, border=1)
        M = None if cmask is None else remap_nearest_u8(cmask, rmap)
        if e["mask"] is not None:
            M = e["mask"] if M is None else (e["mask"] & M)
        It = e["ref"] - W
        itxy = np.stack([It * e["Ix"], It * e["Iy"]], -1)
        if M is not None:
            itxy[M == 0] = 0
        b = self.avgdown(itxy)
        b0, b1 = b[..., 0], b[..., 1]
        cu = e["idet"] * (e["a11"] * b0 - e["a01"] * b1)
        cv = e["idet"] * (e["a00"] * b1 - e["a01"] * b0)
        return resize_cubic(np.stack([cu, cv], -1), w, h)

    def compute_uv(self, frame, rmap=None, mask=None):
        L = len(self.levels)
        cur, cmasks = self._pyramid(frame, mask)
        w0, h0 = self.sizes[0]
        wl, hl = self.sizes[-1]
        dt = self.dt
        if rmap is None:
            uv = np.zeros((hl, wl, 2), dt)
        else:
            yy, xx = np.mgrid[0:h0, 0:w0].astype(dt)
            rmap = np.asarray(rmap, dt)
            flow = np.stack([rmap[..., 0] - xx, rmap[..., 1] - yy], -1).astype(dt)
            uv = scale_flow(resize_cubic(flow, wl, hl), wl / w0, hl / h0)
        for i in range(L - 1, -1, -1):
            w, h = self.sizes[i]
            if i < L - 1:
                pw, ph = self.sizes[i + 1]
                uv = resize_cubic(scale_flow(uv, w / pw, h / ph), w, h)
            for _ in range(self.iters):
                uv = uv + self.iterate(self.levels[i], cur[i], cmasks[i], uv)
        return uv

    def compute(self, frame, rmap=None, mask=None):
        """Returns the refined remap (identity + uv) at the reference size."""
        uv = self.compute_uv(frame, rmap, mask)
        h0, w0 = uv.shape[:2]
        yy, xx = np.mgrid[0:h0, 0:w0].astype(self.dt)
        return np.stack([uv[..., 0] + xx, uv[..., 1] + yy], -1)

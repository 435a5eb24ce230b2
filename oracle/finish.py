"""CPU oracle of the stack finish (test infrastructure only): linear-interpolation
inpainting (core/proc/inpaint/linear_interpolation_inpaint.cc), unsharp_mask
(core/proc/unsharp_mask.cc:15-115), average_pyramid_inpaint, and the frame and remap
upscales of stacking at a larger size (c_image_stacking_pipeline.cc:1869-2000) with
OpenCV's resize / pyrUp on float and 8-bit data restated from its generic C++ paths.

Only tests/, ``__graft_entry__.smoke()`` and bench.py's CPU leg use this module.
The inpainting follows the reference's run loops literally (pure Python over
rows and columns, float32 scalars), so it is meant for small images.
"""
from __future__ import annotations

import math

import numpy as np

f32 = np.float32


def _runs(valid):
    """(start, end) of the runs of invalid entries, as the reference's while loops find them."""
    n = len(valid)
    start = 0
    while start < n:
        while start < n and valid[start]:
            start += 1
        if start >= n:
            break
        end = start + 1
        while end < n and not valid[end]:
            end += 1
        yield start, end
        start = end


def _fill_line(vals, valid, out, dist):
    """One row / column of _interpolate_holes_h2 / _v2 (linear_interpolation_inpaint.cc:14-227)."""
    n = len(valid)
    for start, end in _runs(valid):
        if start > 0 and end < n:
            s, e = start - 1, end
            scale = f32(1.0) / f32(end - start)
            kk = (vals[e] - vals[s]) * scale
            for x in range(start, e):
                out[x] = vals[s] + f32(x - s) * kk
                dist[x] = max(x - s, e - x)
        elif start > 0:
            s = start - 1
            for x in range(start, n):
                out[x] = vals[s]
                dist[x] = x - s
        elif end < n:
            for x in range(start, end):
                out[x] = vals[end]
                dist[x] = end - x


def inpaint_linear(image: np.ndarray, mask: np.ndarray) -> np.ndarray:
    """linear_interpolation_inpaint(image, mask) (linear_interpolation_inpaint.cc:335-361)."""
    img = np.array(image, dtype=np.float32, copy=True)
    if img.ndim == 2:
        img = img[..., None]
    h, w, cn = img.shape
    m = np.array(mask, dtype=np.uint8, copy=True)
    total = int((m == 0).sum())
    while total > 0:
        ih = np.zeros_like(img)
        iv = np.zeros_like(img)
        hd = np.zeros((h, w), np.float32)
        vd = np.zeros((h, w), np.float32)
        valid = m != 0
        for y in range(h):
            _fill_line(img[y], valid[y], ih[y], hd[y])
        for x in range(w):
            col_out = iv[:, x]
            col_d = vd[:, x]
            _fill_line(img[:, x], valid[:, x], col_out, col_d)
        filled = 0
        for y in range(h):
            for x in range(w):
                if m[y, x]:
                    continue
                dh, dv = hd[y, x], vd[y, x]
                if dh > 0 and dv > 0:
                    dd = f32(1.0) / (dh + dv)
                    img[y, x] = dd * (ih[y, x] * dv + iv[y, x] * dh)
                elif dv > 0:
                    img[y, x] = iv[y, x]
                elif dh > 0:
                    img[y, x] = ih[y, x]
                else:
                    continue
                m[y, x] = 255
                filled += 1
        if filled < 1:
            break
        total -= filled
    return img[..., 0] if image.ndim == 2 else img


def lpass_plan(sigma: float, w: int, h: int):
    """create_lpass_image's pyramid choice (unsharp_mask.cc:15-48): (levels, Ci, delta)."""
    level, Ci = 0, 0
    if sigma > 2:
        min_size, imax = min(w, h), 0
        while True:
            min_size >>= 1
            if not min_size:
                break
            imax += 1
        C = int(sigma * sigma / 2)
        while level < imax and (1 + 4 * Ci) <= C:
            Ci = 1 + 4 * Ci
            level += 1
    delta = math.sqrt(sigma * sigma - 2 * Ci) / (1 << level) if level else sigma
    return level, Ci, delta


def _gaussian(ora, plane, sigma):
    n = 2 * max(1, int(sigma * 5)) + 1
    k = ora.gaussian_kernel(n, sigma)
    return ora.sep_filter(plane, k, k, 2)  # BORDER_REFLECT


def lpass(ora, plane: np.ndarray, sigma: float) -> np.ndarray:
    """create_lpass_image (unsharp_mask.cc:15-70) of one f32 plane, over the C oracle's
    sepFilter2D / pyrDown / pyrUp."""
    h, w = plane.shape
    level, Ci, delta = lpass_plan(sigma, w, h)
    if level < 1:
        return _gaussian(ora, plane, sigma)
    sizes = [(w, h)]
    lp = ora.pyr_down(plane, border=2)
    for _ in range(1, level):
        sizes.append((lp.shape[1], lp.shape[0]))
        lp = ora.pyr_down(lp, border=2)
    if delta > 0:
        lp = _gaussian(ora, lp, delta)
    for j in range(len(sizes) - 1, -1, -1):
        lp = ora.pyr_up(lp, sizes[j])
    return lp


def unsharp_mask(ora, image: np.ndarray, sigma: float, alpha: float) -> np.ndarray:
    """unsharp_mask(src, dst, sigma, alpha) (unsharp_mask.cc:72-115) for f32 images."""
    if sigma <= 0 or alpha <= 0:
        return image.copy()
    planes = [image] if image.ndim == 2 else [np.ascontiguousarray(image[..., c]) for c in range(image.shape[2])]
    a, b = f32(1.0 / (1.0 - alpha)), f32(-alpha / (1.0 - alpha))
    out = [(p * a + lpass(ora, p, sigma).astype(np.float32) * b).astype(np.float32) for p in planes]
    return out[0] if image.ndim == 2 else np.stack(out, -1)


# ---- average_pyramid_inpaint (core/proc/inpaint/average_pyramid_inpaint.cc) ------------------------

def _box3(a):
    """cv::boxFilter(ksize 3x3, normalize=false, BORDER_REPLICATE) of an (h, w[, cn]) float32
    array: the 9-term sum in float64 (OpenCV's sum type for float data), rounded to float32."""
    h, w = a.shape[:2]
    pad = ((1, 1), (1, 1)) + ((0, 0),) * (a.ndim - 2)
    p = np.pad(a.astype(np.float64), pad, mode="edge")
    s = np.zeros(a.shape, np.float64)
    for dy in range(3):
        for dx in range(3):
            s += p[dy:dy + h, dx:dx + w]
    return s.astype(np.float32)


def _filter2(src, srcmask, fb_src, fb_mask):
    """_average_pyramid_filter2 (average_pyramid_inpaint.cc:17-58): src (h, w, cn), masks (h, w)."""
    dst = _box3(src)
    dmask = _box3(srcmask)
    fb = fb_mask != 0
    div = ~fb & (dmask != 0)
    scale = np.where(div, f32(1.0) / np.where(div, dmask, f32(1.0)), f32(1.0)).astype(np.float32)
    dst = np.where(fb[..., None], fb_src, np.where(div[..., None], (dst * scale[..., None]).astype(np.float32), dst))
    dmask = np.where(fb | div, f32(1.0), dmask).astype(np.float32)
    return dst.astype(np.float32), dmask


def _downstrike_even(a):
    """downstrike_even (downstrike.cc:34-75): rows and columns min(2i + 1, n - 1)."""
    h, w = a.shape[:2]
    ys = np.minimum(2 * np.arange((h + 1) // 2) + 1, h - 1)
    xs = np.minimum(2 * np.arange((w + 1) // 2) + 1, w - 1)
    return a[ys][:, xs]


def _upject_even(a, h, w):
    """upject_even (downstrike.cc:254-367) to (h, w): values and a float mask of ones at the
    odd-odd positions, zeros elsewhere."""
    out = np.zeros((h, w) + a.shape[2:], a.dtype)
    z = np.zeros((h, w), np.float32)
    ny, nx = len(range(1, h, 2)), len(range(1, w, 2))
    out[1::2, 1::2] = a[:ny, :nx]
    z[1::2, 1::2] = 1
    return out, z


def _recurse(image, mask, max_levels):
    """average_pyramid_recurse (average_pyramid_inpaint.cc:69-94)."""
    h, w = image.shape[:2]
    if min(w, h) > 1 and max_levels > 0:
        fi, fm = _filter2(image, mask, image, mask)
        fi, fm = _downstrike_even(fi), _downstrike_even(fm)
        if np.count_nonzero(fm) < fm.size:
            fi, fm = _recurse(fi, fm, max_levels - 1)
        fi, fm = _upject_even(fi, h, w)
        fi, fm = _filter2(fi, fm, image, mask)
        return fi, fm
    return image, mask


def average_pyramid_inpaint(src, mask, max_levels=100):
    """average_pyramid_inpaint (average_pyramid_inpaint.cc:96-126) of an f32 image (h, w[, cn])
    and u8 mask; returns (image, mask)."""
    src = np.asarray(src, np.float32)
    if mask is None or np.count_nonzero(mask) == mask.size:
        return src.copy(), (None if mask is None else mask.copy())
    img = src if src.ndim == 3 else src[..., None]
    img = np.where((mask != 0)[..., None], img, f32(0)).astype(np.float32)
    msk = (mask.astype(np.float32) * f32(1.0 / 255.0)).astype(np.float32)
    img, msk = _recurse(img, msk, max_levels)
    out_mask = np.clip(np.rint(msk * f32(255.0)), 0, 255).astype(np.uint8)
    return (img if src.ndim == 3 else img[..., 0]), out_mask


# ---- upscale of registration remaps (c_image_stacking_pipeline.cc:1869-1905) -----------------------

UPSCALE_NONE, UPSCALE_X2, UPSCALE_X15, UPSCALE_X3 = 0, 1, 2, 3


def upscale_size(option, w, h):
    return {UPSCALE_X2: (2 * w, 2 * h), UPSCALE_X15: (w * 3 // 2, h * 3 // 2), UPSCALE_X3: (3 * w, 3 * h)}.get(
        option, (w, h))


def _fix11(w):
    """resize.cpp's 8-bit weights: saturate_cast<short>(w * INTER_RESIZE_COEF_SCALE), 11 bits,
    rounded half to even."""
    return np.rint(w.astype(np.float32) * np.float32(2048)).astype(np.int64)


def resize_linear_u8(src, dw, dh):
    """cv::resize INTER_LINEAR of an 8-bit plane (resize.cpp's generic fixed-point path): 11-bit
    weights, the horizontal pass in int (S[x0] * a0 + S[x1] * a1), the vertical pass of the
    8-bit VResizeLinear: ((b0 * (R0 >> 4)) >> 16) + ((b1 * (R1 >> 4)) >> 16) + 2) >> 2."""
    from oracle import lvm

    h, w = src.shape
    x0, x1, a0, a1 = lvm.resize_linear_coeffs(w, dw)
    y0, y1, b0, b1 = lvm.resize_linear_vcoeffs(h, dh)
    s = src.astype(np.int64)
    rows = s[:, x0] * _fix11(a0) + s[:, x1] * _fix11(a1)
    v = (((_fix11(b0)[:, None] * (rows[y0] >> 4)) >> 16) + ((_fix11(b1)[:, None] * (rows[y1] >> 4)) >> 16) + 2) >> 2
    return v.astype(np.uint8)


def _exact_coeffs(ssize, dsize):
    """interpolationLinear::getCoeffs (resize.cpp, bit-exact resize): fval = scale * (d + 0.5)
    - 0.5 in double; inside the source the weights are 8-bit fixed point, c1 = round-half-even
    (256 (fval - floor(fval))), c0 = 256 - c1; before the source or from its last sample on, the
    edge sample with (256, 0)."""
    scale = 1.0 / (dsize / ssize)
    x0 = np.zeros(dsize, np.int64)
    c1 = np.zeros(dsize, np.int64)
    for d in range(dsize):
        f = scale * (d + 0.5) - 0.5
        i = math.floor(f)
        if i >= 0 and ssize > 1:
            if i < ssize - 1:
                x0[d], c1[d] = i, int(np.rint((f - i) * 256.0))
            else:
                x0[d] = ssize - 1
    return x0, np.minimum(x0 + 1, ssize - 1), 256 - c1, c1


def resize_linear_exact_u8(src, dw, dh):
    """cv::resize INTER_LINEAR_EXACT of an 8-bit plane: the horizontal pass exact in 8-bit fixed
    point, the vertical one exact in 16-bit, rounded half up: (sum + 2^15) >> 16."""
    h, w = src.shape
    x0, x1, c0, c1 = _exact_coeffs(w, dw)
    y0, y1, d0, d1 = _exact_coeffs(h, dh)
    s = src.astype(np.int64)
    rows = s[:, x0] * c0 + s[:, x1] * c1
    v = rows[y0] * d0[:, None] + rows[y1] * d1[:, None]
    return ((v + 32768) >> 16).astype(np.uint8)


def pyr_up_u8(src):
    """cv::pyrUp of an 8-bit plane to (2w, 2h): [1 6 1] / [4 4] per axis summed in int, left/top
    REFLECT101, right/bottom edge replicated, (sum + 32) >> 6."""
    h, w = src.shape
    s = src.astype(np.int64)
    if w == 1:
        rows = np.repeat(s * 8, 2, axis=1)
    else:
        left = np.concatenate([s[:, 1:2], s[:, :-1]], axis=1)
        right = np.concatenate([s[:, 1:], s[:, -1:]], axis=1)
        rows = np.empty((h, 2 * w), np.int64)
        rows[:, 0::2] = left + 6 * s + right
        rows[:, 1::2] = 4 * (s + right)
    up = rows[1 if h > 1 else 0][None, :]
    down = rows[-1][None, :]
    prev = np.concatenate([up, rows[:-1]], axis=0)
    nxt = np.concatenate([rows[1:], down], axis=0)
    out = np.empty((2 * h, 2 * w), np.int64)
    out[0::2] = prev + 6 * rows + nxt
    out[1::2] = 4 * (rows + nxt)
    return ((out + 32) >> 6).astype(np.uint8)


def upscale_image(ora, src, mask, option):
    """upscale_image (c_image_stacking_pipeline.cc:1949-2000) of an f32 image (h, w) or (h, w, cn)
    and its optional 8-bit mask: x1.5 cv::resize INTER_LINEAR, x2 cv::pyrUp, x3 INTER_LINEAR_EXACT
    (on float data the same as INTER_LINEAR); the upscaled mask is then compare(>= 255) (0 / 255).
    The none option copies both."""
    from oracle import lvm

    src = np.asarray(src, np.float32)
    if option == UPSCALE_NONE:
        return src.copy(), (None if mask is None else mask.copy())
    h, w = src.shape[:2]
    dw, dh = upscale_size(option, w, h)
    planes = [src] if src.ndim == 2 else [np.ascontiguousarray(src[..., c]) for c in range(src.shape[2])]
    if option == UPSCALE_X2:
        up = [ora.pyr_up(p) for p in planes]
    else:
        up = [lvm.resize_linear(p, dw, dh) for p in planes]
    dst = up[0] if src.ndim == 2 else np.stack(up, -1)
    dmask = None
    if mask is not None:
        m = {UPSCALE_X15: lambda a: resize_linear_u8(a, dw, dh), UPSCALE_X2: pyr_up_u8,
             UPSCALE_X3: lambda a: resize_linear_exact_u8(a, dw, dh)}[option](np.asarray(mask, np.uint8))
        dmask = np.where(m >= 255, 255, 0).astype(np.uint8)
    return dst.astype(np.float32), dmask


def upscale_map(ora, m, option):
    """upscale_remap of an (h, w, 2) f32 map: cv::resize INTER_LINEAR per channel for x1.5 and x3
    (INTER_LINEAR_EXACT on float data is INTER_LINEAR), cv::pyrUp per channel for x2."""
    from oracle import lvm

    h, w = m.shape[:2]
    if option == UPSCALE_NONE:
        return m.copy()
    dw, dh = upscale_size(option, w, h)
    if option == UPSCALE_X2:
        chans = [ora.pyr_up(np.ascontiguousarray(m[..., c])) for c in range(2)]
    else:
        chans = [lvm.resize_linear(np.ascontiguousarray(m[..., c]), dw, dh) for c in range(2)]
    return np.stack(chans, -1).astype(np.float32)

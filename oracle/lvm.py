"""Local-variance sharpness map -- CPU oracle (test infrastructure only).

Restates compute_local_variance_map (core/proc/sharpness_measure/
c_local_variance_sharpness_measure.cc:193-249) for float images, on the FP32
oracle's pyrDown and numpy float32 elementwise arithmetic:

  1. gray input (cv::COLOR_BGR2GRAY for 3 channels: b*0.114 + g*0.587 + r*0.299),
  2. pdownscale(dscale): pyrDown while the image keeps a side >= 4 (:30-52),
  3. G = morphological gradient (dilate - erode) with a (2r+1)^2 square, BORDER_REPLICATE,
  4. W = sum |G| (NORM_L1), Q = depthScale^3 / W * sum G^4 -- summed in float64 here
     (the reference sums in float with an atomic compare-exchange per TBB range, so its
     low bits depend on the scheduling),
  5. map = G^3 * float(depthScale^3) + float(0.05 Q), depthScale = 20 for float images,
  6. cv::resize INTER_LINEAR back to the input size (half-pixel centres; columns outside
     the source take the edge column with weights (1, 0), rows outside it are clipped to
     the edge row with both weights kept, as resize.cpp's generic path does).

8-bit and 16-bit single-channel images stay integer through the pyramid (cv::pyrDown
on integer depths sums the 5x5 taps in int and rounds with (sum + 128) >> 8) and use
depthScale = 20 / maxval(depth) (getMaxValForPixelDepth: 255, 65535).
"""
from __future__ import annotations

import numpy as np

F32 = np.float32


def gray(img):
    if img.ndim == 2:
        return img.astype(F32)
    b, g, r = (img[..., c].astype(F32) for c in range(3))
    return (b * F32(0.114) + g * F32(0.587)) + r * F32(0.299)


def pyr_down_int(m):
    """cv::pyrDown of an 8/16-bit image: [1 4 6 4 1] taps per axis summed in int,
    (sum + 128) >> 8, BORDER_REFLECT_101, dst size ((w + 1) / 2, (h + 1) / 2)."""
    h, w = m.shape
    dh, dw = (h + 1) // 2, (w + 1) // 2
    p = np.pad(m.astype(np.int64), 2, mode="reflect")  # numpy's "reflect" is REFLECT_101
    taps = (1, 4, 6, 4, 1)
    cols = 2 * np.arange(dw)
    rows = sum(t * p[:, cols + k] for k, t in enumerate(taps))
    ys = 2 * np.arange(dh)
    tot = sum(t * rows[ys + k] for k, t in enumerate(taps))
    return np.clip((tot + 128) >> 8, 0, np.iinfo(m.dtype).max).astype(m.dtype)


def pdownscale(ora, m, level):
    down = pyr_down_int if m.dtype.kind == "u" else ora.pyr_down
    if min(m.shape) < 4:
        return m
    m = down(m)
    if min(m.shape) >= 4:
        for _ in range(1, level):
            m = down(m)
            if min(m.shape) < 4:
                break
    return m


def morph_gradient(m, r):
    """dilate - erode with a (2r+1)^2 square, BORDER_REPLICATE; in the image's depth, then
    as float32 (exact for 8/16-bit)."""
    k = 2 * max(1, r) + 1
    p = np.pad(m, k // 2, mode="edge")
    h, w = m.shape
    win = np.lib.stride_tricks.sliding_window_view(p, (k, k))[:h, :w]
    return (win.max(axis=(2, 3)) - win.min(axis=(2, 3))).astype(F32)


def resize_linear_coeffs(ssize, dsize):
    """cv::resize INTER_LINEAR source columns and float weights (imgproc resize.cpp, generic
    path): fx = (float)((dx + 0.5) * scale - 0.5), sx = floor(fx), fx -= sx; a column left of
    the source (sx < 0) or at/after its last one takes that edge column with fx = 0."""
    scale = 1.0 / (dsize / ssize)  # resize.cpp: scale_x = 1 / inv_scale_x
    sx = np.zeros(dsize, np.int64)
    fx = np.zeros(dsize, F32)
    for d in range(dsize):
        f = np.float32((d + 0.5) * scale - 0.5)
        s = int(np.floor(f))
        f = np.float32(f - np.float32(s))
        if s < 0:
            f, s = F32(0), 0
        if s >= ssize - 1:
            f, s = F32(0), ssize - 1
        sx[d], fx[d] = s, f
    return sx, np.minimum(sx + 1, ssize - 1), (F32(1) - fx), fx


def resize_linear_vcoeffs(ssize, dsize):
    """The rows of cv::resize INTER_LINEAR: fy and sy as for the columns, but fy is never
    reset; the two rows sy, sy + 1 are clipped to [0, ssize - 1] and keep the weights
    (1 - fy, fy) (resizeGeneric_Invoker's clip(sy, 0, ssize.height))."""
    scale = 1.0 / (dsize / ssize)
    sy = np.zeros(dsize, np.int64)
    fy = np.zeros(dsize, F32)
    for d in range(dsize):
        f = np.float32((d + 0.5) * scale - 0.5)
        s = int(np.floor(f))
        sy[d], fy[d] = s, np.float32(f - np.float32(s))
    return np.clip(sy, 0, ssize - 1), np.clip(sy + 1, 0, ssize - 1), (F32(1) - fy), fy


def resize_linear(m, dw, dh):
    """cv::resize INTER_LINEAR of a float32 plane: the horizontal pass S[x0] * a0 + S[x1] * a1,
    then the vertical R[y0] * b0 + R[y1] * b1, each in float32."""
    h, w = m.shape
    x0, x1, a0, a1 = resize_linear_coeffs(w, dw)
    y0, y1, b0, b1 = resize_linear_vcoeffs(h, dh)
    rows = m[:, x0] * a0 + m[:, x1] * a1  # horizontal pass, float32
    return (rows[y0] * b0[:, None] + rows[y1] * b1[:, None]).astype(F32)


def local_variance_map(ora, img, kradius=1, dscale=1):
    """Returns (map at the input size, Q); (zeros, 0) when the gradient vanishes."""
    h, w = img.shape[:2]
    integral = img.dtype in (np.uint8, np.uint16)
    M = img if integral else gray(img)
    if dscale > 0:
        M = pdownscale(ora, M, dscale)
    G = morph_gradient(M, kradius)
    W = float(np.abs(G.astype(np.float64)).sum())
    if not W > 0:
        return np.zeros((h, w), F32), 0.0
    ds = 20.0 / float(np.iinfo(img.dtype).max) if integral else 20.0
    g64 = G.astype(np.float64)
    Q = ds ** 3 / W * float((g64 ** 4).sum())
    m = ((G * G) * G) * F32(ds ** 3)
    m = m + F32(0.05 * Q)
    if m.shape != (h, w):
        m = resize_linear(m, w, h)
    return m, Q

"""CPU oracle of the input stage (test infrastructure only): debayer() of
core/io/debayer.cc and the raw-frame conversion of
c_image_stacking_pipeline_base::read_input_frame (c_image_stacking_pipeline_base.cc:124-279).

Only tests/, ``__graft_entry__.smoke()`` and bench.py's CPU leg use this module;
the product never imports it.

numpy restatement.  Integer sources keep the reference's integer arithmetic
(int sums with the rounding constants c1 = 1, c2 = 2; AVGBP 2); float sources
sum in the reference's operand order in float32, then divide by 2 or 4.
"""
from __future__ import annotations

import numpy as np

DISABLE, NN, NN2, VNG, EA, SP, MATRIX, PLANE_0, PLANE_1, PLANE_2, PLANE_3, AVGBP = range(12)
DEFAULT = -1

# COLORID -> layout of the 2x2 cell, B = 0, G = 1, R = 2 (debayer.cc:828-1160); the
# CMY patterns use their RGB counterpart's code path
_CELLS = {
    8: ((2, 1), (1, 0)),   # RGGB
    9: ((1, 2), (0, 1)),   # GRBG
    10: ((1, 0), (2, 1)),  # GBRG
    11: ((0, 1), (1, 2)),  # BGGR
}
_CELLS[19], _CELLS[18], _CELLS[17], _CELLS[16] = _CELLS[8], _CELLS[9], _CELLS[10], _CELLS[11]


def is_bayer(colorid: int) -> bool:
    return colorid in _CELLS


def color_map(colorid: int, h: int, w: int) -> np.ndarray:
    cell = np.array(_CELLS[colorid])
    return np.tile(cell, (h // 2 + 1, w // 2 + 1))[:h, :w]


def _work(src):
    """(working array, integer?)"""
    if np.issubdtype(src.dtype, np.integer):
        return src.astype(np.int64), True
    return src.astype(np.float32), False


def _avg(terms, integer):
    n = len(terms)
    if integer:
        s = np.int64(n // 2)
        for t in terms:
            s = s + t
        return s // n
    s = np.float32(0)
    for t in terms:
        s = (s + t).astype(np.float32)
    return (s / np.float32(n)).astype(np.float32)


def nn2(src: np.ndarray, colorid: int) -> np.ndarray:
    """_debayer_nn2_interpolation (debayer.cc:828-1160): bilinear demosaic, BGR output of
    the source depth; neighbours mirrored without repeating the edge (rows y0 = 1 at the
    top, y2 = h - 2 at the bottom; columns likewise)."""
    h, w = src.shape
    W, integer = _work(src)
    P = np.pad(W, 1, mode="reflect")  # reflect = BORDER_REFLECT_101
    up, dn = P[:-2, 1:-1], P[2:, 1:-1]
    lf, rt = P[1:-1, :-2], P[1:-1, 2:]
    ul, ur, dl, dr = P[:-2, :-2], P[:-2, 2:], P[2:, :-2], P[2:, 2:]
    cm = color_map(colorid, h, w)
    out = np.zeros((h, w, 3), W.dtype)
    cross = _avg([up, lf, rt, dn], integer)
    diag = _avg([ul, ur, dl, dr], integer)
    horz = _avg([lf, rt], integer)
    vert = _avg([up, dn], integer)
    cm_left = np.pad(cm, 1, mode="reflect")[1:-1, :-2]
    cm_up = np.pad(cm, 1, mode="reflect")[:-2, 1:-1]
    for c in range(3):
        own = cm == c
        g_site = cm == 1
        v = np.where(own, W, 0)
        if c == 1:
            v = np.where(~own, cross, v)
        else:
            v = np.where(~own & ~g_site, diag, v)
            v = np.where(g_site & (cm_left == c), horz, v)
            v = np.where(g_site & (cm_up == c), vert, v)
        out[..., c] = v
    return out.astype(src.dtype)


def sp(src: np.ndarray, colorid: int) -> np.ndarray:
    """_debayer_sp (debayer.cc:587-700): one BGR pixel per 2x2 cell, G = (c1 + g + g) / 2."""
    h, w = src.shape
    W, integer = _work(src)
    cell = _CELLS[colorid]
    out = np.zeros((h // 2, w // 2, 3), W.dtype)
    greens = []
    for dy in range(2):
        for dx in range(2):
            v = W[dy::2, dx::2][: h // 2, : w // 2]
            c = cell[dy][dx]
            if c == 1:
                greens.append(v)
            else:
                out[..., c] = v
    out[..., 1] = _avg(greens, integer)
    return out.astype(src.dtype)


def avgbp(src: np.ndarray) -> np.ndarray:
    """_average_bayer_planes (debayer.cc:277-369): (c1 + s00 + s01 + s10 + s11) / 4."""
    W, integer = _work(src)
    t = [W[0::2, 0::2], W[0::2, 1::2], W[1::2, 0::2], W[1::2, 1::2]]
    return _avg(t, integer).astype(src.dtype)


def plane(src: np.ndarray, k: int) -> np.ndarray:
    """_extract_bayer_plane (debayer.cc:157-215): row (k >> 1) & 1, column k & 1 of each cell."""
    return np.ascontiguousarray(src[(k >> 1) & 1::2, k & 1::2])


def matrix(src: np.ndarray, colorid: int) -> np.ndarray:
    """_debayer_matrix (debayer.cc:455-575): each site's value in its own BGR channel."""
    h, w = src.shape
    cm = color_map(colorid, h, w)
    out = np.zeros((h, w, 3), src.dtype)
    for c in range(3):
        out[..., c] = np.where(cm == c, src, 0)
    return out


def _opencv_demosaic(src: np.ndarray, colorid: int, edge_aware: bool) -> np.ndarray:
    """cv::demosaicing of an 8/16-bit Bayer image to BGR (the reference's DEBAYER_NN and
    DEBAYER_EA, debayer.cc:1272-1355): OpenCV's Bayer2RGB_ (COLOR_Bayer*2BGR, bilinear) or
    Bayer2RGB_EdgeAware_T (COLOR_Bayer*2BGR_EA), restated from OpenCV's imgproc demosaicing.cpp
    (third-party, absent here).  Inside (rows and columns 1 .. n-2) each pixel keeps its own
    colour; at a green site the colour of its left/right neighbours is (L + R + 1) >> 1 and that of
    its up/down neighbours (U + D + 1) >> 1; at a red or blue site the opposite colour is
    (UL + UR + DL + DR + 2) >> 2 and green is (U + D + L + R + 2) >> 2 (bilinear) or, edge-aware,
    (U + D + 1) >> 1 when |L - R| > |D - U| and (L + R + 1) >> 1 otherwise.  Columns 0 and w-1
    copy columns 1 and w-2, then rows 0 and h-1 copy rows 1 and h-2; an image of 2 rows or
    columns or fewer comes out 0."""
    h, w = src.shape
    out = np.zeros((h, w, 3), src.dtype)
    if h <= 2 or w <= 2:
        return out
    s = src.astype(np.int64)
    c = s[1:-1, 1:-1]
    U, D_, L, R = s[:-2, 1:-1], s[2:, 1:-1], s[1:-1, :-2], s[1:-1, 2:]
    diag = (s[:-2, :-2] + s[:-2, 2:] + s[2:, :-2] + s[2:, 2:] + 2) >> 2
    hor, ver = (L + R + 1) >> 1, (U + D_ + 1) >> 1
    if edge_aware:
        g = np.where(np.abs(L - R) > np.abs(D_ - U), ver, hor)
    else:
        g = (U + D_ + L + R + 2) >> 2
    cm = color_map(colorid, h, w)
    site = cm[1:-1, 1:-1]
    hcol = cm[1:-1, :-2]  # colour of the left neighbour
    vcol = cm[:-2, 1:-1]  # colour of the upper neighbour
    inner = np.zeros((h - 2, w - 2, 3), np.int64)
    for k in range(3):
        green_site = np.where(hcol == k, hor, np.where(vcol == k, ver, 0))
        other_site = np.where(k == 1, g, diag)
        inner[..., k] = np.where(site == k, c, np.where(site == 1, green_site, other_site))
    full = np.empty((h, w, 3), np.int64)
    full[1:-1, 1:-1] = inner
    full[1:-1, 0] = inner[:, 0]
    full[1:-1, -1] = inner[:, -1]
    full[0] = full[1]
    full[-1] = full[-2]
    return full.astype(src.dtype)


def opencv_bilinear(src: np.ndarray, colorid: int) -> np.ndarray:
    """DEBAYER_NN on 8/16-bit data: cv::demosaicing(COLOR_Bayer*2BGR)."""
    return _opencv_demosaic(src, colorid, False)


def opencv_ea(src: np.ndarray, colorid: int) -> np.ndarray:
    """DEBAYER_EA on 8/16-bit data: cv::demosaicing(COLOR_Bayer*2BGR_EA)."""
    return _opencv_demosaic(src, colorid, True)


def debayer(src: np.ndarray, colorid: int, algo: int = DEFAULT) -> np.ndarray:
    """debayer() (debayer.cc:1197-1355) for one-channel sources."""
    if algo == DEFAULT:
        algo = NN2
    if algo == AVGBP:
        return avgbp(src)
    if PLANE_0 <= algo <= PLANE_3:
        return plane(src, algo - PLANE_0)
    integral = src.dtype in (np.uint8, np.uint16)
    if algo == DISABLE:
        return src.copy()
    if algo == NN and integral:
        return opencv_bilinear(src, colorid)
    if algo == EA and integral:
        return opencv_ea(src, colorid)
    if algo == VNG and src.dtype == np.uint8:
        raise NotImplementedError("OpenCV VNG demosaicing of 8-bit data")
    if algo in (NN, EA, VNG, NN2):
        return nn2(src, colorid)
    if algo == SP:
        return sp(src, colorid)
    if algo == MATRIX:
        return matrix(src, colorid)
    raise ValueError(algo)


def read_input_frame(raw: np.ndarray, colorid: int, bpp: int, algo: int = DEFAULT, dark=None, flat=None):
    """The conversion steps of read_input_frame (c_image_stacking_pipeline_base.cc:124-279):
    with dark/flat, convertTo(CV_32F, 1 / 2^bpp) then subtract / divide; debayer Bayer
    frames; convertTo(CV_32F, 1 / 2^bpp) when the image is not f32 yet."""
    scale = np.float32(1.0 / (1 << bpp))
    img = raw
    if dark is not None or flat is not None:
        if img.dtype != np.float32:
            img = (img.astype(np.float32) * scale).astype(np.float32)
        if dark is not None:
            img = (img - dark).astype(np.float32)
        if flat is not None:
            img = (img / flat).astype(np.float32)
    if is_bayer(colorid) and img.ndim == 2:
        img = debayer(img, colorid, algo)
    if img.dtype != np.float32:
        img = (img.astype(np.float32) * scale).astype(np.float32)
    return img

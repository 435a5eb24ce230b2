"""Focus stacking -- CPU oracle (test infrastructure only).

Restates, on the FP32 oracle's canonical pyrDown / pyrUp / sepFilter2D
(oracle/ssg_oracle.c) and numpy float32 elementwise arithmetic evaluated in
the order the reference writes it:

  * lpg() (core/proc/lpg.cc:63-127, 277-344) with dscale = uscale = 0: channel
    average, the 5x5 gradient and 5x5 laplacian stencils, the edge replication of
    compute_lpg_5x5_parallel, m = alpha*lap + beta*grad + eps, pow(m, p);
  * gaussian smoothing as cv::GaussianBlur(sigma) on float images: kernel size
    cvRound(8 sigma + 1) | 1, getGaussianKernel, BORDER_REFLECT101;
  * build / reconstruct_laplacian_pyramid (core/proc/laplacian_pyramid.cc:44-127),
    levels while the next size's minimum side stays above 8 (or a level cap);
  * c_laplacian_pyramid_focus_stacking::add / compute (core/average/
    c_frame_accumulation.cc:491-681): select_max_energy and weighted_average;
  * the focus configuration's composition: per-pixel weights W = gaussian(lpg),
    W reduced with pyrDown to every level, acc_i += W_i L_i, wsum_i += W_i on every
    level including the top one, result = reconstruct(acc_i / wsum_i).

Multi-channel images are planar lists of channels here.
"""
from __future__ import annotations

import numpy as np

F32 = np.float32
SELECT_MAX_ENERGY, WEIGHTED_AVERAGE, LPG_WEIGHTS = 0, 1, 2


def gray(planes):
    """reduce_color_channels(REDUCE_AVG): (c0 + c1) + c2 in float, times 1/3 in double."""
    if len(planes) == 1:
        return planes[0].astype(F32)
    s = (planes[0] + planes[1]) + planes[2]
    return (s.astype(np.float64) * (1.0 / 3.0)).astype(F32)


def lpg_coefficients(k: float):
    l_norm, g_norm = F32(100.0) / F32(4.0), F32(100.0) / F32(36.0)
    alpha = F32(k / (k + 1)) * (l_norm * l_norm)
    beta = F32(1.0 / (k + 1)) * (g_norm * g_norm)
    return alpha, beta, F32(1e-9)


def lpg(g: np.ndarray, k: float, p: float) -> np.ndarray:
    g = g.astype(F32)
    h, w = g.shape
    alpha, beta, eps = lpg_coefficients(k)
    ys = np.clip(np.arange(h), 2, h - 3)
    xs = np.clip(np.arange(w), 2, w - 3)

    def R(dy, dx):  # r_{2+dy}[x + dx] at every (clamped) output pixel
        return g[np.ix_(ys + dy, xs + dx)]

    two, four = F32(2), F32(4)
    col = lambda dx: (((R(-2, dx) + two * R(-1, dx)) + four * R(0, dx)) + two * R(1, dx)) + R(2, dx)
    row = lambda dy: (((R(dy, -2) + two * R(dy, -1)) + four * R(dy, 0)) + two * R(dy, 1)) + R(dy, 2)
    c3 = lambda dx: (R(-1, dx) + two * R(0, dx)) + R(1, dx)
    r3 = lambda dy: (R(dy, -1) + two * R(dy, 0)) + R(dy, 1)
    gx = (col(2) - col(-2)) + two * (c3(1) - c3(-1))
    gy = (row(2) - row(-2)) + two * (r3(1) - r3(-1))
    grad = gx * gx + gy * gy
    t = (((F32(16) * R(0, 0) - two * (((R(-1, 0) + R(1, 0)) + R(0, -1)) + R(0, 1)))
          - F32(1) * (((R(-1, -1) + R(-1, 1)) + R(1, -1)) + R(1, 1)))
         - F32(1) * (((R(-2, 0) + R(2, 0)) + R(0, -2)) + R(0, 2)))
    m = (alpha * (t * t) + beta * grad) + eps
    if p == 2:
        m = m * m
    elif p != 0 and p != 1:
        m = np.power(m, F32(p)).astype(F32)
    return m


def gaussian_ksize(sigma: float) -> int:
    return int(np.floor(sigma * 8 + 1 + 0.5)) | 1


def gaussian_blur(ora, img, sigma: float):
    n = gaussian_ksize(sigma)
    kern = ora.gaussian_kernel(n, sigma)
    return ora.sep_filter(img, kern, kern, 4)


def level_sizes(w: int, h: int, max_levels: int = 0, min_size: int = 8):
    sizes = [(w, h)]
    while True:
        nw, nh = (sizes[-1][0] + 1) // 2, (sizes[-1][1] + 1) // 2
        if min(nw, nh) <= min_size or (max_levels and len(sizes) >= max_levels):
            break
        sizes.append((nw, nh))
    return sizes


def build_laplacian(ora, img, sizes):
    lp = []
    cur = img.astype(F32)
    for (w, h), (nw, nh) in zip(sizes[:-1], sizes[1:]):
        down = ora.pyr_down(cur, (nw, nh))
        up = ora.pyr_up(down, (w, h))
        lp.append(cur - up)
        cur = down
    lp.append(cur)
    return lp


def reconstruct(ora, pyr):
    cur = pyr[-1]
    for lvl in reversed(pyr[:-1]):
        cur = ora.pyr_up(cur, (lvl.shape[1], lvl.shape[0])) + lvl
    return cur


class FocusStack:
    """c_laplacian_pyramid_focus_stacking (and the lpg-weight composition), planar channels."""

    def __init__(self, ora, policy=SELECT_MAX_ENERGY, avgchannel=True, kradius=0, ksigma=0.0, lpg_k=6.0, lpg_p=2.0,
                 lpg_sigma=5.0, max_levels=0):
        assert ora.bits == 32
        self.o = ora
        self.policy, self.avgc = policy, avgchannel
        self.G = ora.gaussian_kernel(max(0, 2 * kradius + 1), max(0.0, ksigma)) if (ksigma > 0 or kradius > 0) else None
        self.k, self.p, self.sigma, self.max_levels = lpg_k, lpg_p, lpg_sigma, max_levels
        self.acc = None

    def _sep(self, img, delta=0.0):
        return self.o.sep_filter(img, self.G, self.G, 4, delta) if self.G is not None else img

    def _energy(self, lap):
        """compute_energy (:511-531): w per channel, or the squared channel sum; lap *= w."""
        if not self.avgc or len(lap) == 1:
            ws = [l * l for l in lap]
            ws = [self._sep(x, 1e-12) if self.G is not None else x for x in ws]
            return [l * x for l, x in zip(lap, ws)], ws
        s = (lap[0] + lap[1]) + lap[2]
        w = s * s
        if self.G is not None:
            w = self._sep(w, 1e-12)
        return [l * w for l in lap], [w]

    @staticmethod
    def _graystdev(ora, planes):
        return sum(ora.mean_stddev(p)[1] for p in planes)

    def add(self, planes):
        o = self.o
        h, w = planes[0].shape
        sizes = level_sizes(w, h, self.max_levels)
        pyr = [build_laplacian(o, p, sizes) for p in planes]  # [channel][level]
        L = len(sizes)
        if self.policy == LPG_WEIGHTS:
            W = gaussian_blur(o, lpg(gray(planes), self.k, self.p), self.sigma)
            Ws = [W]
            for (nw, nh) in sizes[1:]:
                Ws.append(o.pyr_down(Ws[-1], (nw, nh)))
            if self.acc is None:
                self.acc = [[np.zeros_like(pyr[c][i]) for i in range(L)] for c in range(len(planes))]
                self.wsum = [np.zeros_like(Ws[i]) for i in range(L)]
            for i in range(L):
                for c in range(len(planes)):
                    self.acc[c][i] = self.acc[c][i] + Ws[i] * pyr[c][i]
                self.wsum[i] = self.wsum[i] + Ws[i]
            return
        if self.acc is None:
            self.acc = pyr
            if self.policy == WEIGHTED_AVERAGE:
                self.wwp = []
                for i in range(L - 1):
                    lap, ws = self._energy([pyr[c][i] for c in range(len(planes))])
                    for c in range(len(planes)):
                        self.acc[c][i] = lap[c]
                    self.wwp.append(ws)
            return
        s0 = self._graystdev(o, [self.acc[c][-1] for c in range(len(planes))])
        s1 = self._graystdev(o, [pyr[c][-1] for c in range(len(planes))])
        a0, a1 = F32(s0 / (s0 + s1)), F32(s1 / (s0 + s1))
        for c in range(len(planes)):
            self.acc[c][-1] = self.acc[c][-1] * a0 + pyr[c][-1] * a1
        for i in range(L - 1):
            if self.policy == SELECT_MAX_ENERGY:
                for c in range(len(planes)):
                    w0, w1 = np.abs(self.acc[c][i]), np.abs(pyr[c][i])
                    if self.G is not None:
                        w0, w1 = self._sep(w0), self._sep(w1)
                    self.acc[c][i] = np.where(w1 > w0, pyr[c][i], self.acc[c][i])
            else:
                lap, ws = self._energy([pyr[c][i] for c in range(len(planes))])
                for c in range(len(planes)):
                    self.acc[c][i] = lap[c] + self.acc[c][i]
                self.wwp[i] = [ww + wp for ww, wp in zip(ws, self.wwp[i])]

    def compute(self):
        o = self.o
        out = []
        for c in range(len(self.acc)):
            if self.policy == SELECT_MAX_ENERGY:
                lv = self.acc[c]
            elif self.policy == WEIGHTED_AVERAGE:
                lv = [self.acc[c][i] / self.wwp[i][c if len(self.wwp[i]) > 1 else 0] for i in range(len(self.acc[c]) - 1)]
                lv.append(self.acc[c][-1])
            else:
                lv = [self.acc[c][i] / self.wsum[i] for i in range(len(self.acc[c]))]
            out.append(reconstruct(o, lv))
        return out

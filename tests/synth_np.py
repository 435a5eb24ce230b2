"""Small seeded planetary-disk frames in numpy (test inputs), and the warp measures the
registration tests compare with.

A limb-darkened disk (I = 1 - 0.6 (1 - mu)) textured by gaussian-filtered
uniform noise at 10%, warped by a known affine map, optionally blurred and
with gaussian noise.  ``frame(x) = scene(T x)``; the registration of a frame
onto the un-warped scene is ``W = T^-1`` (returned as ``truth``).
"""
from __future__ import annotations

import numpy as np
from scipy import ndimage


def scene(w: int, h: int, radius: float, seed: int = 1, tex_sigma: float = 3.0, amp: float = 0.1,
          margin: int = 32):
    rng = np.random.default_rng(seed)
    tw, th = w + 2 * margin, h + 2 * margin
    tex = ndimage.gaussian_filter(rng.uniform(-1, 1, (th, tw)), tex_sigma)
    tex /= np.abs(tex).max()
    cx, cy = 0.5 * (w - 1), 0.5 * (h - 1)

    def render(A):
        yy, xx = np.mgrid[0:h, 0:w].astype(np.float64)
        u = A[0, 0] * xx + A[0, 1] * yy + A[0, 2]
        v = A[1, 0] * xx + A[1, 1] * yy + A[1, 2]
        r = np.hypot(u - cx, v - cy) / radius
        mu = np.sqrt(np.clip(1 - r * r, 0, 1))
        disk = np.where(r < 1, 1 - 0.6 * (1 - mu), 0.0)
        t = ndimage.map_coordinates(tex, [v + margin, u + margin], order=3, mode="nearest")
        return disk * (1 + amp * t)

    return render


def random_affine(rng, w, h, shift, rot_deg, scale_dev=0.0, shear=0.0):
    cx, cy = 0.5 * (w - 1), 0.5 * (h - 1)
    th = np.deg2rad(rng.uniform(-rot_deg, rot_deg))
    s = 1 + rng.uniform(-scale_dev, scale_dev)
    hs = rng.uniform(-shear, shear)
    R = np.array([[np.cos(th), -np.sin(th)], [np.sin(th), np.cos(th)]])
    M = R @ np.array([[s, hs], [0, s]])
    t = rng.uniform(-shift, shift, 2)
    A = np.zeros((2, 3))
    A[:, :2] = M
    A[:, 2] = np.array([cx, cy]) + t - M @ np.array([cx, cy])
    return A


def invert(A):
    M = np.vstack([A, [0, 0, 1]])
    return np.linalg.inv(M)[:2]


def frames(n, w, h, radius, seed=7, shift=4.0, rot_deg=0.5, scale_dev=0.0, shear=0.0, blur=(0.0, 0.0),
           noise=0.0, scale=1.0, dtype=np.float32):
    """Returns (reference, frames list, truth list of 2x3 ref->frame maps)."""
    render = scene(w, h, radius, seed)
    rng = np.random.default_rng(seed + 1)
    ref = render(np.array([[1.0, 0, 0], [0, 1.0, 0]]))
    if blur[1] > 0:
        ref = ndimage.gaussian_filter(ref, 0.5 * (blur[0] + blur[1]))
    out, truth = [], []
    for _ in range(n):
        A = random_affine(rng, w, h, shift, rot_deg, scale_dev, shear)
        f = render(A)
        if blur[1] > 0:
            f = ndimage.gaussian_filter(f, rng.uniform(*blur))
        if noise > 0:
            f = f + rng.normal(0, noise, f.shape)
        out.append((f * scale).astype(dtype))
        truth.append(invert(A))
    return (ref * scale).astype(dtype), out, truth


def max_displacement(m1, m2, w, h):
    """Max over the frame corners (the extremes of an affine difference) of |m1 x - m2 x|."""
    m1 = np.asarray(m1, np.float64).reshape(2, 3)
    m2 = np.asarray(m2, np.float64).reshape(2, 3)
    c = np.array([[0, 0, 1], [w - 1, 0, 1], [0, h - 1, 1], [w - 1, h - 1, 1]], np.float64).T
    d = (m1 - m2) @ c
    return float(np.sqrt((d ** 2).sum(0)).max())


def blob_scene(w: int, h: int, seed: int, n: int = 60, smin: float = 3.0, smax: float = 7.0):
    """A band-limited analytic scene: a sum of gaussian blobs (sigma smin..smax px).
    ``render(T)`` evaluates scene(T x) exactly at every pixel, so a frame rendered
    with T is registered onto render(identity) by W = T^-1 with no sampling bias."""
    r = np.random.default_rng(seed)
    xs, ys = r.uniform(-10, w + 10, n), r.uniform(-10, h + 10, n)
    ss, aa = r.uniform(smin, smax, n), r.uniform(0.2, 1.0, n)

    def render(T):
        T = np.asarray(T, np.float64).reshape(2, 3)
        yy, xx = np.mgrid[0:h, 0:w].astype(np.float64)
        u = T[0, 0] * xx + T[0, 1] * yy + T[0, 2]
        v = T[1, 0] * xx + T[1, 1] * yy + T[1, 2]
        return at(u, v)

    def at(u, v):
        """The scene at arbitrary float64 coordinates (arrays u, v)."""
        out = np.zeros(np.shape(u))
        for x, y, s, a in zip(xs, ys, ss, aa):
            out += a * np.exp(-((u - x) ** 2 + (v - y) ** 2) / (2 * s * s))
        return out

    render.at = at
    return render


def warp_bar(d32, tol=0.01):
    """The warp bar of a registration run to convergence: `tol` px, or twice the FP32 spread
    where that is larger."""
    return max(tol, 2.0 * d32)


def fp32_spread(oreg32, o64, f, mask, w, h, seed, copies=8):
    """Farthest end point of the FP32 oracle from the FP64 one, over the frame and `copies`
    copies with each pixel scaled by (1 +- 2^-23): how far the reference algorithm at its own
    precision is determined on this frame."""
    rng = np.random.default_rng(seed)
    d = 0.0
    for k in range(copies + 1):
        fk = f if k == 0 else (f * (1.0 + rng.choice([-1.0, 1.0], f.shape) * 2.0**-23)).astype(np.float32)
        o32 = oreg32.register_frame(fk, mask)
        if o32.ok:
            d = max(d, max_displacement(o32.matrix, o64.matrix, w, h))
    return d

"""FFT phase-correlation translation estimate -- CPU oracle (test infrastructure only).

SerStacker has no phase correlation on this path (its nearest code is
core/proc/fft.cc:17-160: fftGetOptimalSize, fftCopyMakeBorder,
fftImageToSpectrum); the GPU path defines it and this numpy float64
restatement is its oracle:

  1. optional 2x2 binning (mean of each cell; the luminance of a raw Bayer frame),
  2. mean removal and separable Hann apodisation w(i) = 0.5 - 0.5 cos(2 pi i / (n - 1))
     (without the mean, the window itself -- common to both images -- correlates at
     zero shift),
  3. zero padding to the next powers of two (rows N, columns M),
  4. R = F conj(G) / |F conj(G)| (F: frame spectrum, G: reference spectrum), weighted
     by exp(-2 pi^2 s^2 (kx^2 + ky^2)) (kx, ky in cycles per binned pixel, s = 1): the
     correlation peak becomes a gaussian of sigma 1 px and the noise-only high
     frequencies lose their weight,
  5. r = real(ifft2(R)); integer peak; 3-point gaussian refinement per axis
     (parabola through the logarithms; a plain parabola if a value is not positive),
  6. shifts wrap to (-M/2, M/2] and (-N/2, N/2], times the binning.

The result T is the translation of the registration convention ref(x) =
frame(x + T) (the translation transform's parameters, c_image_transform.cc:150-170).
"""
from __future__ import annotations

import numpy as np


def next_pow2(n: int) -> int:
    p = 1
    while p < n:
        p *= 2
    return p


def hann(n: int) -> np.ndarray:
    if n == 1:
        return np.ones(1)
    return 0.5 - 0.5 * np.cos(2.0 * np.pi * np.arange(n) / (n - 1))


def bin2(img: np.ndarray) -> np.ndarray:
    h, w = img.shape[0] // 2, img.shape[1] // 2
    a = np.asarray(img, np.float64)[:2 * h, :2 * w]
    return 0.25 * (a[0::2, 0::2] + a[0::2, 1::2] + a[1::2, 0::2] + a[1::2, 1::2])


def prepare(img: np.ndarray, binning: int = 1, window: bool = True) -> np.ndarray:
    a = bin2(img) if binning == 2 else np.asarray(img, np.float64)
    h, w = a.shape
    if window:
        a = (a - a.mean()) * np.outer(hann(h), hann(w))
    out = np.zeros((next_pow2(h), next_pow2(w)))
    out[:h, :w] = a
    return out


def spectrum(img: np.ndarray, binning: int = 1, window: bool = True) -> np.ndarray:
    return np.fft.fft2(prepare(img, binning, window))


PEAK_SIGMA = 1.0


def parabolic(rm: float, r0: float, rp: float) -> float:
    if rm > 0 and r0 > 0 and rp > 0:
        rm, r0, rp = np.log(rm), np.log(r0), np.log(rp)
    d = rm - 2.0 * r0 + rp
    return 0.0 if d == 0.0 else 0.5 * (rm - rp) / d


def peak_weight(N: int, M: int) -> np.ndarray:
    ky = np.fft.fftfreq(N)[:, None]
    kx = np.fft.fftfreq(M)[None, :]
    return np.exp(-2.0 * (np.pi * PEAK_SIGMA) ** 2 * (kx * kx + ky * ky))


def estimate(ref_spec: np.ndarray, frame: np.ndarray, binning: int = 1, window: bool = True):
    """Returns (dx, dy, peak) for one frame against a reference spectrum."""
    F = spectrum(frame, binning, window)
    X = F * np.conj(ref_spec)
    mag = np.abs(X)
    R = np.where(mag > 0, X / np.where(mag > 0, mag, 1), 0) * peak_weight(*X.shape)
    r = np.real(np.fft.ifft2(R))
    N, M = r.shape
    py, px = np.unravel_index(int(np.argmax(r)), r.shape)
    cx = parabolic(r[py, (px - 1) % M], r[py, px], r[py, (px + 1) % M])
    cy = parabolic(r[(py - 1) % N, px], r[py, px], r[(py + 1) % N, px])
    sx = px - M if px > M // 2 else px
    sy = py - N if py > N // 2 else py
    return (sx + cx) * binning, (sy + cy) * binning, float(r[py, px])

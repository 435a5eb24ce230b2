// ssg_stencil.hpp -- per-pixel stencil formulas shared by the pyramid, focus and
// multi-sequence focus kernels, each in the canonical order of ssg_oracle.c:
// sepFilter2D row / column passes, pyrUp, the lpg 5x5 measure and the channel average.
#pragma once

#include "ssg_engine.hpp"

namespace ssg {

// ---------------------------------------------------------------------------
// Separable filter, canonical order of ssg_oracle.c:ora_sep_filter
// (the OpenCV split SymmRowSmallFilter / RowFilter / SymmColumnFilter,
// filter.simd.hpp, with fused multiply-adds).
// ---------------------------------------------------------------------------
__device__ __forceinline__ float row_filter(const float* c, const SepKernel& K) {
  // c points at the centre tap
  const int n = K.nx, a = n / 2;
  const float* k = K.kx;
  if (K.tx == 1 && n <= 5) {
    float s = c[0] * k[a];
    if (n >= 3) s = fmaf(c[-1] + c[1], k[a + 1], s);
    if (n >= 5) s = fmaf(c[-2] + c[2], k[a + 2], s);
    return s;
  }
  if (K.tx == 2 && n >= 3 && n <= 5) {
    float s = (c[1] - c[-1]) * k[a + 1];
    if (n >= 5) s = fmaf(c[2] - c[-2], k[a + 2], s);
    return s;
  }
  float s = 0.f;
  for (int i = 0; i < n; ++i) s = fmaf(c[i - a], k[i], s);
  return s;
}

__device__ __forceinline__ float col_filter(const float* c, int stride, const SepKernel& K) {
  const int n = K.ny, a = n / 2;
  const float* k = K.ky;
  if (K.ty == 1) {
    float s = fmaf(c[0], k[a], K.delta);
    for (int i = 1; i <= a; ++i) s = fmaf(c[-i * stride] + c[i * stride], k[a + i], s);
    return s;
  }
  if (K.ty == 2) {
    float s = K.delta;
    for (int i = 1; i <= a; ++i) s = fmaf(c[i * stride] - c[-i * stride], k[a + i], s);
    return s;
  }
  float s = K.delta;
  for (int i = 0; i < n; ++i) s = fmaf(c[(i - a) * stride], k[i], s);
  return s;
}


// cv::pyrUp (float) at output pixel (xo, yo) of a dw x dh image from the sw x sh plane S
// (rows of `pitch` floats): horizontally upsampled rows (left REFLECT101, right replicated),
// then the vertical pass; an odd extra column repeats the last one, an odd extra row repeats
// row 2*sh-2.  The oracle's canonical form (ssg_oracle.c:ora_pyr_up).
__device__ __forceinline__ float pyrup_at(const float* S, int pitch, int sw, int sh, int dh, int xo, int yo) {
  const int ye = (dh > 2 * sh && yo == 2 * sh) ? 2 * sh - 2 : yo;  // the extra odd row copies row 2 sh - 2
  const int y = ye >> 1;
  auto srow = [&](int r) {  // source rows -1..sh: REFLECT101 above, replicated below
    if (r < 0) r = sh > 1 ? 1 : 0;
    if (r >= sh) r = sh - 1;
    return S + (size_t)r * pitch;
  };
  const float scale = 1.f / 64.f;
  if (ye & 1) {
    // a row 2y+1 that exists only as min(2y+1, dh-1) when dh is odd and smaller is
    // written by the even branch (the oracle writes t1 first, then t0)
    return (hup(srow(y), sw, xo) + hup(srow(y + 1), sw, xo)) * 4.f * scale;
  }
  const float r0 = hup(srow(y - 1), sw, xo), r1 = hup(srow(y), sw, xo), r2 = hup(srow(y + 1), sw, xo);
  return fmaf(r1, 6.f, r0 + r2) * scale;
}

// compute_lpg_5x5_parallel (lpg.cc:63-127) at pixel (x, y) of the w x h plane g (rows of `pitch`
// floats), on the clamped grid (its edge rows/columns copy the nearest computed pixel), then
// pow(m, p)
__device__ __forceinline__ float lpg_at(const float* g, int pitch, int w, int h, int x, int y, float alpha, float beta,
                                        float eps, float p) {
  const int cy = min(max(y, 2), h - 3), cx = min(max(x, 2), w - 3);
  auto R = [&](int dy, int dx) { return g[(size_t)(cy + dy) * pitch + cx + dx]; };
  const float two = 2.f, four = 4.f;
  auto col = [&](int dx) { return (((R(-2, dx) + two * R(-1, dx)) + four * R(0, dx)) + two * R(1, dx)) + R(2, dx); };
  auto row = [&](int dy) { return (((R(dy, -2) + two * R(dy, -1)) + four * R(dy, 0)) + two * R(dy, 1)) + R(dy, 2); };
  auto c3 = [&](int dx) { return (R(-1, dx) + two * R(0, dx)) + R(1, dx); };
  auto r3 = [&](int dy) { return (R(dy, -1) + two * R(dy, 0)) + R(dy, 1); };
  const float gx = (col(2) - col(-2)) + two * (c3(1) - c3(-1));
  const float gy = (row(2) - row(-2)) + two * (r3(1) - r3(-1));
  const float grad = gx * gx + gy * gy;
  const float t = ((16.f * R(0, 0) - two * (((R(-1, 0) + R(1, 0)) + R(0, -1)) + R(0, 1))) -
                   1.f * (((R(-1, -1) + R(-1, 1)) + R(1, -1)) + R(1, 1))) -
                  1.f * (((R(-2, 0) + R(2, 0)) + R(0, -2)) + R(0, 2));
  float v = (alpha * (t * t) + beta * grad) + eps;
  if (p == 2.f) v = v * v;
  else if (p != 0.f && p != 1.f) v = powf(v, p);
  return v;
}

// reduce_color_channels(REDUCE_AVG) of one pixel: (c0 + c1) + c2 in float, times 1/3 in double
__device__ __forceinline__ float gray3(float c0, float c1, float c2) {
  const float s = (c0 + c1) + c2;
  return (float)((double)s * (1.0 / 3.0));
}

// ---- per-channel bilinear samplers (cv::remap INTER_LINEAR of interleaved CN-channel rows) --------------------------------------------------------------------
template <int CN>
__device__ __forceinline__ void lin_fast(const float* __restrict__ img, int pitch, int X, int Y, float* out) {
  const int ix = X >> 5, iy = Y >> 5;
  float w0, w1, w2, w3;
  bilinear_w(X & 31, Y & 31, w0, w1, w2, w3);
  const float* r0 = img + (size_t)iy * pitch + ix * CN;
  const float* r1 = r0 + pitch;
#pragma unroll
  for (int c = 0; c < CN; ++c) out[c] = bilinear_sum(r0[c], r0[CN + c], r1[c], r1[CN + c], w0, w1, w2, w3);
}

template <int CN>
__device__ __forceinline__ bool lin_general(const float* __restrict__ img, int pitch, int w, int h, int X, int Y,
                                            int border, float cval, float* out) {
  const int ix = sat_short(X >> 5), iy = sat_short(Y >> 5);
  if ((unsigned)ix < (unsigned)(w - 1) && (unsigned)iy < (unsigned)(h - 1)) {
    lin_fast<CN>(img, pitch, X, Y, out);
    return true;
  }
  if (border == BORDER_CONSTANT && (ix >= w || ix + 1 < 0 || iy >= h || iy + 1 < 0)) {
#pragma unroll
    for (int c = 0; c < CN; ++c) out[c] = cval;
    return true;
  }
  if (border == BORDER_TRANSPARENT) return false;
  float w0, w1, w2, w3;
  bilinear_w(X & 31, Y & 31, w0, w1, w2, w3);
  const int x0 = border_interp(ix, w, border), x1 = border_interp(ix + 1, w, border);
  const int y0 = border_interp(iy, h, border), y1 = border_interp(iy + 1, h, border);
#pragma unroll
  for (int c = 0; c < CN; ++c) {
    const float s00 = (x0 >= 0 && y0 >= 0) ? img[(size_t)y0 * pitch + x0 * CN + c] : cval;
    const float s01 = (x1 >= 0 && y0 >= 0) ? img[(size_t)y0 * pitch + x1 * CN + c] : cval;
    const float s10 = (x0 >= 0 && y1 >= 0) ? img[(size_t)y1 * pitch + x0 * CN + c] : cval;
    const float s11 = (x1 >= 0 && y1 >= 0) ? img[(size_t)y1 * pitch + x1 * CN + c] : cval;
    out[c] = bilinear_sum(s00, s01, s10, s11, w0, w1, w2, w3);
  }
  return true;
}

}  // namespace ssg

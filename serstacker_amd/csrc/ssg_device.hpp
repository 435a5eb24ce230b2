// ssg_device.hpp -- device building blocks shared by the gfx950 kernels.
//
// Canonical arithmetic: every per-pixel formula below fixes one order of
// operations (explicit fmaf, compiled with -ffp-contract=off) so that results
// do not depend on compiler contraction.  The orders follow OpenCV 4.5.x
// (imgwarp.cpp remapBilinear/remapBicubic, filter.simd.hpp, pyramids.cpp)
// with its multiply-adds fused.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace ssg {

enum : int { BORDER_CONSTANT = 0, BORDER_REPLICATE = 1, BORDER_REFLECT = 2, BORDER_WRAP = 3,
             BORDER_REFLECT101 = 4, BORDER_TRANSPARENT = 5 };
enum : int { INTER_NEAREST = 0, INTER_LINEAR = 1, INTER_CUBIC = 2 };
enum : int { MOTION_TRANSLATION = 0, MOTION_EUCLIDEAN = 1, MOTION_SCALED_EUCLIDEAN = 2, MOTION_AFFINE = 3 };
// internal: a dense per-frame remap field (SSG_32FC2) instead of a parametric transform
constexpr int MOTION_MAP = 4;

constexpr int WAVE = 64;

// Raw buffer access: a 32-bit per-lane byte offset plus a wave-uniform byte
// offset (SGPR) and a small immediate, instead of a 64-bit address per lane.
// Callers keep every access inside [0, bytes): the hardware range check does
// not cover the uniform part, so it is never relied on.
using rsrc_t = __amdgpu_buffer_rsrc_t;
__device__ __forceinline__ rsrc_t make_rsrc(const void* base, uint32_t bytes) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), (short)0, (int)bytes, 0x00020000);
}
__device__ __forceinline__ float bload_f32(rsrc_t r, uint32_t voff, uint32_t soff) {
  return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, voff, soff, 0));
}

// cv::borderInterpolate; -1 for CONSTANT/TRANSPARENT
__device__ __forceinline__ int border_interp(int p, int len, int border) {
  if ((unsigned)p < (unsigned)len) return p;
  if (border == BORDER_REPLICATE) return p < 0 ? 0 : len - 1;
  if (border == BORDER_REFLECT || border == BORDER_REFLECT101) {
    const int delta = border == BORDER_REFLECT101;
    if (len == 1) return 0;
    // bounded: each step maps p strictly closer to [0, len) for |p| < 2^30
    for (int it = 0; it < 64 && (unsigned)p >= (unsigned)len; ++it) {
      if (p < 0) p = -p - 1 + delta;
      else p = len - 1 - (p - len) - delta;
    }
    return ((unsigned)p < (unsigned)len) ? p : 0;
  }
  if (border == BORDER_WRAP) {
    if (p < 0) p -= ((p - len + 1) / len) * len;
    if (p >= len) p %= len;
    return p;
  }
  return -1;
}

// cvRound(v * 32) with saturation to +-2^30 (NaN maps to the lower bound)
__device__ __forceinline__ int quant32(float v) {
  float s = v * 32.f;
  s = fminf(fmaxf(s, -1073741824.f), 1073741824.f);
  return __float2int_rn(s);
}

// quant32 for |v| < 2^17 as the bits of one fma: v*32 is exact, and adding 1.5*2^23 rounds it to
// the nearest integer, ties to even, into the low mantissa bits, so the result's bits are
// 0x4B400000 + quant32(v): bits 0..4 are quant32(v) & 31, and ubfe(bits, 5, 18) is
// (quant32(v) >> 5) + 2^17
__device__ __forceinline__ uint32_t quant32_bits(float v) { return __float_as_uint(fmaf(v, 32.f, 12582912.f)); }

__device__ __forceinline__ int sat_short(int v) { return min(max(v, -32768), 32767); }

// INTER_NEAREST position of a float map coordinate: saturate_cast<short>(float),
// round to nearest even; out-of-range values (NaN included) land outside any image
__device__ __forceinline__ int round_coord(float v) { return __float2int_rn(fminf(fmaxf(v, -1e9f), 1e9f)); }

// cv::resize INTER_LINEAR tap of one axis (resize.cpp resize(): fx = (d+0.5)*scale - 0.5, the
// taps clamped to the source)
__device__ __forceinline__ void lin_tap(int d, double scale, int ssize, int& s0, int& s1, float& a0, float& a1) {
  float f = (float)(((double)d + 0.5) * scale - 0.5);
  int s = (int)floorf(f);
  f -= (float)s;
  if (s < 0) {
    f = 0.f;
    s = 0;
  }
  if (s >= ssize - 1) {
    f = 0.f;
    s = ssize - 1;
  }
  s0 = s;
  s1 = min(s + 1, ssize - 1);
  a0 = 1.f - f;
  a1 = f;
}

// The rows of the same resize: fy is never reset; the rows sy and sy + 1 are clipped to the
// source and keep the weights (1 - fy, fy) (resize.cpp resizeGeneric_Invoker, clip(sy, 0, h))
__device__ __forceinline__ void lin_tap_v(int d, double scale, int ssize, int& s0, int& s1, float& b0, float& b1) {
  float f = (float)(((double)d + 0.5) * scale - 0.5);
  const int s = (int)floorf(f);
  f -= (float)s;
  s0 = min(max(s, 0), ssize - 1);
  s1 = min(max(s + 1, 0), ssize - 1);
  b0 = 1.f - f;
  b1 = f;
}

// cv::pyrUp (float), horizontal pass: the upsampled value at output column xo of source row S
// (ST floats between pixels): [1 6 1] at even columns, [4 4] at odd ones, left REFLECT101, right
// replicated; an odd extra column repeats the last one
template <int ST = 1>
__device__ __forceinline__ float hup(const float* S, int sw, int xo) {
  if (sw == 1) return S[0] * 8.f;
  if (xo >= 2 * sw) xo = 2 * sw - 1;
  const int x = xo >> 1;
  if (xo & 1) {
    if (x == sw - 1) return S[(sw - 1) * ST] * 8.f;
    return (S[x * ST] + S[(x + 1) * ST]) * 4.f;
  }
  if (x == 0) return fmaf(S[0], 6.f, S[ST] + S[ST]);
  if (x == sw - 1) return fmaf(S[(sw - 1) * ST], 6.f, S[(sw - 2) * ST] + S[(sw - 1) * ST]);
  return fmaf(S[x * ST], 6.f, S[(x - 1) * ST] + S[(x + 1) * ST]);
}

// sum of p[t * elem_stride] for t = t0, t0 + stride, ... < n, added in that order, with eight
// loads in flight ahead of the additions (the same sum as the plain loop, bit for bit)
__device__ __forceinline__ double ordered_strided_sum(const double* p, int t0, int n, int stride, int elem_stride) {
  double s = 0;
  int t = t0;
  for (; t + 7 * stride < n; t += 8 * stride) {
    double v[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) v[u] = p[(size_t)(t + u * stride) * elem_stride];
#pragma unroll
    for (int u = 0; u < 8; ++u) s += v[u];
  }
  for (; t < n; t += stride) s += p[(size_t)t * elem_stride];
  return s;
}

// bilinear float weights at fraction indices (fx, fy) in 1/32 px
__device__ __forceinline__ void bilinear_w(int fx, int fy, float& w0, float& w1, float& w2, float& w3) {
  const float tx = (float)fx * (1.f / 32.f), ty = (float)fy * (1.f / 32.f);
  const float ox = 1.f - tx, oy = 1.f - ty;
  w0 = oy * ox;
  w1 = oy * tx;
  w2 = ty * ox;
  w3 = ty * tx;
}

// the weights of the 32 x 32 sub-pixel positions, entry fy * 32 + fx (cv::remap's bilinear table),
// filled by the nt threads of a block into LDS; the caller synchronises before reading it
__device__ __forceinline__ void fill_bilinear_table(float4* t, int tid, int nt) {
  for (int i = tid; i < 1024; i += nt) {
    float w0, w1, w2, w3;
    bilinear_w(i & 31, i >> 5, w0, w1, w2, w3);
    t[i] = make_float4(w0, w1, w2, w3);
  }
}

__device__ __forceinline__ float bilinear_sum(float s00, float s01, float s10, float s11, float w0, float w1,
                                              float w2, float w3) {
  float v = s00 * w0;
  v = fmaf(s01, w1, v);
  v = fmaf(s10, w2, v);
  v = fmaf(s11, w3, v);
  return v;
}

// INTER_LINEAR sample of a float image at quantised position (X, Y), BORDER_REPLICATE
__device__ __forceinline__ float sample_linear_replicate(const float* __restrict__ img, int pitch, int w, int h,
                                                         int X, int Y) {
  const int ix = sat_short(X >> 5), iy = sat_short(Y >> 5);
  float w0, w1, w2, w3;
  bilinear_w(X & 31, Y & 31, w0, w1, w2, w3);
  const int x0 = min(max(ix, 0), w - 1), x1 = min(max(ix + 1, 0), w - 1);
  const int y0 = min(max(iy, 0), h - 1), y1 = min(max(iy + 1, 0), h - 1);
  const float* r0 = img + (size_t)y0 * pitch;
  const float* r1 = img + (size_t)y1 * pitch;
  return bilinear_sum(r0[x0], r0[x1], r1[x0], r1[x1], w0, w1, w2, w3);
}

// INTER_LINEAR sample with a general border.  Returns false for TRANSPARENT
// pixels that need an outside tap (the destination is left untouched).
__device__ __forceinline__ bool sample_linear(const float* __restrict__ img, int pitch, int w, int h, int X, int Y,
                                              int border, float cval, float& out) {
  const int ix = sat_short(X >> 5), iy = sat_short(Y >> 5);
  if ((unsigned)ix < (unsigned)(w - 1) && (unsigned)iy < (unsigned)(h - 1)) {
    float w0, w1, w2, w3;
    bilinear_w(X & 31, Y & 31, w0, w1, w2, w3);
    const float* r0 = img + (size_t)iy * pitch + ix;
    out = bilinear_sum(r0[0], r0[1], r0[pitch], r0[pitch + 1], w0, w1, w2, w3);
    return true;
  }
  if (border == BORDER_CONSTANT && (ix >= w || ix + 1 < 0 || iy >= h || iy + 1 < 0)) {
    out = cval;
    return true;
  }
  float w0, w1, w2, w3;
  bilinear_w(X & 31, Y & 31, w0, w1, w2, w3);
  const int x0 = border_interp(ix, w, border), x1 = border_interp(ix + 1, w, border);
  const int y0 = border_interp(iy, h, border), y1 = border_interp(iy + 1, h, border);
  if (border == BORDER_TRANSPARENT) return false;
  const float s00 = (x0 >= 0 && y0 >= 0) ? img[(size_t)y0 * pitch + x0] : cval;
  const float s01 = (x1 >= 0 && y0 >= 0) ? img[(size_t)y0 * pitch + x1] : cval;
  const float s10 = (x0 >= 0 && y1 >= 0) ? img[(size_t)y1 * pitch + x0] : cval;
  const float s11 = (x1 >= 0 && y1 >= 0) ? img[(size_t)y1 * pitch + x1] : cval;
  out = bilinear_sum(s00, s01, s10, s11, w0, w1, w2, w3);
  return true;
}

// Keys cubic (a = -0.75) coefficients at t = i/32, computed on the host with
// the same float expression as OpenCV's interpolateCubic and uploaded by
// set_cubic_table() of the translation unit that samples with it.
static __constant__ float c_cubic_tab[32][4];

__device__ __forceinline__ bool sample_cubic(const float* __restrict__ img, int pitch, int w, int h, int X, int Y,
                                             int border, float cval, float& out) {
  const int bx = sat_short(X >> 5) - 1, by = sat_short(Y >> 5) - 1;
  const float* cx = c_cubic_tab[X & 31];
  const float* cy = c_cubic_tab[Y & 31];
  if ((unsigned)bx < (unsigned)(w - 3) && (unsigned)by < (unsigned)(h - 3)) {
    const float* r = img + (size_t)by * pitch + bx;
    float sum = 0.f;
#pragma unroll
    for (int j = 0; j < 4; ++j, r += pitch) {
      float t = r[0] * (cy[j] * cx[0]);
      t = fmaf(r[1], cy[j] * cx[1], t);
      t = fmaf(r[2], cy[j] * cx[2], t);
      t = fmaf(r[3], cy[j] * cx[3], t);
      sum = j == 0 ? t : sum + t;
    }
    out = sum;
    return true;
  }
  if (border == BORDER_CONSTANT && (bx >= w || bx + 4 <= 0 || by >= h || by + 4 <= 0)) {
    out = cval;
    return true;
  }
  if (border == BORDER_TRANSPARENT) return false;
  int xs[4];
#pragma unroll
  for (int i = 0; i < 4; ++i) xs[i] = border_interp(bx + i, w, border);
  float sum = 0.f;
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    const int yy = border_interp(by + j, h, border);
    const float* r = img + (size_t)(yy < 0 ? 0 : yy) * pitch;
    float v[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) v[i] = (yy >= 0 && xs[i] >= 0) ? r[xs[i]] : cval;
    float t = v[0] * (cy[j] * cx[0]);
    t = fmaf(v[1], cy[j] * cx[1], t);
    t = fmaf(v[2], cy[j] * cx[2], t);
    t = fmaf(v[3], cy[j] * cx[3], t);
    sum = j == 0 ? t : sum + t;
  }
  out = sum;
  return true;
}

// 15-bit fixed-point bilinear weights of cv::remap on CV_8U (BilinearTab_i).
// The entry for fx = fy = 0 is {32767, 0, 0, 1}: the float weight 1.0
// saturates to SHRT_MAX and the missing unit goes to the last tap.
__device__ __forceinline__ void bilinear_itab(int fx, int fy, int& w0, int& w1, int& w2, int& w3) {
  w0 = (32 - fx) * (32 - fy) * 32;
  w1 = fx * (32 - fy) * 32;
  w2 = (32 - fx) * fy * 32;
  w3 = fx * fy * 32;
  if ((fx | fy) == 0) {
    w0 = 32767;
    w3 = 1;
  }
}

// remap(mask, INTER_LINEAR, BORDER_CONSTANT 0) evaluated at (X, Y); mask may be
// null, meaning an all-255 mask of size w x h.  Exact integer arithmetic:
// (sum + 2^14) >> 15, saturated to [0, 255].
__device__ __forceinline__ int remap_mask_u8(const uint8_t* __restrict__ m, int pitch, int w, int h, int X, int Y) {
  const int ix = sat_short(X >> 5), iy = sat_short(Y >> 5);
  if (ix >= w || ix + 1 < 0 || iy >= h || iy + 1 < 0) return 0;
  int w0, w1, w2, w3;
  bilinear_itab(X & 31, Y & 31, w0, w1, w2, w3);
  const bool bx0 = (unsigned)ix < (unsigned)w, bx1 = (unsigned)(ix + 1) < (unsigned)w;
  const bool by0 = (unsigned)iy < (unsigned)h, by1 = (unsigned)(iy + 1) < (unsigned)h;
  int v00, v01, v10, v11;
  if (m) {
    v00 = (bx0 && by0) ? m[(size_t)iy * pitch + ix] : 0;
    v01 = (bx1 && by0) ? m[(size_t)iy * pitch + ix + 1] : 0;
    v10 = (bx0 && by1) ? m[(size_t)(iy + 1) * pitch + ix] : 0;
    v11 = (bx1 && by1) ? m[(size_t)(iy + 1) * pitch + ix + 1] : 0;
  } else {
    v00 = (bx0 && by0) ? 255 : 0;
    v01 = (bx1 && by0) ? 255 : 0;
    v10 = (bx0 && by1) ? 255 : 0;
    v11 = (bx1 && by1) ? 255 : 0;
  }
  const int s = v00 * w0 + v01 * w1 + v10 * w2 + v11 * w3;
  return min(max((s + (1 << 14)) >> 15, 0), 255);
}

// Warp maps of the image transforms (c_image_transform.cc).  mc[] holds the
// per-frame map coefficients:
//   translation : {t0, t1}                                        (:150-170)
//   euclidean   : {sa, ca, s, Tx, Ty, Cx, Cy}, sa/ca = (float)sin/cos((double)angle)
//                 u = fma(ca,xx,-(sa*yy)); v = fma(sa,xx,ca*yy);
//                 sx = fma(s,u,Tx); sy = fma(s,v,Ty)              (:587-614)
//   affine      : {a00,a01,a02,a10,a11,a12}
//                 sx = fma(a01,y,a00*x)+a02; sy = fma(a11,y,a10*x)+a12  (:926-946)
struct MapCoef {
  float c[8];
};

template <int MOTION>
__device__ __forceinline__ void warp_xy(const MapCoef& m, int x, int y, float& sx, float& sy) {
  const float fx = (float)x, fy = (float)y;
  if constexpr (MOTION == MOTION_TRANSLATION) {
    sx = fx + m.c[0];
    sy = fy + m.c[1];
  } else if constexpr (MOTION == MOTION_AFFINE) {
    sx = fmaf(m.c[1], fy, m.c[0] * fx) + m.c[2];
    sy = fmaf(m.c[4], fy, m.c[3] * fx) + m.c[5];
  } else {
    const float xx = fx - m.c[5], yy = fy - m.c[6];
    const float u = fmaf(m.c[1], xx, -(m.c[0] * yy));
    const float v = fmaf(m.c[0], xx, m.c[1] * yy);
    sx = fmaf(m.c[2], u, m.c[3]);
    sy = fmaf(m.c[2], v, m.c[4]);
  }
}

// ---- reductions -----------------------------------------------------------
__device__ __forceinline__ float wave_sum(float v) {
#pragma unroll
  for (int o = WAVE / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, WAVE);
  return v;
}

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int o = WAVE / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, WAVE);
  return v;
}

// Reduce-scatter of N per-lane values across a wave by recursive halving: at
// step k (mask = 1<<k) each lane keeps the half of its current value set
// selected by its lane bit and adds the partner's copy of that half.  After
// log2(64) steps lane l holds the wave total of value index slot_of(l)
// (or nothing when l >= N).  Cost ~2N shuffles instead of 6N.
template <int N>
struct ReduceScatter {
  static constexpr int n1 = (N + 1) / 2, n2 = (n1 + 1) / 2, n3 = (n2 + 1) / 2, n4 = (n3 + 1) / 2,
                       n5 = (n4 + 1) / 2, n6 = (n5 + 1) / 2;
  static_assert(n6 == 1, "ReduceScatter supports up to 64 values");

  template <int M, int MASK>
  __device__ __forceinline__ static void step(const float* in, float* out, int lane) {
    // value set of size M; lane with bit clear keeps [0, ceil(M/2)), bit set keeps [ceil(M/2), M)
    constexpr int H = (M + 1) / 2;
    const bool hi = (lane & MASK) != 0;
#pragma unroll
    for (int i = 0; i < H; ++i) {
      const float lo_v = in[i];
      const float hi_v = (i + H < M) ? in[i + H] : 0.f;
      const float keep = hi ? hi_v : lo_v;
      const float send = hi ? lo_v : hi_v;
      out[i] = keep + __shfl_xor(send, MASK, WAVE);
    }
  }

  // Slot owned by a lane after the six steps, or -1.  idx accumulates the
  // upper-half offsets; r tracks how many real (non-padding) values the lane's
  // set still holds, so padding slots are never reported.
  __host__ __device__ static constexpr int slot_of(int lane) {
    const int sizes[6] = {N, n1, n2, n3, n4, n5};
    int idx = 0, r = N;
    for (int k = 0; k < 6; ++k) {
      const int h = (sizes[k] + 1) / 2;
      if (lane & (1 << k)) {
        idx += h;
        r = r - h > 0 ? r - h : 0;
      } else {
        r = r < h ? r : h;
      }
    }
    return r >= 1 ? idx : -1;
  }

  // returns the wave total of value slot_of(lane) (meaningless when slot_of(lane) < 0)
  __device__ __forceinline__ static float run(const float* v, int lane) {
    float a[n1], b[n2], c[n3], d[n4], e[n5], f[n6];
    step<N, 1>(v, a, lane);
    step<n1, 2>(a, b, lane);
    step<n2, 4>(b, c, lane);
    step<n3, 8>(c, d, lane);
    step<n4, 16>(d, e, lane);
    step<n5, 32>(e, f, lane);
    return f[0];
  }
};

// The same recursive halving with the masks in wave_sum's order (32, 16, ..., 1) and any
// associative-commutative operation: each step adds the two set-sums wave_sum's butterfly adds
// at that level, so the total of a value is bit for bit wave_sum's (which is the same on every
// lane).  Lane l then holds value slot_of(l) (or nothing, -1).
template <int N>
struct ReduceScatterDesc {
  static constexpr int n1 = (N + 1) / 2, n2 = (n1 + 1) / 2, n3 = (n2 + 1) / 2, n4 = (n3 + 1) / 2,
                       n5 = (n4 + 1) / 2, n6 = (n5 + 1) / 2;
  static_assert(n6 == 1, "ReduceScatterDesc supports up to 64 values");

  template <int M, int MASK, typename T, typename Op>
  __device__ __forceinline__ static void step(const T* in, T* out, int lane, Op op) {
    constexpr int H = (M + 1) / 2;
    const bool hi = (lane & MASK) != 0;
#pragma unroll
    for (int i = 0; i < H; ++i) {
      const T lo_v = in[i];
      const T hi_v = (i + H < M) ? in[i + H] : lo_v;  // a padding slot is never reported
      const T keep = hi ? hi_v : lo_v;
      const T send = hi ? lo_v : hi_v;
      out[i] = op(keep, __shfl_xor(send, MASK, WAVE));
    }
  }

  __host__ __device__ static constexpr int slot_of(int lane) {
    const int sizes[6] = {N, n1, n2, n3, n4, n5};
    int idx = 0, r = N;
    for (int k = 0; k < 6; ++k) {
      const int h = (sizes[k] + 1) / 2;
      if (lane & (32 >> k)) {
        idx += h;
        r = r - h > 0 ? r - h : 0;
      } else {
        r = r < h ? r : h;
      }
    }
    return r >= 1 ? idx : -1;
  }

  template <typename T, typename Op>
  __device__ __forceinline__ static T run(const T* v, int lane, Op op) {
    T a[n1], b[n2], c[n3], d[n4], e[n5], f[n6];
    step<N, 32>(v, a, lane, op);
    step<n1, 16>(a, b, lane, op);
    step<n2, 8>(b, c, lane, op);
    step<n3, 4>(c, d, lane, op);
    step<n4, 2>(d, e, lane, op);
    step<n5, 1>(e, f, lane, op);
    return f[0];
  }
};

}  // namespace ssg

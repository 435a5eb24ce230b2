// ssg_finish.hip -- the finish of a stack on the GPU: inpainting of masked-out pixels
// (core/proc/inpaint/average_pyramid_inpaint.cc, linear_interpolation_inpaint.cc) and the
// weighted blend of unsharp_mask (core/proc/unsharp_mask.cc:72-115).
//
// average_pyramid_inpaint: the recursion's filter + downstrike of a level is one kernel over
// the coarser level's pixels, its upject + filter one kernel over the finer level's pixels;
// the host walks the levels and reads one count per level for the recursion test.
//
// Inpainting, one iteration: every row is scanned by one wave (prefix max /
// suffix min of the valid columns give each hole its bounding valid pixels),
// every column by one thread (the reference's run loop, coalesced across the
// columns of a wave), then each hole takes the distance-weighted mix of the two
// fills.  The host repeats until no hole is filled.
#include "ssg_engine.hpp"

namespace ssg {
namespace {

// the horizontal fill of _interpolate_holes_h2 (linear_interpolation_inpaint.cc:14-117)
__global__ void __launch_bounds__(256) k_inpaint_rows(InpaintArgs a) {
  const int lane = threadIdx.x & 63, y = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (y >= a.h) return;  // whole waves leave together
  const int w = a.w, cn = a.cn;
  const uint8_t* mrow = a.mask + (size_t)y * a.mpitch;
  const float* irow = a.img + (size_t)y * a.pitch;
  int* srow = a.s_tmp + (size_t)y * w;
  // last valid column at or before x
  int carry = -1;
  for (int x0 = 0; x0 < w; x0 += 64) {
    const int x = x0 + lane;
    int v = (x < w && mrow[x]) ? x : -1;
    for (int d = 1; d < 64; d <<= 1) {
      const int t = __shfl_up(v, d, 64);
      if (lane >= d) v = max(v, t);
    }
    v = max(v, carry);
    if (x < w) srow[x] = v;
    carry = __shfl(v, 63, 64);
  }
  // first valid column at or after x, and the fill
  carry = w;
  for (int x0 = ((w - 1) / 64) * 64; x0 >= 0; x0 -= 64) {
    const int x = x0 + lane;
    int v = (x < w && mrow[x]) ? x : w;
    for (int d = 1; d < 64; d <<= 1) {
      const int t = __shfl_down(v, d, 64);
      if (lane + d < 64) v = min(v, t);
    }
    v = min(v, carry);
    carry = __shfl(v, 0, 64);
    if (x < w) {
      float dist = 0.f;
      float* out = a.ih + ((size_t)y * w + x) * cn;
      if (!mrow[x]) {
        const int s = srow[x], e = v;
        if (s >= 0 && e < w) {
          const float scale = 1.0f / (float)(e - s - 1);
          const float factor = (float)(x - s);
          for (int c = 0; c < cn; ++c) {
            const float sv = irow[s * cn + c], ev = irow[e * cn + c];
            const float kk = (ev - sv) * scale;
            out[c] = sv + factor * kk;
          }
          dist = (float)max(x - s, e - x);
        } else if (s >= 0) {
          for (int c = 0; c < cn; ++c) out[c] = irow[s * cn + c];
          dist = (float)(x - s);
        } else if (e < w) {
          for (int c = 0; c < cn; ++c) out[c] = irow[e * cn + c];
          dist = (float)(e - x);
        }
      }
      a.hd[(size_t)y * w + x] = dist;
    }
  }
}

// the vertical fill of _interpolate_holes_v2 (linear_interpolation_inpaint.cc:119-227)
__global__ void __launch_bounds__(256) k_inpaint_cols(InpaintArgs a) {
  const int x = blockIdx.x * 256 + threadIdx.x;
  if (x >= a.w) return;
  const int w = a.w, h = a.h, cn = a.cn;
  auto M = [&](int y) { return a.mask[(size_t)y * a.mpitch + x]; };
  auto I = [&](int y, int c) { return a.img[(size_t)y * a.pitch + x * cn + c]; };
  for (int y = 0; y < h; ++y) a.vd[(size_t)y * w + x] = 0.f;
  for (int start = 0; start < h;) {
    while (start < h && M(start)) ++start;
    if (start >= h) break;
    int end = start + 1;
    while (end < h && !M(end)) ++end;
    if (start > 0 && end < h) {
      const int s = start - 1, e = end;
      const float scale = 1.0f / (float)(end - start);
      for (int y = start; y < e; ++y) {
        const float factor = (float)(y - s);
        float* out = a.iv + ((size_t)y * w + x) * cn;
        a.vd[(size_t)y * w + x] = (float)max(y - s, e - y);
        for (int c = 0; c < cn; ++c) {
          const float sv = I(s, c), ev = I(e, c);
          const float kk = (ev - sv) * scale;
          out[c] = sv + factor * kk;
        }
      }
    } else if (start > 0) {
      const int s = start - 1;
      for (int y = start; y < h; ++y) {
        a.vd[(size_t)y * w + x] = (float)(y - s);
        for (int c = 0; c < cn; ++c) a.iv[((size_t)y * w + x) * cn + c] = I(s, c);
      }
    } else if (end < h) {
      const int e = end;
      for (int y = start; y < e; ++y) {
        a.vd[(size_t)y * w + x] = (float)(e - y);
        for (int c = 0; c < cn; ++c) a.iv[((size_t)y * w + x) * cn + c] = I(e, c);
      }
    }
    start = end;
  }
}

// _fill_holes2 (linear_interpolation_inpaint.cc:229-296)
__global__ void __launch_bounds__(256) k_inpaint_fill(InpaintArgs a) {
  const int x = blockIdx.x * 64 + (threadIdx.x & 63), y = blockIdx.y * 4 + (threadIdx.x >> 6);
  bool filled = false;
  if (x < a.w && y < a.h) {
    uint8_t& m = a.mask[(size_t)y * a.mpitch + x];
    if (!m) {
      const size_t o = (size_t)y * a.w + x;
      const float dh = a.hd[o], dv = a.vd[o];
      float* px = a.img + (size_t)y * a.pitch + x * a.cn;
      const float* hp = a.ih + o * a.cn;
      const float* vp = a.iv + o * a.cn;
      if (dh > 0 && dv > 0) {
        const float dd = 1.0f / (dh + dv);
        for (int c = 0; c < a.cn; ++c) px[c] = dd * (hp[c] * dv + vp[c] * dh);
        filled = true;
      } else if (dv > 0) {
        for (int c = 0; c < a.cn; ++c) px[c] = vp[c];
        filled = true;
      } else if (dh > 0) {
        for (int c = 0; c < a.cn; ++c) px[c] = hp[c];
        filled = true;
      }
      if (filled) m = 255;
    }
  }
  const unsigned long long b = __ballot(filled);
  if ((threadIdx.x & 63) == 0 && b) atomicAdd(a.filled, (unsigned)__popcll(b));
}

// addWeighted(src, alpha, lpass, beta, 0) of unsharp_mask: (a * alpha + b * beta) in float
__global__ void __launch_bounds__(256) k_unsharp_blend(const float* src, int spitch, const float* lp, int lpitch,
                                                       float* dst, int dpitch, int wcn, int h, float alpha,
                                                       float beta) {
  const int x = blockIdx.x * 256 + threadIdx.x, y = blockIdx.y;
  if (x >= wcn || y >= h) return;
  dst[(size_t)y * dpitch + x] = src[(size_t)y * spitch + x] * alpha + lp[(size_t)y * lpitch + x] * beta;
}

// ---- average_pyramid_inpaint (core/proc/inpaint/average_pyramid_inpaint.cc) ----------------------
//
// A level is a dense interleaved f32 image [h][w][cn] and an f32 mask [h][w].  The 3x3 box sums
// of cv::boxFilter(normalize = false, BORDER_REPLICATE) are accumulated in double, as OpenCV's
// box filter does for float data (sum type CV_64F), and rounded once to float.

// the per-pixel rule of _average_pyramid_filter2 (:29-55) with the 3x3 sums b[] and bm
template <int CN>
__device__ __forceinline__ float apr_rule(float fm, const float* fb, const double* b, double bm, float* out) {
  if (fm != 0.f) {
    for (int c = 0; c < CN; ++c) out[c] = fb[c];
    return 1.f;
  }
  const float m = (float)bm;
  if (m != 0.f) {
    const float scale = 1.0f / m;
    for (int c = 0; c < CN; ++c) out[c] = (float)b[c] * scale;
    return 1.f;
  }
  for (int c = 0; c < CN; ++c) out[c] = (float)b[c];
  return m;
}

// level 0: src.copyTo(img, mask), msk = mask * (float)(1 / 255.) (:114-115); counts the mask's
// non-zero pixels
template <int CN>
__global__ void __launch_bounds__(256) k_apr_init(const float* src, int spitch, const uint8_t* mask, int mpitch, int w,
                                                   int h, float* img, float* msk, unsigned* nz) {
  const int x = blockIdx.x * 64 + (threadIdx.x & 63), y = blockIdx.y * 4 + (threadIdx.x >> 6);
  bool valid = false;
  if (x < w && y < h) {
    const uint8_t m = mask[(size_t)y * mpitch + x];
    valid = m != 0;
    const float* sp = src + (size_t)y * spitch + (size_t)x * CN;
    float* dp = img + ((size_t)y * w + x) * CN;
    for (int c = 0; c < CN; ++c) dp[c] = valid ? sp[c] : 0.f;
    msk[(size_t)y * w + x] = (float)m * (float)(1.0 / 255.0);
  }
  const unsigned long long b = __ballot(valid);
  if ((threadIdx.x & 63) == 0 && b) atomicAdd(nz, (unsigned)__popcll(b));
}

// filter2 of level l evaluated at the points downstrike_even keeps, (min(2x+1, w-1),
// min(2y+1, h-1)) (downstrike.cc:34-75), written as level l+1; counts its non-zero mask
template <int CN>
__global__ void __launch_bounds__(256) k_apr_down(const float* img, const float* msk, int w, int h, float* dimg,
                                                   float* dmsk, int dw, int dh, unsigned* nz) {
  const int x = blockIdx.x * 64 + (threadIdx.x & 63), y = blockIdx.y * 4 + (threadIdx.x >> 6);
  bool valid = false;
  if (x < dw && y < dh) {
    const int sx = min(2 * x + 1, w - 1), sy = min(2 * y + 1, h - 1);
    double b[CN];
    double bm = 0;
    for (int c = 0; c < CN; ++c) b[c] = 0;
    const float fm = msk[(size_t)sy * w + sx];
    if (fm == 0.f) {
      for (int dy = -1; dy <= 1; ++dy) {
        const int r = min(max(sy + dy, 0), h - 1);
        for (int dx = -1; dx <= 1; ++dx) {
          const int q = min(max(sx + dx, 0), w - 1);
          const size_t o = (size_t)r * w + q;
          for (int c = 0; c < CN; ++c) b[c] += (double)img[o * CN + c];
          bm += (double)msk[o];
        }
      }
    }
    float out[CN];
    const float om = apr_rule<CN>(fm, img + ((size_t)sy * w + sx) * CN, b, bm, out);
    float* dp = dimg + ((size_t)y * dw + x) * CN;
    for (int c = 0; c < CN; ++c) dp[c] = out[c];
    dmsk[(size_t)y * dw + x] = om;
    valid = om != 0.f;
  }
  const unsigned long long bb = __ballot(valid);
  if ((threadIdx.x & 63) == 0 && bb) atomicAdd(nz, (unsigned)__popcll(bb));
}

// upject_even of level l+1 to the size of level l (values and the zmask of ones at the
// odd-odd positions, downstrike.cc:254-367), then filter2 with level l as the fallback,
// in place on level l
template <int CN>
__global__ void __launch_bounds__(256) k_apr_up(float* img, float* msk, int w, int h, const float* simg, int sw) {
  const int x = blockIdx.x * 64 + (threadIdx.x & 63), y = blockIdx.y * 4 + (threadIdx.x >> 6);
  if (x >= w || y >= h) return;
  double b[CN];
  double bm = 0;
  for (int c = 0; c < CN; ++c) b[c] = 0;
  const size_t o = (size_t)y * w + x;
  const float fm = msk[o];
  if (fm == 0.f) {
    for (int dy = -1; dy <= 1; ++dy) {
      const int r = min(max(y + dy, 0), h - 1);
      if (!(r & 1)) continue;
      for (int dx = -1; dx <= 1; ++dx) {
        const int q = min(max(x + dx, 0), w - 1);
        if (!(q & 1)) continue;
        const float* sp = simg + ((size_t)(r >> 1) * sw + (q >> 1)) * CN;
        for (int c = 0; c < CN; ++c) b[c] += (double)sp[c];
        bm += 1.0;
      }
    }
  }
  float out[CN];
  const float om = apr_rule<CN>(fm, img + o * CN, b, bm, out);
  for (int c = 0; c < CN; ++c) img[o * CN + c] = out[c];
  msk[o] = om;
}

// msk.convertTo(dstmask, CV_8U, 255.0) (:123-125): round to nearest even, saturate
__global__ void __launch_bounds__(256) k_apr_mask_out(const float* msk, int w, int h, uint8_t* dst, int dpitch) {
  const int x = blockIdx.x * 256 + threadIdx.x, y = blockIdx.y;
  if (x >= w || y >= h) return;
  const float v = rintf(msk[(size_t)y * w + x] * 255.f);
  dst[(size_t)y * dpitch + x] = (uint8_t)fminf(fmaxf(v, 0.f), 255.f);
}

}  // namespace

hipError_t launch_apr_init(const float* src, int spitch, const uint8_t* mask, int mpitch, int w, int h, int cn,
                           float* img, float* msk, unsigned* nz, hipStream_t s) {
  const dim3 g((w + 63) / 64, (h + 3) / 4);
  if (cn == 3)
    hipLaunchKernelGGL(k_apr_init<3>, g, dim3(256), 0, s, src, spitch, mask, mpitch, w, h, img, msk, nz);
  else
    hipLaunchKernelGGL(k_apr_init<1>, g, dim3(256), 0, s, src, spitch, mask, mpitch, w, h, img, msk, nz);
  return hipGetLastError();
}

hipError_t launch_apr_down(const float* img, const float* msk, int w, int h, int cn, float* dimg, float* dmsk, int dw,
                           int dh, unsigned* nz, hipStream_t s) {
  const dim3 g((dw + 63) / 64, (dh + 3) / 4);
  if (cn == 3)
    hipLaunchKernelGGL(k_apr_down<3>, g, dim3(256), 0, s, img, msk, w, h, dimg, dmsk, dw, dh, nz);
  else
    hipLaunchKernelGGL(k_apr_down<1>, g, dim3(256), 0, s, img, msk, w, h, dimg, dmsk, dw, dh, nz);
  return hipGetLastError();
}

hipError_t launch_apr_up(float* img, float* msk, int w, int h, int cn, const float* simg, int sw, hipStream_t s) {
  const dim3 g((w + 63) / 64, (h + 3) / 4);
  if (cn == 3)
    hipLaunchKernelGGL(k_apr_up<3>, g, dim3(256), 0, s, img, msk, w, h, simg, sw);
  else
    hipLaunchKernelGGL(k_apr_up<1>, g, dim3(256), 0, s, img, msk, w, h, simg, sw);
  return hipGetLastError();
}

hipError_t launch_apr_mask_out(const float* msk, int w, int h, uint8_t* dst, int dpitch, hipStream_t s) {
  hipLaunchKernelGGL(k_apr_mask_out, dim3((w + 255) / 256, h), dim3(256), 0, s, msk, w, h, dst, dpitch);
  return hipGetLastError();
}

hipError_t launch_inpaint_pass(const InpaintArgs& a, hipStream_t s) {
  hipLaunchKernelGGL(k_inpaint_rows, dim3((a.h + 3) / 4), dim3(256), 0, s, a);
  hipLaunchKernelGGL(k_inpaint_cols, dim3((a.w + 255) / 256), dim3(256), 0, s, a);
  hipLaunchKernelGGL(k_inpaint_fill, dim3((a.w + 63) / 64, (a.h + 3) / 4), dim3(256), 0, s, a);
  return hipGetLastError();
}

hipError_t launch_unsharp_blend(const float* src, int spitch, const float* lp, int lpitch, float* dst, int dpitch,
                                int wcn, int h, float alpha, float beta, hipStream_t s) {
  hipLaunchKernelGGL(k_unsharp_blend, dim3((wcn + 255) / 256, h), dim3(256), 0, s, src, spitch, lp, lpitch, dst,
                     dpitch, wcn, h, alpha, beta);
  return hipGetLastError();
}

}  // namespace ssg

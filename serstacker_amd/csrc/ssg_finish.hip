// ssg_finish.hip -- the finish of a stack on the GPU: linear-interpolation
// inpainting of masked-out pixels (core/proc/inpaint/linear_interpolation_inpaint.cc)
// and the weighted blend of unsharp_mask (core/proc/unsharp_mask.cc:72-115).
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

}  // namespace

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

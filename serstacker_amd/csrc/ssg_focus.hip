// ssg_focus.hip -- focus stacking on gfx950: the lpg sharpness map
// (core/proc/lpg.cc:63-127, 277-344), laplacian pyramids
// (core/proc/laplacian_pyramid.cc:44-127) and the per-level fusion of
// c_laplacian_pyramid_focus_stacking (core/average/c_frame_accumulation.cc:491-681).
//
// Images are planar f32 (one plane per channel, dense rows).  Every elementwise
// formula keeps the reference's order of operations with no fused multiply-add
// (the library builds with -ffp-contract=off), so the results reproduce the
// FP32 oracle bit for bit.
#include "ssg_stencil.hpp"

namespace ssg {

namespace {

__device__ __forceinline__ size_t gidx() { return (size_t)blockIdx.x * 256 + threadIdx.x; }

// interleaved cn-channel rows (pitch in floats) -> cn dense planes
__global__ void __launch_bounds__(256) k_split(const float* src, int pitch, int w, int h, int cn, float* planes) {
  const size_t i = gidx();
  const size_t n = (size_t)w * h;
  if (i >= n) return;
  const int y = (int)(i / w), x = (int)(i - (size_t)y * w);
  for (int c = 0; c < cn; ++c) planes[(size_t)c * n + i] = src[(size_t)y * pitch + (size_t)x * cn + c];
}

__global__ void __launch_bounds__(256) k_merge(const float* planes, int w, int h, int cn, float* dst, int pitch) {
  const size_t i = gidx();
  const size_t n = (size_t)w * h;
  if (i >= n) return;
  const int y = (int)(i / w), x = (int)(i - (size_t)y * w);
  for (int c = 0; c < cn; ++c) dst[(size_t)y * pitch + (size_t)x * cn + c] = planes[(size_t)c * n + i];
}

// reduce_color_channels(REDUCE_AVG): (c0 + c1) + c2 in float, times 1/3 in double
__global__ void __launch_bounds__(256) k_gray(const float* planes, size_t n, int cn, float* g) {
  const size_t i = gidx();
  if (i >= n) return;
  if (cn == 1) {
    g[i] = planes[i];
    return;
  }
  g[i] = gray3(planes[i], planes[n + i], planes[2 * n + i]);
}

// compute_lpg_5x5_parallel (lpg.cc:63-127) at every pixel of the clamped grid, then pow(m, p)
__global__ void __launch_bounds__(256) k_lpg(const float* g, int w, int h, float alpha, float beta, float eps,
                                             float p, float* m) {
  const size_t i = gidx();
  if (i >= (size_t)w * h) return;
  const int y = (int)(i / w), x = (int)(i - (size_t)y * w);
  m[i] = lpg_at(g, w, w, h, x, y, alpha, beta, eps, p);
}

// cv::pyrUp (float) of nplanes planes (pyrup_at, the oracle's canonical form)
__global__ void __launch_bounds__(256) k_pyrup(const float* src, long long src_fstride, int sw, int sh, float* dst,
                                               long long dst_fstride, int dw, int dh) {
  const size_t i = gidx();
  if (i >= (size_t)dw * dh) return;
  const int yo = (int)(i / dw), xo = (int)(i - (size_t)yo * dw);
  const float* S = src + blockIdx.y * src_fstride;
  float* D = dst + blockIdx.y * dst_fstride;
  D[(size_t)yo * dw + xo] = pyrup_at(S, sw, sw, sh, dh, xo, yo);
}

// elementwise helpers over n elements
__global__ void __launch_bounds__(256) k_sub(const float* a, const float* b, float* d, size_t n) {
  const size_t i = gidx();
  if (i < n) d[i] = a[i] - b[i];
}
__global__ void __launch_bounds__(256) k_add(const float* a, const float* b, float* d, size_t n) {
  const size_t i = gidx();
  if (i < n) d[i] = a[i] + b[i];
}
// d[c*n + i] = a[c*n + i] / w[(wc > 1 ? c : 0)*n + i]
__global__ void __launch_bounds__(256) k_div(const float* a, const float* w, int wc, float* d, size_t n, int cn) {
  const size_t i = gidx();
  if (i >= n) return;
  for (int c = 0; c < cn; ++c) d[(size_t)c * n + i] = a[(size_t)c * n + i] / w[(size_t)(wc > 1 ? c : 0) * n + i];
}
// lpg weights: acc[c] = acc[c] + W * L[c]; wsum = wsum + W
__global__ void __launch_bounds__(256) k_wacc(float* acc, float* wsum, const float* W, const float* L, size_t n,
                                              int cn) {
  const size_t i = gidx();
  if (i >= n) return;
  const float w = W[i];
  for (int c = 0; c < cn; ++c) acc[(size_t)c * n + i] = acc[(size_t)c * n + i] + w * L[(size_t)c * n + i];
  wsum[i] = wsum[i] + w;
}
// |x| of cn planes
__global__ void __launch_bounds__(256) k_abs(const float* a, float* d, size_t n) {
  const size_t i = gidx();
  if (i < n) d[i] = fabsf(a[i]);
}
// select_max_energy: acc = w1 > w0 ? pyr : acc
__global__ void __launch_bounds__(256) k_select(float* acc, const float* pyr, const float* w0, const float* w1,
                                                size_t n) {
  const size_t i = gidx();
  if (i < n && w1[i] > w0[i]) acc[i] = pyr[i];
}
// compute_energy, first half: w = lap^2 per channel, or (sum of channels)^2
__global__ void __launch_bounds__(256) k_energy(const float* lap, float* w, size_t n, int cn, int avgc) {
  const size_t i = gidx();
  if (i >= n) return;
  if (!avgc || cn == 1) {
    for (int c = 0; c < cn; ++c) w[(size_t)c * n + i] = lap[(size_t)c * n + i] * lap[(size_t)c * n + i];
  } else {
    const float s = (lap[i] + lap[n + i]) + lap[2 * n + i];
    w[i] = s * s;
  }
}
// lap[c] = lap[c] * w[c or 0]
__global__ void __launch_bounds__(256) k_mulw(float* lap, const float* w, int wc, size_t n, int cn) {
  const size_t i = gidx();
  if (i >= n) return;
  for (int c = 0; c < cn; ++c) lap[(size_t)c * n + i] = lap[(size_t)c * n + i] * w[(size_t)(wc > 1 ? c : 0) * n + i];
}
// addWeighted(a, alpha, b, beta, 0): (a*alpha + b*beta) in float
__global__ void __launch_bounds__(256) k_blend(float* a, const float* b, float alpha, float beta, size_t n) {
  const size_t i = gidx();
  if (i < n) a[i] = a[i] * alpha + b[i] * beta;
}
// meanStdDev of each plane, sequential double sums (small top levels)
__global__ void k_plane_stdev(const float* planes, size_t n, int nplanes, double* out) {
  const int c = threadIdx.x;
  if (c >= nplanes) return;
  double s = 0, sq = 0;
  for (size_t i = 0; i < n; ++i) {
    const double v = planes[(size_t)c * n + i];
    s += v;
    sq += v * v;
  }
  const double scale = 1.0 / (double)n;
  s *= scale;
  const double var = sq * scale - s * s;
  out[c] = sqrt(var > 0 ? var : 0.0);
}

inline dim3 g1(size_t n, int y = 1) { return dim3((unsigned)((n + 255) / 256), y); }

}  // namespace

hipError_t launch_split(const float* src, int pitch, int w, int h, int cn, float* planes, hipStream_t s) {
  hipLaunchKernelGGL(k_split, g1((size_t)w * h), dim3(256), 0, s, src, pitch, w, h, cn, planes);
  return hipGetLastError();
}
hipError_t launch_merge(const float* planes, int w, int h, int cn, float* dst, int pitch, hipStream_t s) {
  hipLaunchKernelGGL(k_merge, g1((size_t)w * h), dim3(256), 0, s, planes, w, h, cn, dst, pitch);
  return hipGetLastError();
}
hipError_t launch_gray(const float* planes, size_t n, int cn, float* g, hipStream_t s) {
  hipLaunchKernelGGL(k_gray, g1(n), dim3(256), 0, s, planes, n, cn, g);
  return hipGetLastError();
}
hipError_t launch_lpg(const float* g, int w, int h, float alpha, float beta, float eps, float p, float* m,
                      hipStream_t s) {
  if (w < 5 || h < 5) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_lpg, g1((size_t)w * h), dim3(256), 0, s, g, w, h, alpha, beta, eps, p, m);
  return hipGetLastError();
}
hipError_t launch_pyrup(const float* src, long long src_fstride, int sw, int sh, float* dst, long long dst_fstride,
                        int dw, int dh, int nplanes, hipStream_t s) {
  hipLaunchKernelGGL(k_pyrup, g1((size_t)dw * dh, nplanes), dim3(256), 0, s, src, src_fstride, sw, sh, dst,
                     dst_fstride, dw, dh);
  return hipGetLastError();
}
// ---- batches of frames: the per-pixel fusion of m frames in frame order, one launch per level ----
// Level planes of a batch are [frame][channel][pixel]; every loop below applies the one-frame
// update of the kernels above frame after frame, so a batch gives the sums of m single adds.

// select_max_energy without smoothing: acc = |L_k| > |acc| ? L_k : acc, frame after frame
__global__ void __launch_bounds__(256) k_select_frames(float* acc, const float* L, size_t ne, size_t fstride, int m,
                                                       int first) {
  const size_t e = gidx();
  if (e >= ne) return;
  float A = acc[e];
  int k = 0;
  if (first) {
    A = L[e];
    k = 1;
  }
  for (; k < m; ++k) {
    const float v = L[(size_t)k * fstride + e];
    if (fabsf(v) > fabsf(A)) A = v;
  }
  acc[e] = A;
}

// weighted_average without smoothing: w = compute_energy(L_k), acc += L_k w, wsum += w
__global__ void __launch_bounds__(256) k_wavg_frames(float* acc, float* wsum, const float* L, size_t n, int cn,
                                                     int avgc, int wc, int m, int first) {
  const size_t i = gidx();
  if (i >= n) return;
  for (int k = 0; k < m; ++k) {
    const float* Lk = L + (size_t)k * cn * n;
    float w[3];
    if (!avgc || cn == 1) {
      for (int c = 0; c < cn; ++c) w[c] = Lk[(size_t)c * n + i] * Lk[(size_t)c * n + i];
    } else {
      const float sm = (Lk[i] + Lk[n + i]) + Lk[2 * n + i];
      w[0] = sm * sm;
    }
    const bool init = first && k == 0;
    for (int c = 0; c < cn; ++c) {
      const float lw = Lk[(size_t)c * n + i] * w[wc > 1 ? c : 0];
      acc[(size_t)c * n + i] = init ? lw : lw + acc[(size_t)c * n + i];
    }
    for (int c = 0; c < wc; ++c) wsum[(size_t)c * n + i] = init ? w[c] : w[c] + wsum[(size_t)c * n + i];
  }
}

// lpg weights: acc += W_k L_k, wsum += W_k (k_wacc per frame)
__global__ void __launch_bounds__(256) k_wacc_frames(float* acc, float* wsum, const float* W, const float* L, size_t n,
                                                     int cn, int m) {
  const size_t i = gidx();
  if (i >= n) return;
  for (int k = 0; k < m; ++k) {
    const float w = W[(size_t)k * n + i];
    const float* Lk = L + (size_t)k * cn * n;
    for (int c = 0; c < cn; ++c) acc[(size_t)c * n + i] = acc[(size_t)c * n + i] + w * Lk[(size_t)c * n + i];
    wsum[i] = wsum[i] + w;
  }
}

// the top level of select / weighted average, frame after frame in one block: the blend weighted
// by the channel-summed standard deviations of the accumulator and the frame (k_plane_stdev and
// k_blend with the weights the host would form)
__global__ void __launch_bounds__(256) k_top_frames(float* acc, const float* L, size_t n, int cn, int m, int first) {
  __shared__ double sd[6];
  __shared__ float ab[2];
  int k = 0;
  if (first) {
    for (size_t e = threadIdx.x; e < (size_t)cn * n; e += 256) acc[e] = L[e];
    __syncthreads();
    k = 1;
  }
  for (; k < m; ++k) {
    const float* Lk = L + (size_t)k * cn * n;
    const int t = threadIdx.x;
    if (t < 2 * cn) {
      const float* pl = t < cn ? acc + (size_t)t * n : Lk + (size_t)(t - cn) * n;
      double sm = 0, sq = 0;
      for (size_t i = 0; i < n; ++i) {
        const double v = pl[i];
        sm += v;
        sq += v * v;
      }
      const double scale = 1.0 / (double)n;
      sm *= scale;
      const double var = sq * scale - sm * sm;
      sd[t] = sqrt(var > 0 ? var : 0.0);
    }
    __syncthreads();
    if (t == 0) {
      double s0 = 0, s1 = 0;
      for (int c = 0; c < cn; ++c) {
        s0 += sd[c];
        s1 += sd[cn + c];
      }
      ab[0] = (float)(s0 / (s0 + s1));
      ab[1] = (float)(s1 / (s0 + s1));
    }
    __syncthreads();
    const float alpha = ab[0], beta = ab[1];
    for (size_t e = t; e < (size_t)cn * n; e += 256) acc[e] = acc[e] * alpha + Lk[e] * beta;
    __syncthreads();
  }
}

hipError_t launch_select_frames(float* acc, const float* L, size_t ne, size_t fstride, int m, bool first,
                                hipStream_t s) {
  hipLaunchKernelGGL(k_select_frames, g1(ne), dim3(256), 0, s, acc, L, ne, fstride, m, first ? 1 : 0);
  return hipGetLastError();
}
hipError_t launch_wavg_frames(float* acc, float* wsum, const float* L, size_t n, int cn, int avgc, int wc, int m,
                              bool first, hipStream_t s) {
  hipLaunchKernelGGL(k_wavg_frames, g1(n), dim3(256), 0, s, acc, wsum, L, n, cn, avgc, wc, m, first ? 1 : 0);
  return hipGetLastError();
}
hipError_t launch_wacc_frames(float* acc, float* wsum, const float* W, const float* L, size_t n, int cn, int m,
                              hipStream_t s) {
  hipLaunchKernelGGL(k_wacc_frames, g1(n), dim3(256), 0, s, acc, wsum, W, L, n, cn, m);
  return hipGetLastError();
}
hipError_t launch_top_frames(float* acc, const float* L, size_t n, int cn, int m, bool first, hipStream_t s) {
  hipLaunchKernelGGL(k_top_frames, dim3(1), dim3(256), 0, s, acc, L, n, cn, m, first ? 1 : 0);
  return hipGetLastError();
}

hipError_t launch_ew(int op, const float* a, const float* b, float* d, size_t n, hipStream_t s) {
  switch (op) {
    case 0: hipLaunchKernelGGL(k_sub, g1(n), dim3(256), 0, s, a, b, d, n); break;
    case 1: hipLaunchKernelGGL(k_add, g1(n), dim3(256), 0, s, a, b, d, n); break;
    case 2: hipLaunchKernelGGL(k_abs, g1(n), dim3(256), 0, s, a, d, n); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}
hipError_t launch_div(const float* a, const float* w, int wc, float* d, size_t n, int cn, hipStream_t s) {
  hipLaunchKernelGGL(k_div, g1(n), dim3(256), 0, s, a, w, wc, d, n, cn);
  return hipGetLastError();
}
hipError_t launch_wacc(float* acc, float* wsum, const float* W, const float* L, size_t n, int cn, hipStream_t s) {
  hipLaunchKernelGGL(k_wacc, g1(n), dim3(256), 0, s, acc, wsum, W, L, n, cn);
  return hipGetLastError();
}
hipError_t launch_select(float* acc, const float* pyr, const float* w0, const float* w1, size_t n, hipStream_t s) {
  hipLaunchKernelGGL(k_select, g1(n), dim3(256), 0, s, acc, pyr, w0, w1, n);
  return hipGetLastError();
}
hipError_t launch_energy(const float* lap, float* w, size_t n, int cn, int avgc, hipStream_t s) {
  hipLaunchKernelGGL(k_energy, g1(n), dim3(256), 0, s, lap, w, n, cn, avgc);
  return hipGetLastError();
}
hipError_t launch_mulw(float* lap, const float* w, int wc, size_t n, int cn, hipStream_t s) {
  hipLaunchKernelGGL(k_mulw, g1(n), dim3(256), 0, s, lap, w, wc, n, cn);
  return hipGetLastError();
}
hipError_t launch_blend(float* a, const float* b, float alpha, float beta, size_t n, hipStream_t s) {
  hipLaunchKernelGGL(k_blend, g1(n), dim3(256), 0, s, a, b, alpha, beta, n);
  return hipGetLastError();
}
hipError_t launch_plane_stdev(const float* planes, size_t n, int nplanes, double* out, hipStream_t s) {
  hipLaunchKernelGGL(k_plane_stdev, dim3(1), dim3(64), 0, s, planes, n, nplanes, out);
  return hipGetLastError();
}

}  // namespace ssg

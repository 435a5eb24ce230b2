// ssg_lvm.hip -- the local-variance sharpness map on gfx950
// (compute_local_variance_map, core/proc/sharpness_measure/
// c_local_variance_sharpness_measure.cc:193-249): morphological gradient of the
// (pyrDown'ed) gray frame, its L1 norm and fourth-power sum, the cubed map plus
// 0.05 Q, and cv::resize INTER_LINEAR back to the frame size.
#include "ssg_engine.hpp"

namespace ssg {

namespace {

__device__ __forceinline__ size_t gid() { return (size_t)blockIdx.x * 256 + threadIdx.x; }

// cv::COLOR_BGR2GRAY on float: b*0.114 + g*0.587 + r*0.299
__global__ void __launch_bounds__(256) k_bgr2gray(const float* src, int pitch, int w, int h, float* g) {
  const size_t i = gid();
  if (i >= (size_t)w * h) return;
  const int y = (int)(i / w), x = (int)(i - (size_t)y * w);
  const float* p = src + (size_t)y * pitch + 3 * x;
  g[i] = (p[0] * 0.114f + p[1] * 0.587f) + p[2] * 0.299f;
}

// dilate - erode with a (2r+1)^2 square, BORDER_REPLICATE
__global__ void __launch_bounds__(256) k_morph_grad(const float* m, int w, int h, int r, float* g) {
  const size_t i = gid();
  if (i >= (size_t)w * h) return;
  const int y = (int)(i / w), x = (int)(i - (size_t)y * w);
  float mx = -3.4e38f, mn = 3.4e38f;
  for (int dy = -r; dy <= r; ++dy) {
    const float* row = m + (size_t)min(max(y + dy, 0), h - 1) * w;
    for (int dx = -r; dx <= r; ++dx) {
      const float v = row[min(max(x + dx, 0), w - 1)];
      mx = fmaxf(mx, v);
      mn = fminf(mn, v);
    }
  }
  g[i] = mx - mn;
}

// per-block partial sums of |G| and G^4 in double
__global__ void __launch_bounds__(256) k_lvm_sums(const float* g, size_t n, double* partial) {
  __shared__ double s1[256], s4[256];
  double a = 0, b = 0;
  for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256) {
    const double v = g[i];
    a += fabs(v);
    const double v2 = v * v;
    b += v2 * v2;
  }
  s1[threadIdx.x] = a;
  s4[threadIdx.x] = b;
  __syncthreads();
  for (int k = 128; k > 0; k >>= 1) {
    if (threadIdx.x < k) {
      s1[threadIdx.x] += s1[threadIdx.x + k];
      s4[threadIdx.x] += s4[threadIdx.x + k];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    partial[2 * blockIdx.x] = s1[0];
    partial[2 * blockIdx.x + 1] = s4[0];
  }
}

// map = G^3 * mapScale + add
__global__ void __launch_bounds__(256) k_lvm_map(const float* g, size_t n, float map_scale, float add, float* m) {
  const size_t i = gid();
  if (i >= n) return;
  const float v = g[i];
  m[i] = ((v * v) * v) * map_scale + add;
}

__global__ void __launch_bounds__(256) k_resize_linear(const float* src, int sw, int sh, float* dst, int dst_pitch,
                                                       int dw, int dh, double scale_x, double scale_y) {
  const size_t i = gid();
  if (i >= (size_t)dw * dh) return;
  const int y = (int)(i / dw), x = (int)(i - (size_t)y * dw);
  int x0, x1, y0, y1;
  float a0, a1, b0, b1;
  lin_tap(x, scale_x, sw, x0, x1, a0, a1);
  lin_tap_v(y, scale_y, sh, y0, y1, b0, b1);
  const float* r0 = src + (size_t)y0 * sw;
  const float* r1 = src + (size_t)y1 * sw;
  const float h0 = r0[x0] * a0 + r0[x1] * a1, h1 = r1[x0] * a0 + r1[x1] * a1;
  dst[(size_t)y * dst_pitch + x] = h0 * b0 + h1 * b1;
}

}  // namespace

namespace {
// an 8/16-bit image as dense f32 (exact)
__global__ void __launch_bounds__(256) k_int_to_f32(const uint8_t* src, int pitch, int w, int h, int u16, float* d) {
  const size_t i = gid();
  if (i >= (size_t)w * h) return;
  const int x = (int)(i % w), y = (int)(i / w);
  const uint8_t* row = src + (size_t)y * pitch;
  d[i] = u16 ? (float)((const uint16_t*)row)[x] : (float)row[x];
}
// cv::pyrDown's integer rounding on 8/16-bit data, (sum + 128) >> 8, applied to the float
// pyramid level (whose values are exact multiples of 1/256 for integer inputs)
__global__ void __launch_bounds__(256) k_round_half_up(float* v, size_t n) {
  const size_t i = gid();
  if (i < n) v[i] = floorf(v[i] + 0.5f);
}
}  // namespace

hipError_t launch_int_to_f32(const void* src, int pitch, int w, int h, bool u16, float* dst, hipStream_t s) {
  hipLaunchKernelGGL(k_int_to_f32, dim3((unsigned)(((size_t)w * h + 255) / 256)), dim3(256), 0, s,
                     (const uint8_t*)src, pitch, w, h, u16 ? 1 : 0, dst);
  return hipGetLastError();
}
hipError_t launch_round_half_up(float* v, size_t n, hipStream_t s) {
  hipLaunchKernelGGL(k_round_half_up, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, v, n);
  return hipGetLastError();
}

hipError_t launch_bgr2gray(const float* src, int pitch, int w, int h, float* g, hipStream_t s) {
  hipLaunchKernelGGL(k_bgr2gray, dim3((unsigned)(((size_t)w * h + 255) / 256)), dim3(256), 0, s, src, pitch, w, h, g);
  return hipGetLastError();
}
hipError_t launch_morph_grad(const float* m, int w, int h, int r, float* g, hipStream_t s) {
  hipLaunchKernelGGL(k_morph_grad, dim3((unsigned)(((size_t)w * h + 255) / 256)), dim3(256), 0, s, m, w, h, r, g);
  return hipGetLastError();
}
int lvm_sum_blocks(size_t n) { return (int)std::min<size_t>(1024, (n + 255) / 256); }
hipError_t launch_lvm_sums(const float* g, size_t n, double* partial, hipStream_t s) {
  hipLaunchKernelGGL(k_lvm_sums, dim3(lvm_sum_blocks(n)), dim3(256), 0, s, g, n, partial);
  return hipGetLastError();
}
hipError_t launch_lvm_map(const float* g, size_t n, float map_scale, float add, float* m, hipStream_t s) {
  hipLaunchKernelGGL(k_lvm_map, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, g, n, map_scale, add, m);
  return hipGetLastError();
}
hipError_t launch_resize_linear(const float* src, int sw, int sh, float* dst, int dst_pitch, int dw, int dh,
                                hipStream_t s) {
  const double sx = 1.0 / ((double)dw / sw), sy = 1.0 / ((double)dh / sh);
  hipLaunchKernelGGL(k_resize_linear, dim3((unsigned)(((size_t)dw * dh + 255) / 256)), dim3(256), 0, s, src, sw, sh,
                     dst, dst_pitch, dw, dh, sx, sy);
  return hipGetLastError();
}

}  // namespace ssg

// ssg_upscale.hip -- upscale of registration remaps for stacking at a larger size
// (c_image_stacking_pipeline.cc:1869-1905, upscale_remap): a batch of interleaved
// 2-channel f32 fields per launch (blockIdx.y = field), one thread per output pixel.
//   x1.5, x3: cv::resize INTER_LINEAR (for float data INTER_LINEAR_EXACT is INTER_LINEAR);
//   x2:       cv::pyrUp.
// The map values are source positions in the frame and are not scaled.
#include "ssg_engine.hpp"

namespace ssg {
namespace {

__global__ void __launch_bounds__(256) k_upscale_linear(const float* const* src, int spitch, int sw, int sh,
                                                        float* const* dst, int dpitch, int dw, int dh, double scale_x,
                                                        double scale_y) {
  const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= (size_t)dw * dh) return;
  const int y = (int)(i / dw), x = (int)(i - (size_t)y * dw);
  const float* S = src[blockIdx.y];
  float* D = dst[blockIdx.y];
  int x0, x1, y0, y1;
  float a0, a1, b0, b1;
  lin_tap(x, scale_x, sw, x0, x1, a0, a1);
  lin_tap(y, scale_y, sh, y0, y1, b0, b1);
  const float* r0 = S + (size_t)y0 * spitch;
  const float* r1 = S + (size_t)y1 * spitch;
  float2 v;
  {
    const float h0 = r0[2 * x0] * a0 + r0[2 * x1] * a1, h1 = r1[2 * x0] * a0 + r1[2 * x1] * a1;
    v.x = h0 * b0 + h1 * b1;
  }
  {
    const float h0 = r0[2 * x0 + 1] * a0 + r0[2 * x1 + 1] * a1, h1 = r1[2 * x0 + 1] * a0 + r1[2 * x1 + 1] * a1;
    v.y = h0 * b0 + h1 * b1;
  }
  *(float2*)(D + (size_t)y * dpitch + 2 * x) = v;
}

// cv::pyrUp to exactly (2 sw, 2 sh): rows -1 and sh come from REFLECT101 above and replication below
__global__ void __launch_bounds__(256) k_upscale_pyrup(const float* const* src, int spitch, int sw, int sh,
                                                       float* const* dst, int dpitch) {
  const int dw = 2 * sw, dh = 2 * sh;
  const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= (size_t)dw * dh) return;
  const int yo = (int)(i / dw), xo = (int)(i - (size_t)yo * dw);
  const float* S = src[blockIdx.y];
  float* D = dst[blockIdx.y];
  const int y = yo >> 1;
  auto srow = [&](int r) {
    if (r < 0) r = sh > 1 ? 1 : 0;
    if (r >= sh) r = sh - 1;
    return S + (size_t)r * spitch;
  };
  const float scale = 1.f / 64.f;
  float out[2];
#pragma unroll
  for (int c = 0; c < 2; ++c) {
    if (yo & 1) {
      out[c] = (hup<2>(srow(y) + c, sw, xo) + hup<2>(srow(y + 1) + c, sw, xo)) * 4.f * scale;
    } else {
      const float q0 = hup<2>(srow(y - 1) + c, sw, xo), q1 = hup<2>(srow(y) + c, sw, xo),
                  q2 = hup<2>(srow(y + 1) + c, sw, xo);
      out[c] = fmaf(q1, 6.f, q0 + q2) * scale;
    }
  }
  *(float2*)(D + (size_t)yo * dpitch + 2 * xo) = make_float2(out[0], out[1]);
}

}  // namespace

void upscale_size(int option, int w, int h, int* dw, int* dh) {
  switch (option) {
    case UPSCALE_X2: *dw = 2 * w; *dh = 2 * h; break;
    case UPSCALE_X15: *dw = w * 3 / 2; *dh = h * 3 / 2; break;
    case UPSCALE_X3: *dw = 3 * w; *dh = 3 * h; break;
    default: *dw = w; *dh = h; break;
  }
}

hipError_t launch_upscale_maps(int option, int n, const float* const* src, int spitch, int sw, int sh,
                               float* const* dst, int dpitch, hipStream_t s) {
  int dw, dh;
  upscale_size(option, sw, sh, &dw, &dh);
  const dim3 g((unsigned)(((size_t)dw * dh + 255) / 256), n);
  if (option == UPSCALE_X2) {
    hipLaunchKernelGGL(k_upscale_pyrup, g, dim3(256), 0, s, src, spitch, sw, sh, dst, dpitch);
  } else {
    // resize(): inv_scale = dsize / ssize, scale = 1 / inv_scale, in double
    const double sx = 1.0 / ((double)dw / sw), sy = 1.0 / ((double)dh / sh);
    hipLaunchKernelGGL(k_upscale_linear, g, dim3(256), 0, s, src, spitch, sw, sh, dst, dpitch, dw, dh, sx, sy);
  }
  return hipGetLastError();
}

}  // namespace ssg

// ssg_upscale.hip -- the upscales of stacking at a larger size (c_image_stacking_pipeline.cc:
// 1869-2000): upscale_remap of registration remaps (interleaved 2-channel f32 fields) and
// upscale_image of frames (1- or 3-channel f32, weight maps included) and their 8-bit masks.
// A batch of same-size planes per launch (blockIdx.z = plane), one thread per output pixel of a
// float plane (64 x 4 per block, so each wave writes one row segment) and per four adjacent
// outputs of a mask.
//   x1.5: cv::resize INTER_LINEAR;  x3: INTER_LINEAR_EXACT (on float data INTER_LINEAR);
//   x2:   cv::pyrUp.
// Float data follows resize.cpp's generic float path: the horizontal pass S[x0] a0 + S[x1] a1,
// then the vertical R[y0] b0 + R[y1] b1 (rows clipped, weights kept).  Masks follow the 8-bit
// paths -- the fixed-point INTER_LINEAR (11-bit weights), the bit-exact INTER_LINEAR_EXACT
// (8-bit weights, exact sums, rounded half up), the integer pyrUp -- and then
// compare(dstmask, 255, CMP_GE).  Remap values (source positions) are not scaled.
#include "ssg_engine.hpp"

namespace ssg {
namespace {

constexpr int UP_BX = 64, UP_BY = 4;

template <int CN>
__device__ __forceinline__ void store_px(float* d, const float (&v)[CN]) {
  if constexpr (CN == 2) {
    *(float2*)d = make_float2(v[0], v[1]);
  } else {
#pragma unroll
    for (int c = 0; c < CN; ++c) d[c] = v[c];
  }
}

template <int CN>
__global__ void __launch_bounds__(256) k_up_linear(const float* const* src, int spitch, int sw, int sh,
                                                   float* const* dst, int dpitch, int dw, int dh, double scale_x,
                                                   double scale_y) {
  const int x = blockIdx.x * UP_BX + threadIdx.x, y = blockIdx.y * UP_BY + threadIdx.y;
  if (x >= dw || y >= dh) return;
  const float* S = src[blockIdx.z];
  int x0, x1, y0, y1;
  float a0, a1, b0, b1;
  lin_tap(x, scale_x, sw, x0, x1, a0, a1);
  lin_tap_v(y, scale_y, sh, y0, y1, b0, b1);
  const float* r0 = S + (size_t)y0 * spitch;
  const float* r1 = S + (size_t)y1 * spitch;
  float v[CN];
#pragma unroll
  for (int c = 0; c < CN; ++c) {
    const float h0 = r0[CN * x0 + c] * a0 + r0[CN * x1 + c] * a1;
    const float h1 = r1[CN * x0 + c] * a0 + r1[CN * x1 + c] * a1;
    v[c] = h0 * b0 + h1 * b1;
  }
  store_px<CN>(dst[blockIdx.z] + (size_t)y * dpitch + CN * x, v);
}

// cv::pyrUp to exactly (2 sw, 2 sh): rows -1 and sh come from REFLECT101 above and replication below
template <int CN>
__global__ void __launch_bounds__(256) k_up_pyrup(const float* const* src, int spitch, int sw, int sh,
                                                  float* const* dst, int dpitch) {
  const int xo = blockIdx.x * UP_BX + threadIdx.x, yo = blockIdx.y * UP_BY + threadIdx.y;
  if (xo >= 2 * sw || yo >= 2 * sh) return;
  const float* S = src[blockIdx.z];
  const int y = yo >> 1;
  auto srow = [&](int r) {
    if (r < 0) r = sh > 1 ? 1 : 0;
    if (r >= sh) r = sh - 1;
    return S + (size_t)r * spitch;
  };
  const float scale = 1.f / 64.f;
  float out[CN];
#pragma unroll
  for (int c = 0; c < CN; ++c) {
    if (yo & 1) {
      out[c] = (hup<CN>(srow(y) + c, sw, xo) + hup<CN>(srow(y + 1) + c, sw, xo)) * 4.f * scale;
    } else {
      const float q0 = hup<CN>(srow(y - 1) + c, sw, xo), q1 = hup<CN>(srow(y) + c, sw, xo),
                  q2 = hup<CN>(srow(y + 1) + c, sw, xo);
      out[c] = fmaf(q1, 6.f, q0 + q2) * scale;
    }
  }
  store_px<CN>(dst[blockIdx.z] + (size_t)yo * dpitch + CN * xo, out);
}

// interpolationLinear::getCoeffs of the bit-exact resize (8-bit fixed-point weights): fval =
// scale (d + 0.5) - 0.5 in double; inside the source c1 = round-half-even(256 frac(fval)),
// c0 = 256 - c1; before the source or from its last sample on, that edge sample alone
__device__ __forceinline__ void exact_tap(int d, double scale, int n, int& s0, int& s1, int& c0, int& c1) {
  const double f = scale * ((double)d + 0.5) - 0.5;
  const double fl = floor(f);
  const int i = (int)fl;
  s0 = 0;
  c1 = 0;
  if (i >= 0 && n > 1) {
    if (i < n - 1) {
      s0 = i;
      c1 = (int)rint((f - fl) * 256.0);
    } else {
      s0 = n - 1;
    }
  }
  c0 = 256 - c1;
  s1 = min(s0 + 1, n - 1);
}

// the integer pyrUp's horizontal pass at output column xo of row R
__device__ __forceinline__ int hup_u8(const uint8_t* R, int sw, int xo) {
  if (sw == 1) return R[0] * 8;
  const int x = xo >> 1;
  if (xo & 1) return x == sw - 1 ? R[x] * 8 : (R[x] + R[x + 1]) * 4;
  if (x == 0) return R[0] * 6 + R[1] * 2;
  if (x == sw - 1) return R[x - 1] + R[x] * 7;
  return R[x - 1] + R[x] * 6 + R[x + 1];
}

// MODE 0: x1.5 INTER_LINEAR (fixed point), 1: x3 INTER_LINEAR_EXACT, 2: x2 pyrUp; then >= 255.
// Each thread writes MX_PT adjacent outputs of one row (the row's taps formed once), as one 4-byte
// store where the row allows it.
constexpr int MX_PT = 4;

template <int MODE>
__global__ void __launch_bounds__(256) k_up_mask(const uint8_t* const* src, int spitch, int sw, int sh,
                                                 uint8_t* const* dst, int dpitch, int dw, int dh, double scale_x,
                                                 double scale_y) {
  const int xb = (blockIdx.x * UP_BX + threadIdx.x) * MX_PT, y = blockIdx.y * UP_BY + threadIdx.y;
  if (xb >= dw || y >= dh) return;
  const uint8_t* S = src[blockIdx.z];
  uint8_t* D = dst[blockIdx.z] + (size_t)y * dpitch;
  uint32_t packed = 0;
  int yt0, yt1, wy0 = 0, wy1 = 0;  // the row's taps and weights
  if constexpr (MODE == 0) {
    float b0, b1;
    lin_tap_v(y, scale_y, sh, yt0, yt1, b0, b1);
    // saturate_cast<short>(w * INTER_RESIZE_COEF_SCALE): round half to even
    wy0 = __float2int_rn(b0 * 2048.f);
    wy1 = __float2int_rn(b1 * 2048.f);
  } else if constexpr (MODE == 1) {
    exact_tap(y, scale_y, sh, yt0, yt1, wy0, wy1);
  } else {
    yt0 = yt1 = 0;
  }
  const int n = min(MX_PT, dw - xb);
#pragma unroll
  for (int k = 0; k < MX_PT; ++k) {
    if (k >= n) break;
    const int x = xb + k;
    int v;
    if constexpr (MODE == 0) {
      int x0, x1;
      float a0, a1;
      lin_tap(x, scale_x, sw, x0, x1, a0, a1);
      const int ia0 = __float2int_rn(a0 * 2048.f), ia1 = __float2int_rn(a1 * 2048.f);
      const uint8_t* r0 = S + (size_t)yt0 * spitch;
      const uint8_t* r1 = S + (size_t)yt1 * spitch;
      const int h0 = r0[x0] * ia0 + r0[x1] * ia1, h1 = r1[x0] * ia0 + r1[x1] * ia1;
      v = (((wy0 * (h0 >> 4)) >> 16) + ((wy1 * (h1 >> 4)) >> 16) + 2) >> 2;
    } else if constexpr (MODE == 1) {
      int x0, x1, c0, c1;
      exact_tap(x, scale_x, sw, x0, x1, c0, c1);
      const uint8_t* r0 = S + (size_t)yt0 * spitch;
      const uint8_t* r1 = S + (size_t)yt1 * spitch;
      const int h0 = r0[x0] * c0 + r0[x1] * c1, h1 = r1[x0] * c0 + r1[x1] * c1;
      v = (h0 * wy0 + h1 * wy1 + 32768) >> 16;
    } else {
      auto srow = [&](int r) {
        if (r < 0) r = sh > 1 ? 1 : 0;
        if (r >= sh) r = sh - 1;
        return S + (size_t)r * spitch;
      };
      const int yy = y >> 1;
      if (y & 1)
        v = ((hup_u8(srow(yy), sw, x) + hup_u8(srow(yy + 1), sw, x)) * 4 + 32) >> 6;
      else
        v = (hup_u8(srow(yy - 1), sw, x) + hup_u8(srow(yy), sw, x) * 6 + hup_u8(srow(yy + 1), sw, x) + 32) >> 6;
    }
    packed |= (v >= 255 ? 255u : 0u) << (8 * k);
  }
  if (n == MX_PT && (((uintptr_t)(D + xb)) & 3u) == 0) {
    *(uint32_t*)(D + xb) = packed;
  } else {
    for (int k = 0; k < n; ++k) D[xb + k] = (uint8_t)(packed >> (8 * k));
  }
}

// resize(): inv_scale = dsize / ssize, scale = 1 / inv_scale, both in double
inline double resize_scale(int d, int s) { return 1.0 / ((double)d / s); }

}  // namespace

void upscale_size(int option, int w, int h, int* dw, int* dh) {
  switch (option) {
    case UPSCALE_X2: *dw = 2 * w; *dh = 2 * h; break;
    case UPSCALE_X15: *dw = w * 3 / 2; *dh = h * 3 / 2; break;
    case UPSCALE_X3: *dw = 3 * w; *dh = 3 * h; break;
    default: *dw = w; *dh = h; break;
  }
}

hipError_t launch_upscale_f32(int option, int cn, int n, const float* const* src, int spitch, int sw, int sh,
                              float* const* dst, int dpitch, hipStream_t s) {
  int dw, dh;
  upscale_size(option, sw, sh, &dw, &dh);
  const dim3 g((unsigned)((dw + UP_BX - 1) / UP_BX), (unsigned)((dh + UP_BY - 1) / UP_BY), (unsigned)n),
      b(UP_BX, UP_BY);
  if (option == UPSCALE_X2) {
    switch (cn) {
      case 1: hipLaunchKernelGGL(k_up_pyrup<1>, g, b, 0, s, src, spitch, sw, sh, dst, dpitch); break;
      case 2: hipLaunchKernelGGL(k_up_pyrup<2>, g, b, 0, s, src, spitch, sw, sh, dst, dpitch); break;
      case 3: hipLaunchKernelGGL(k_up_pyrup<3>, g, b, 0, s, src, spitch, sw, sh, dst, dpitch); break;
      default: return hipErrorInvalidValue;
    }
  } else {
    const double sx = resize_scale(dw, sw), sy = resize_scale(dh, sh);
    switch (cn) {
      case 1: hipLaunchKernelGGL(k_up_linear<1>, g, b, 0, s, src, spitch, sw, sh, dst, dpitch, dw, dh, sx, sy); break;
      case 2: hipLaunchKernelGGL(k_up_linear<2>, g, b, 0, s, src, spitch, sw, sh, dst, dpitch, dw, dh, sx, sy); break;
      case 3: hipLaunchKernelGGL(k_up_linear<3>, g, b, 0, s, src, spitch, sw, sh, dst, dpitch, dw, dh, sx, sy); break;
      default: return hipErrorInvalidValue;
    }
  }
  return hipGetLastError();
}

hipError_t launch_upscale_maps(int option, int n, const float* const* src, int spitch, int sw, int sh,
                               float* const* dst, int dpitch, hipStream_t s) {
  return launch_upscale_f32(option, 2, n, src, spitch, sw, sh, dst, dpitch, s);
}

hipError_t launch_upscale_masks(int option, int n, const uint8_t* const* src, int spitch, int sw, int sh,
                                uint8_t* const* dst, int dpitch, hipStream_t s) {
  int dw, dh;
  upscale_size(option, sw, sh, &dw, &dh);
  const dim3 g((unsigned)((dw + UP_BX * MX_PT - 1) / (UP_BX * MX_PT)), (unsigned)((dh + UP_BY - 1) / UP_BY),
               (unsigned)n),
      b(UP_BX, UP_BY);
  const double sx = resize_scale(dw, sw), sy = resize_scale(dh, sh);
  switch (option) {
    case UPSCALE_X15:
      hipLaunchKernelGGL(k_up_mask<0>, g, b, 0, s, src, spitch, sw, sh, dst, dpitch, dw, dh, sx, sy);
      break;
    case UPSCALE_X3:
      hipLaunchKernelGGL(k_up_mask<1>, g, b, 0, s, src, spitch, sw, sh, dst, dpitch, dw, dh, sx, sy);
      break;
    case UPSCALE_X2:
      hipLaunchKernelGGL(k_up_mask<2>, g, b, 0, s, src, spitch, sw, sh, dst, dpitch, dw, dh, sx, sy);
      break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

}  // namespace ssg

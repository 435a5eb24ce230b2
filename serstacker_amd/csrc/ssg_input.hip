// ssg_input.hip -- the input stage of the stacking pipeline on the GPU: raw
// frames (u8 / u16 / f32) to the pipeline's f32 images, fused with the
// darkbayer subtraction, flatbayer division and debayer of
// c_image_stacking_pipeline_base::read_input_frame
// (c_image_stacking_pipeline_base.cc:124-279), and the debayer algorithms of
// core/io/debayer.cc on their own (NN2, SP, MATRIX, PLANE_k, AVGBP, and OpenCV's bilinear and
// edge-aware demosaicing behind DEBAYER_NN / DEBAYER_EA on 8/16-bit data).
//
// One thread per output pixel; a launch covers a batch of frames (grid.z).
// The reference's arithmetic is kept exactly: integer sources are averaged in
// int with the rounding constants c1 = 1, c2 = 2 (AVGBP: 2), float sources in
// float in the reference's operand order, BORDER_REFLECT_101 at the edges.
#include <type_traits>

#include "ssg_engine.hpp"

namespace ssg {
namespace {

// colour of the Bayer site (x, y): 0 = B, 1 = G, 2 = R (debayer.cc:828-1160);
// the CMY patterns use the layout of their RGB counterpart (MYYC = RGGB, ...)
__device__ __forceinline__ int site_color(int pat, int x, int y) {
  const int q = ((y & 1) << 1) | (x & 1);
  switch (pat) {
    case 0: return q == 0 ? 2 : (q == 3 ? 0 : 1);  // RGGB
    case 1: return q == 1 ? 2 : (q == 2 ? 0 : 1);  // GRBG
    case 2: return q == 2 ? 2 : (q == 1 ? 0 : 1);  // GBRG
    default: return q == 3 ? 2 : (q == 0 ? 0 : 1); // BGGR
  }
}

__device__ __forceinline__ int refl101(int i, int n) { return i < 0 ? -i : (i >= n ? 2 * n - 2 - i : i); }

template <typename TI>
__device__ __forceinline__ TI ld_raw(const InputArgs& a, int f, int x, int y) {
  const uint8_t* row = (const uint8_t*)a.src + (size_t)f * a.src_fstride + (size_t)y * a.src_pitch;
  return ((const TI*)row)[x];
}

// sample in the working type: the integer sample, or the pipeline's float sample
// ((float)raw * 2^-bpp for integer raw, then - dark, then / flat)
template <typename TI, bool FW>
struct Sampler {
  using W = typename std::conditional<FW, float, int>::type;
  __device__ __forceinline__ static W at(const InputArgs& a, int f, int x, int y) {
    if constexpr (FW) {
      float v = (float)ld_raw<TI>(a, f, x, y);
      if (a.convert) v = v * a.scale;
      if (a.dark) v = v - a.dark[(size_t)y * a.dark_pitch + x];
      if (a.flat) v = v / a.flat[(size_t)y * a.flat_pitch + x];
      return v;
    } else {
      return (int)ld_raw<TI>(a, f, x, y);
    }
  }
};

// the reference's averages: (c + a + b) / 2 and (c + a + b + d + e) / 4, left to right
template <typename W>
__device__ __forceinline__ W avg2(W p, W q) {
  if constexpr (std::is_integral<W>::value) return (1 + p + q) / 2;
  else return (0.f + p + q) / 2;
}
template <typename W>
__device__ __forceinline__ W avg4(W p, W q, W r, W s) {
  if constexpr (std::is_integral<W>::value) return (2 + p + q + r + s) / 4;
  else return (0.f + p + q + r + s) / 4;
}

// store of one output value: native type (saturate_cast), or f32 (scaled for integer
// sources converted after the debayer: convertTo(CV_32F, 1 / 2^bpp))
template <typename TO, typename W>
__device__ __forceinline__ TO out_value(const InputArgs& a, W v) {
  if constexpr (std::is_same<TO, float>::value) {
    if constexpr (std::is_same<W, float>::value) return v;
    else return a.convert ? (float)v * a.scale : (float)v;
  } else {
    if constexpr (std::is_same<W, float>::value) {
      const float r = rintf(v);
      return (TO)fminf(fmaxf(r, 0.f), (float)((1 << (8 * sizeof(TO))) - 1));
    } else {
      return (TO)min(max(v, 0), (1 << (8 * sizeof(TO))) - 1);
    }
  }
}

template <typename TO>
__device__ __forceinline__ TO* out_row(const InputArgs& a, int f, int y) {
  return (TO*)((uint8_t*)a.dst + (size_t)f * a.dst_fstride + (size_t)y * a.dst_pitch);
}

template <typename TI, typename TO, bool FW>
__global__ void __launch_bounds__(256) k_input(InputArgs a) {
  using S = Sampler<TI, FW>;
  using W = typename S::W;
  const int x = blockIdx.x * 64 + (threadIdx.x & 63), y = blockIdx.y * 4 + (threadIdx.x >> 6), f = blockIdx.z;
  if (x >= a.ow || y >= a.oh) return;
  const int w = a.w, h = a.h;
  TO* o = out_row<TO>(a, f, y);
  switch (a.method) {
    case INPUT_COPY: {
      o[x] = out_value<TO, W>(a, S::at(a, f, x, y));
      return;
    }
    case INPUT_NN2: {
      // _debayer_nn2_interpolation (debayer.cc:828-1160)
      const int xl = refl101(x - 1, w), xr = refl101(x + 1, w), yu = refl101(y - 1, h), yd = refl101(y + 1, h);
      const int c = site_color(a.pattern, x, y);
      W bgr[3];
      bgr[c] = S::at(a, f, x, y);
      if (c == 1) {
        // green site: the horizontal neighbours' colour from left + right, the vertical one from up + down
        const int ch = site_color(a.pattern, xl, y), cv = site_color(a.pattern, x, yu);
        bgr[ch] = avg2<W>(S::at(a, f, xl, y), S::at(a, f, xr, y));
        bgr[cv] = avg2<W>(S::at(a, f, x, yu), S::at(a, f, x, yd));
      } else {
        bgr[1] = avg4<W>(S::at(a, f, x, yu), S::at(a, f, xl, y), S::at(a, f, xr, y), S::at(a, f, x, yd));
        bgr[2 - c] = avg4<W>(S::at(a, f, xl, yu), S::at(a, f, xr, yu), S::at(a, f, xl, yd), S::at(a, f, xr, yd));
      }
      for (int k = 0; k < 3; ++k) o[3 * x + k] = out_value<TO, W>(a, bgr[k]);
      return;
    }
    case INPUT_MATRIX: {
      // _debayer_matrix (debayer.cc:455-575): the site's value in its own channel
      const int c = site_color(a.pattern, x, y);
      const W v = S::at(a, f, x, y);
      for (int k = 0; k < 3; ++k) o[3 * x + k] = out_value<TO, W>(a, k == c ? v : (W)0);
      return;
    }
    case INPUT_SP: {
      // _debayer_sp (debayer.cc:587-700): one BGR pixel per 2x2 cell, G = (c1 + g + g) / 2
      const int X = 2 * x, Y = 2 * y;
      W bgr[3];
      W g[2];
      int ng = 0;
      for (int q = 0; q < 4; ++q) {
        const int xx = X + (q & 1), yy = Y + (q >> 1);
        const int c = site_color(a.pattern, xx, yy);
        const W v = S::at(a, f, xx, yy);
        if (c == 1) g[ng++] = v;
        else bgr[c] = v;
      }
      bgr[1] = avg2<W>(g[0], g[1]);
      for (int k = 0; k < 3; ++k) o[3 * x + k] = out_value<TO, W>(a, bgr[k]);
      return;
    }
    case INPUT_AVGBP: {
      // _average_bayer_planes (debayer.cc:277-369): (c1 + s00 + s01 + s10 + s11) / 4
      const int X = 2 * x, Y = 2 * y;
      o[x] = out_value<TO, W>(a, avg4<W>(S::at(a, f, X, Y), S::at(a, f, X + 1, Y), S::at(a, f, X, Y + 1),
                                         S::at(a, f, X + 1, Y + 1)));
      return;
    }
    case INPUT_CV_BILINEAR:
    case INPUT_CV_EA: {
      // cv::demosaicing COLOR_Bayer*2BGR / COLOR_Bayer*2BGR_EA (debayer.cc:1272-1355; OpenCV's
      // Bayer2RGB_ and Bayer2RGB_EdgeAware_T): the interior from the 3x3 neighbourhood in int;
      // columns 0 and w-1 copy columns 1 and w-2, then rows 0 and h-1 copy rows 1 and h-2, so a
      // border pixel is its clamped interior pixel; images of 2 rows or columns or fewer are 0
      if (w <= 2 || h <= 2) {
        for (int k = 0; k < 3; ++k) o[3 * x + k] = out_value<TO, W>(a, (W)0);
        return;
      }
      const int xi = min(max(x, 1), w - 2), yi = min(max(y, 1), h - 2);
      const int c = site_color(a.pattern, xi, yi);
      const W L = S::at(a, f, xi - 1, yi), R = S::at(a, f, xi + 1, yi), U = S::at(a, f, xi, yi - 1),
              D = S::at(a, f, xi, yi + 1);
      W bgr[3];
      bgr[c] = S::at(a, f, xi, yi);
      if (c == 1) {
        bgr[site_color(a.pattern, xi - 1, yi)] = avg2<W>(L, R);
        bgr[site_color(a.pattern, xi, yi - 1)] = avg2<W>(U, D);
      } else {
        if (a.method == INPUT_CV_EA) {
          const W dh = L > R ? L - R : R - L, dv = D > U ? D - U : U - D;
          bgr[1] = dh > dv ? avg2<W>(U, D) : avg2<W>(L, R);
        } else {
          bgr[1] = avg4<W>(U, L, R, D);
        }
        bgr[2 - c] = avg4<W>(S::at(a, f, xi - 1, yi - 1), S::at(a, f, xi + 1, yi - 1), S::at(a, f, xi - 1, yi + 1),
                             S::at(a, f, xi + 1, yi + 1));
      }
      for (int k = 0; k < 3; ++k) o[3 * x + k] = out_value<TO, W>(a, bgr[k]);
      return;
    }
    default: {
      // _extract_bayer_plane (debayer.cc:157-215): plane k = site (k & 1, k >> 1) of each cell
      const int k = a.method - INPUT_PLANE0;
      o[x] = out_value<TO, W>(a, S::at(a, f, 2 * x + (k & 1), 2 * y + ((k >> 1) & 1)));
      return;
    }
  }
}

template <typename TI, typename TO, bool FW>
hipError_t launch_k(const InputArgs& a, int nframes, hipStream_t s) {
  dim3 grid((a.ow + 63) / 64, (a.oh + 3) / 4, nframes);
  hipLaunchKernelGGL((k_input<TI, TO, FW>), grid, dim3(256), 0, s, a);
  return hipGetLastError();
}

}  // namespace

// in_type: SSG_8UC1 / SSG_16UC1 / SSG_32FC1; out: native (same depth) or f32
hipError_t launch_input(const InputArgs& a, int in_depth, bool out_f32, bool float_work, int nframes, hipStream_t s) {
  if (nframes <= 0 || a.ow <= 0 || a.oh <= 0) return hipErrorInvalidValue;
  if (in_depth == 0) {  // u8
    if (out_f32) return float_work ? launch_k<uint8_t, float, true>(a, nframes, s) : launch_k<uint8_t, float, false>(a, nframes, s);
    return launch_k<uint8_t, uint8_t, false>(a, nframes, s);
  }
  if (in_depth == 2) {  // u16
    if (out_f32) return float_work ? launch_k<uint16_t, float, true>(a, nframes, s) : launch_k<uint16_t, float, false>(a, nframes, s);
    return launch_k<uint16_t, uint16_t, false>(a, nframes, s);
  }
  if (in_depth == 5) return launch_k<float, float, true>(a, nframes, s);  // f32
  return hipErrorInvalidValue;
}

}  // namespace ssg

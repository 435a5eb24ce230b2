// ssg_pyr.hip -- image pyramid kernels for gfx950: separable smoothing
// (cv::sepFilter2D), pyrDown, mask preparation, small reductions.
//
// The tiled kernels load a tile plus halo into LDS once with coalesced reads,
// filter rows into a second LDS buffer and write the column-filtered output
// with coalesced stores. The registration pyramid is built by column strips
// instead (k_strip_pyrdown): one wave per strip, neighbours by DPP, windows in
// registers, level 0 smoothed and level 1 formed in the same pass.
#include <type_traits>

#include "ssg_stencil.hpp"

namespace ssg {

// ---------------------------------------------------------------------------
// Separable filter, canonical order of ssg_oracle.c:ora_sep_filter
// (the OpenCV split SymmRowSmallFilter / RowFilter / SymmColumnFilter,
// filter.simd.hpp, with fused multiply-adds).
// ---------------------------------------------------------------------------
constexpr int SF_TW = 64, SF_TH = 16, SF_MAXR = (SEP_MAX_TAPS - 1) / 2;

__global__ void __launch_bounds__(256) k_sepfilter(const float* const* __restrict__ src_ptrs, int src_pitch,
                                                   float* __restrict__ dst, long long dst_fstride, int dst_pitch,
                                                   int w, int h, SepKernel K, int border) {
  __shared__ float sm[SF_TH + 2 * SF_MAXR][SF_TW + 2 * SF_MAXR];
  __shared__ float rb[SF_TH + 2 * SF_MAXR][SF_TW];
  const int f = blockIdx.z;
  const float* __restrict__ S = src_ptrs[f];
  float* __restrict__ D = dst + f * dst_fstride;
  const int rx = K.nx / 2, ry = K.ny / 2;
  const int x0 = blockIdx.x * SF_TW, y0 = blockIdx.y * SF_TH;
  const int TW = SF_TW + 2 * rx, TH = SF_TH + 2 * ry;
  for (int i = threadIdx.x; i < TW * TH; i += blockDim.x) {
    const int ly = i / TW, lx = i - ly * TW;
    const int bx = border_interp(x0 - rx + lx, w, border), by = border_interp(y0 - ry + ly, h, border);
    sm[ly][lx] = (bx < 0 || by < 0) ? 0.f : S[(size_t)by * src_pitch + bx];
  }
  __syncthreads();
  for (int i = threadIdx.x; i < TH * SF_TW; i += blockDim.x) {
    const int ly = i / SF_TW, lx = i - ly * SF_TW;
    rb[ly][lx] = row_filter(&sm[ly][lx + rx], K);
  }
  __syncthreads();
  const int lx = threadIdx.x & (SF_TW - 1);
  for (int ly = threadIdx.x / SF_TW; ly < SF_TH; ly += blockDim.x / SF_TW) {
    const int gx = x0 + lx, gy = y0 + ly;
    if (gx < w && gy < h) D[(size_t)gy * dst_pitch + gx] = col_filter(&rb[ly + ry][lx], SF_TW, K);
  }
}

// Same filter on 64x64 output tiles for square kernels of radius R <= 7: the
// source tile is gathered in batches of 4 loads per thread, the column pass
// slides a (2R+1)-row register window down a 16-row strip per thread.
constexpr int ST_TW = 64, ST_TH = 64;

template <int R>
__device__ __forceinline__ float row_filter_r(const float* c, const SepKernel& K) {
  constexpr int n = 2 * R + 1;
  const float* k = K.kx;
  if (K.tx == 1 && n <= 5) {
    float s = c[0] * k[R];
    if (n >= 3) s = fmaf(c[-1] + c[1], k[R + 1], s);
    if (n >= 5) s = fmaf(c[-2] + c[2], k[R + 2], s);
    return s;
  }
  if (K.tx == 2 && n >= 3 && n <= 5) {
    float s = (c[1] - c[-1]) * k[R + 1];
    if (n >= 5) s = fmaf(c[2] - c[-2], k[R + 2], s);
    return s;
  }
  float s = 0.f;
#pragma unroll
  for (int i = 0; i < n; ++i) s = fmaf(c[i - R], k[i], s);
  return s;
}

template <int R>
__device__ __forceinline__ float col_filter_r(const float* win, const SepKernel& K) {
  // win[0..2R]: rows -R..R
  constexpr int n = 2 * R + 1;
  const float* k = K.ky;
  if (K.ty == 1) {
    float s = fmaf(win[R], k[R], K.delta);
#pragma unroll
    for (int i = 1; i <= R; ++i) s = fmaf(win[R - i] + win[R + i], k[R + i], s);
    return s;
  }
  if (K.ty == 2) {
    float s = K.delta;
#pragma unroll
    for (int i = 1; i <= R; ++i) s = fmaf(win[R + i] - win[R - i], k[R + i], s);
    return s;
  }
  float s = K.delta;
#pragma unroll
  for (int i = 0; i < n; ++i) s = fmaf(win[i], k[i], s);
  return s;
}

template <int R>
__global__ void __launch_bounds__(256) k_sepfilter_t(const float* const* __restrict__ src_ptrs, int src_pitch,
                                                     float* __restrict__ dst, long long dst_fstride, int dst_pitch,
                                                     int w, int h, SepKernel K, int border) {
  constexpr int TW = ST_TW + 2 * R, TH = ST_TH + 2 * R;
  __shared__ float sm[TH][TW];
  __shared__ float rb[TH][ST_TW];
  const int f = blockIdx.z, tid = threadIdx.x;
  const float* __restrict__ S = src_ptrs[f];
  float* __restrict__ D = dst + f * dst_fstride;
  const int x0 = blockIdx.x * ST_TW, y0 = blockIdx.y * ST_TH;
  {
    // tiles whose halo lies inside the image skip the border mapping (the same samples)
    constexpr int NS = TW * TH, PER = (NS + 255) / 256, FB = 8;
    float* sf = &sm[0][0];
    auto load_tile = [&](auto inside_c) {
      constexpr bool INSIDE = decltype(inside_c)::value;
#pragma nounroll
      for (int b0 = 0; b0 < PER; b0 += FB) {
        int off[FB], dsti[FB];
#pragma unroll
        for (int j = 0; j < FB; ++j) {
          const int i0 = tid + (b0 + j) * 256;
          const bool ok = b0 + j < PER && i0 < NS;
          const int i = ok ? i0 : 0;
          const int ly = i / TW, lx = i - ly * TW;
          if (INSIDE) {
            off[j] = (y0 - R + ly) * src_pitch + (x0 - R + lx);
          } else {
            const int bx = border_interp(x0 - R + lx, w, border), by = border_interp(y0 - R + ly, h, border);
            off[j] = (bx < 0 || by < 0) ? -1 : by * src_pitch + bx;
          }
          dsti[j] = ok ? i : -1;
        }
        float v[FB];
#pragma unroll
        for (int j = 0; j < FB; ++j) v[j] = (INSIDE || off[j] >= 0) ? S[off[j]] : 0.f;
#pragma unroll
        for (int j = 0; j < FB; ++j)
          if (dsti[j] >= 0) sf[dsti[j]] = v[j];
      }
    };
    if (x0 - R >= 0 && y0 - R >= 0 && x0 + ST_TW + R <= w && y0 + ST_TH + R <= h)
      load_tile(std::true_type{});
    else
      load_tile(std::false_type{});
  }
  __syncthreads();
  for (int i = tid; i < TH * ST_TW; i += 256) {
    const int ly = i / ST_TW, lx = i - ly * ST_TW;
    rb[ly][lx] = row_filter_r<R>(&sm[ly][lx + R], K);
  }
  __syncthreads();
  const int lx = tid & 63, r0 = (tid >> 6) * 16, gx = x0 + lx;
  float win[2 * R + 1];
#pragma unroll
  for (int i = 0; i < 2 * R; ++i) win[i + 1] = rb[r0 + i][lx];
#pragma unroll
  for (int dy = 0; dy < 16; ++dy) {
#pragma unroll
    for (int i = 0; i < 2 * R; ++i) win[i] = win[i + 1];
    win[2 * R] = rb[r0 + dy + 2 * R][lx];
    const int gy = y0 + r0 + dy;
    if (gx < w && gy < h) D[(size_t)gy * dst_pitch + gx] = col_filter_r<R>(win, K);
  }
}

hipError_t launch_sepfilter(const float* const* src_ptrs, int src_pitch, float* dst, long long dst_fstride,
                            int dst_pitch, int w, int h, int nframes, const SepKernel& k, int border,
                            hipStream_t s) {
  if (k.nx > 2 * SF_MAXR + 1 || k.ny > 2 * SF_MAXR + 1 || nframes <= 0) return hipErrorInvalidValue;
  if (k.nx == k.ny && (k.nx & 1) && k.nx <= 15) {
    dim3 grid((w + ST_TW - 1) / ST_TW, (h + ST_TH - 1) / ST_TH, nframes);
#define SSG_SF(RR) hipLaunchKernelGGL((k_sepfilter_t<RR>), grid, dim3(256), 0, s, src_ptrs, src_pitch, dst, dst_fstride, dst_pitch, w, h, k, border)
    switch (k.nx / 2) {
      case 1: SSG_SF(1); break;
      case 2: SSG_SF(2); break;
      case 3: SSG_SF(3); break;
      case 4: SSG_SF(4); break;
      case 5: SSG_SF(5); break;
      case 6: SSG_SF(6); break;
      default: SSG_SF(7); break;
    }
#undef SSG_SF
    return hipGetLastError();
  }
  dim3 grid((w + SF_TW - 1) / SF_TW, (h + SF_TH - 1) / SF_TH, nframes);
  hipLaunchKernelGGL(k_sepfilter, grid, dim3(256), 0, s, src_ptrs, src_pitch, dst, dst_fstride, dst_pitch, w, h, k,
                     border);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// pyrDown (pyramids.cpp pyrDown_): [1 4 6 4 1]/16 at even source pixels.
// Canonical: r = fma(s-1 + s1, 4, fma(s0, 6, s-2 + s2)); d = fma(r1 + r3, 4,
// fma(r2, 6, r0 + r4)) * (1/256), the order of ssg_oracle.c:ora_pyr_down.
// ---------------------------------------------------------------------------
constexpr int PD_TW = 64, PD_TH = 32;
constexpr int PD_SW = 2 * PD_TW + 4, PD_SH = 2 * PD_TH + 4;

__global__ void __launch_bounds__(256) k_pyrdown(const float* __restrict__ src, long long src_fstride, int src_pitch,
                                                 int sw, int sh, float* __restrict__ dst, long long dst_fstride,
                                                 int dst_pitch, int dw, int dh, int border) {
  __shared__ float sm[PD_SH][PD_SW + 1];
  __shared__ float rb[PD_SH][PD_TW];
  const int f = blockIdx.z, tid = threadIdx.x;
  const float* __restrict__ S = src + f * src_fstride;
  float* __restrict__ D = dst + f * dst_fstride;
  const int x0 = blockIdx.x * PD_TW, y0 = blockIdx.y * PD_TH;
  const int sx0 = 2 * x0 - 2, sy0 = 2 * y0 - 2;
  {
    // tiles whose source box lies inside the level skip the border mapping (the same samples)
    constexpr int NS = PD_SW * PD_SH, PER = (NS + 255) / 256, FB = 8;
    auto load_tile = [&](auto inside_c) {
      constexpr bool INSIDE = decltype(inside_c)::value;
#pragma nounroll
      for (int b0 = 0; b0 < PER; b0 += FB) {
        int off[FB], ly[FB], lx[FB];
#pragma unroll
        for (int j = 0; j < FB; ++j) {
          const int i0 = tid + (b0 + j) * 256;
          const bool ok = b0 + j < PER && i0 < NS;
          const int i = ok ? i0 : 0;
          ly[j] = ok ? i / PD_SW : -1;
          lx[j] = i - (i / PD_SW) * PD_SW;
          if (INSIDE) {
            off[j] = (sy0 + i / PD_SW) * src_pitch + (sx0 + lx[j]);
          } else {
            const int bx = border_interp(sx0 + lx[j], sw, border), by = border_interp(sy0 + i / PD_SW, sh, border);
            off[j] = (bx < 0 || by < 0) ? -1 : by * src_pitch + bx;
          }
        }
        float v[FB];
#pragma unroll
        for (int j = 0; j < FB; ++j) v[j] = (INSIDE || off[j] >= 0) ? S[off[j]] : 0.f;
#pragma unroll
        for (int j = 0; j < FB; ++j)
          if (ly[j] >= 0) sm[ly[j]][lx[j]] = v[j];
      }
    };
    if (sx0 >= 0 && sy0 >= 0 && sx0 + PD_SW <= sw && sy0 + PD_SH <= sh)
      load_tile(std::true_type{});
    else
      load_tile(std::false_type{});
  }
  __syncthreads();
  for (int i = threadIdx.x; i < PD_SH * PD_TW; i += blockDim.x) {
    const int ly = i / PD_TW, lx = i - ly * PD_TW;
    const float* c = &sm[ly][2 * lx + 2];
    rb[ly][lx] = fmaf(c[-1] + c[1], 4.f, fmaf(c[0], 6.f, c[-2] + c[2]));
  }
  __syncthreads();
  const int lx = threadIdx.x & (PD_TW - 1);
  for (int ly = threadIdx.x / PD_TW; ly < PD_TH; ly += blockDim.x / PD_TW) {
    const int gx = x0 + lx, gy = y0 + ly;
    if (gx < dw && gy < dh) {
      const int r = 2 * ly + 2;
      const float v = fmaf(rb[r - 1][lx] + rb[r + 1][lx], 4.f, fmaf(rb[r][lx], 6.f, rb[r - 2][lx] + rb[r + 2][lx]));
      D[(size_t)gy * dst_pitch + gx] = v * (1.f / 256.f);
    }
  }
}

hipError_t launch_pyrdown(const float* src, long long src_fstride, int src_pitch, int sw, int sh, float* dst,
                          long long dst_fstride, int dst_pitch, int dw, int dh, int nframes, int border,
                          hipStream_t s) {
  if (nframes <= 0) return hipErrorInvalidValue;
  dim3 grid((dw + PD_TW - 1) / PD_TW, (dh + PD_TH - 1) / PD_TH, nframes);
  hipLaunchKernelGGL(k_pyrdown, grid, dim3(256), 0, s, src, src_fstride, src_pitch, sw, sh, dst, dst_fstride,
                     dst_pitch, dw, dh, border);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// pyrDown of a level with no LDS (R == 0), or the registration pyramid's
// smoothed level 0 (a symmetric gaussian of radius R with BORDER_REPLICATE)
// and level 1 in one pass (R > 0), so that level 0 is stored once and never
// read back. Each wave owns a strip of 240 source columns (4 per lane, 8 more
// either side as halo) and slides down 128 source rows. A lane takes its
// neighbours' columns with DPP lane shifts for the row filters; the smoothing
// column window and pyrDown's five rows (BORDER_REFLECT101) stay in
// registers; source rows are loaded SW_D rows ahead. Same expressions in the
// same order as k_sepfilter_t and k_pyrdown, so the same values.
// ---------------------------------------------------------------------------
constexpr int SW_COLS = 240;  // source columns a strip stores (lanes 2..61)
constexpr int SW_RB = 128;    // source rows a wave covers (fewer for small batches, down to 16)
constexpr int SW_D = 4;       // source rows in flight per wave

__device__ __forceinline__ float from_lane_below(float v) {  // lane i receives lane i-1 (wave_shr:1)
  return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x138, 0xF, 0xF, true));
}
__device__ __forceinline__ float from_lane_above(float v) {  // lane i receives lane i+1 (wave_shl:1)
  return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x130, 0xF, 0xF, true));
}

struct StripArgs {
  const float* const* src_ptrs;  // frame f's source (always when R > 0), else
  const float* src;              // frame f at src + f * src_fstride
  long long src_fstride;
  int src_pitch, w, h;
  float* d0;                     // R > 0: the smoothed level, frame f at d0 + f * d0_fstride
  long long d0_fstride;
  int pitch0;
  float* d1;                     // pyrDown output, frame f at d1 + f * d1_fstride
  long long d1_fstride;
  int pitch1, dw, dh;
  int rb;                        // source rows per wave (even)
  int nstrips, nchunks, nwaves, per_xcd;
};

template <int R>
__global__ void __launch_bounds__(256) k_strip_pyrdown(StripArgs a, SepKernel K) {
  static_assert(R >= 0 && R <= 4, "a lane's neighbours hold 4 columns each");
  constexpr bool SMOOTH = R > 0;
  const int lane = threadIdx.x & 63;
  // consecutive waves (neighbouring strips and chunks, which share halo rows and columns) on one XCD
  const int lb = (blockIdx.x & 7) * a.per_xcd + (blockIdx.x >> 3);
  const int gw = __builtin_amdgcn_readfirstlane(lb * 4 + (threadIdx.x >> 6));
  if (gw >= a.nwaves) return;
  const int w = a.w, h = a.h, dw = a.dw, dh = a.dh;
  const int sx = gw % a.nstrips, t = gw / a.nstrips, cy = t % a.nchunks, f = t / a.nchunks;
  const float* __restrict__ S = (SMOOTH || a.src_ptrs) ? a.src_ptrs[f] : a.src + f * a.src_fstride;
  float* __restrict__ D0 = SMOOTH ? a.d0 + f * a.d0_fstride : nullptr;
  float* __restrict__ D1 = a.d1 + f * a.d1_fstride;
  const int xb = sx * SW_COLS - 8, c = xb + 4 * lane;
  const bool vec = ((uintptr_t)S & 15) == 0 && c >= 0 && c + 3 < w;
  const int cx0 = min(max(c, 0), w - 1), cx1 = min(max(c + 1, 0), w - 1), cx2 = min(max(c + 2, 0), w - 1),
            cx3 = min(max(c + 3, 0), w - 1);
  // rows and columns outside the frame read the nearest edge sample (the smoothing's
  // BORDER_REPLICATE; pyrDown alone reads none that it uses)
  auto load_row = [&](int r) -> float4 {
    const float* row = S + (size_t)min(max(r, 0), h - 1) * a.src_pitch;
    if (vec) return *reinterpret_cast<const float4*>(row + c);
    return make_float4(row[cx0], row[cx1], row[cx2], row[cx3]);
  };
  auto row_filt = [&](float4 q) -> float4 {
    constexpr int RR = SMOOTH ? R : 1;
    const float lo[4] = {from_lane_below(q.x), from_lane_below(q.y), from_lane_below(q.z), from_lane_below(q.w)};
    const float hi[4] = {from_lane_above(q.x), from_lane_above(q.y), from_lane_above(q.z), from_lane_above(q.w)};
    float e[4 + 2 * RR];
#pragma unroll
    for (int i = 0; i < RR; ++i) {
      e[i] = lo[4 - RR + i];
      e[RR + 4 + i] = hi[i];
    }
    e[RR] = q.x;
    e[RR + 1] = q.y;
    e[RR + 2] = q.z;
    e[RR + 3] = q.w;
    float4 o;
    o.x = row_filter_r<RR>(&e[RR], K);
    o.y = row_filter_r<RR>(&e[RR + 1], K);
    o.z = row_filter_r<RR>(&e[RR + 2], K);
    o.w = row_filter_r<RR>(&e[RR + 3], K);
    return o;
  };
  const bool xedge = xb < 0 || xb + 256 > w;  // pyrDown reads reflected columns (BORDER_REFLECT101)
  // pyrDown's row pass at this lane's two even columns c and c + 2
  auto pyr_row = [&](float4 s) -> float2 {
    float e[7] = {from_lane_below(s.z), from_lane_below(s.w), s.x, s.y, s.z, s.w, from_lane_above(s.x)};
    if (xedge) {
      float m[7];
#pragma unroll
      for (int i = 0; i < 7; ++i) {
        const int j = border_interp(c - 2 + i, w, BORDER_REFLECT101) - (c - 2);
        float v = e[i];
#pragma unroll
        for (int k = 0; k < 7; ++k) v = (j == k) ? e[k] : v;
        m[i] = v;
      }
#pragma unroll
      for (int i = 0; i < 7; ++i) e[i] = m[i];
    }
    return make_float2(fmaf(e[1] + e[3], 4.f, fmaf(e[2], 6.f, e[0] + e[4])),
                       fmaf(e[3] + e[5], 4.f, fmaf(e[4], 6.f, e[2] + e[6])));
  };
  const bool st_lane = lane >= 2 && lane < 62;
  const int X = c >> 1;
  auto store1 = [&](int Y, float2 r0, float2 r1, float2 r2, float2 r3, float2 r4) {
    if (st_lane && X < dw) {
      float2 v;
      v.x = fmaf(r1.x + r3.x, 4.f, fmaf(r2.x, 6.f, r0.x + r4.x)) * (1.f / 256.f);
      v.y = fmaf(r1.y + r3.y, 4.f, fmaf(r2.y, 6.f, r0.y + r4.y)) * (1.f / 256.f);
      float* d = D1 + (size_t)Y * a.pitch1 + X;
      if (X + 1 < dw)
        *reinterpret_cast<float2*>(d) = v;
      else
        d[0] = v.x;
    }
  };
  const int ys = cy * a.rb, y_end0 = min(ys + a.rb, h);
  const int Ya = ys >> 1, Yb = min((ys + a.rb) >> 1, dh);
  // source rows formed: s_first .. s_stop-1, those the stores and the level-1 rows read
  const int s_first = ys == 0 ? 0 : ys - 2, s_stop = min(max(SMOOTH ? y_end0 : 0, 2 * Yb + 1), h);
  constexpr int NW = SMOOTH ? 2 * R + 1 : 1;
  float4 wq[NW], pq[SW_D];
  if constexpr (SMOOTH) {
    float4 init[2 * R];
#pragma unroll
    for (int i = 0; i < 2 * R; ++i) init[i] = load_row(s_first - R + i);
#pragma unroll
    for (int d = 0; d < SW_D; ++d) pq[d] = load_row(s_first + R + d);
#pragma unroll
    for (int i = 0; i < 2 * R; ++i) wq[i + 1] = row_filt(init[i]);
  } else {
#pragma unroll
    for (int d = 0; d < SW_D; ++d) pq[d] = load_row(s_first + d);
  }
  float2 hist[5];
#pragma unroll
  for (int i = 0; i < 5; ++i) hist[i] = make_float2(0.f, 0.f);
#pragma nounroll
  for (int sr = s_first; sr < s_stop; ++sr) {
    const float4 raw = pq[0];
#pragma unroll
    for (int d = 0; d < SW_D - 1; ++d) pq[d] = pq[d + 1];
    pq[SW_D - 1] = load_row(sr + R + SW_D);
    float4 s;
    if constexpr (SMOOTH) {
#pragma unroll
      for (int i = 0; i < 2 * R; ++i) wq[i] = wq[i + 1];
      wq[2 * R] = row_filt(raw);
      float a0[NW], a1[NW], a2[NW], a3[NW];
#pragma unroll
      for (int i = 0; i < NW; ++i) {
        a0[i] = wq[i].x;
        a1[i] = wq[i].y;
        a2[i] = wq[i].z;
        a3[i] = wq[i].w;
      }
      s = make_float4(col_filter_r<R>(a0, K), col_filter_r<R>(a1, K), col_filter_r<R>(a2, K), col_filter_r<R>(a3, K));
      if (sr >= ys && sr < y_end0 && st_lane && c < w) {
        float* d = D0 + (size_t)sr * a.pitch0 + c;
        if (c + 3 < w) {
          *reinterpret_cast<float4*>(d) = s;
        } else {
          d[0] = s.x;
          if (c + 1 < w) d[1] = s.y;
          if (c + 2 < w) d[2] = s.z;
        }
      }
    } else {
      s = raw;
    }
#pragma unroll
    for (int i = 0; i < 4; ++i) hist[i] = hist[i + 1];
    hist[4] = pyr_row(s);
    // level-1 row Y reads source rows 2Y-2 .. 2Y+2 (rows -2, -1 reflect to 2, 1)
    if (sr >= 2 && !(sr & 1)) {
      const int Y = (sr - 2) >> 1;
      if (Y >= Ya && Y < Yb) {
        if (Y == 0)
          store1(Y, hist[4], hist[3], hist[2], hist[3], hist[4]);
        else
          store1(Y, hist[0], hist[1], hist[2], hist[3], hist[4]);
      }
    }
  }
  // the last level-1 row may read rows h and h+1, which reflect to h-2 and h-3
  if (s_stop == h) {
    const int Y = (h - 1) >> 1;
    if (Y >= Ya && Y < Yb) {
      if (h & 1)
        store1(Y, hist[2], hist[3], hist[4], hist[3], hist[2]);
      else
        store1(Y, hist[1], hist[2], hist[3], hist[4], hist[3]);
    }
  }
}

namespace {
StripArgs strip_grid(StripArgs a, int nframes, dim3& g) {
  a.nstrips = (a.w + SW_COLS - 1) / SW_COLS;
  // shorter chunks while the launch would have fewer than 8 waves per CU (a single frame)
  a.rb = SW_RB;
  while (a.rb > 16 && (long long)a.nstrips * ((a.h + a.rb - 1) / a.rb) * nframes < 2048) a.rb >>= 1;
  a.nchunks = (a.h + a.rb - 1) / a.rb;
  const long long nw = (long long)a.nstrips * a.nchunks * nframes, nb = (nw + 3) / 4, per = (nb + 7) / 8;
  a.nwaves = (int)nw;
  a.per_xcd = (int)per;
  g = dim3((unsigned)(per * 8));
  return a;
}
bool strip_grid_fits(int w, int h, int nframes) {
  const long long nw = (long long)((w + SW_COLS - 1) / SW_COLS) * ((h + 15) / 16) * nframes;  // the most chunks
  return nw + 32 < (long long)INT32_MAX;
}
}  // namespace

bool strip_pyrdown_supported(const float* src, long long src_fstride, int src_pitch, int w, int h, const float* dst,
                             long long dst_fstride, int dst_pitch, int dw, int dh, int nframes) {
  // src null: the frames come through a pointer table (each frame's alignment is checked on the device)
  return nframes > 0 && w >= 8 && h >= 8 && src_pitch % 4 == 0 && src_pitch >= w &&
         (!src || (src_fstride % 4 == 0 && ((uintptr_t)src & 15) == 0)) && ((uintptr_t)dst & 7) == 0 &&
         dst_fstride % 2 == 0 && dst_pitch % 2 == 0 && dst_pitch >= dw && dw >= 1 && dh >= 1 && 2 * dw <= w + 1 &&
         2 * dh <= h + 1 && strip_grid_fits(w, h, nframes);
}

hipError_t launch_strip_pyrdown(const float* const* src_ptrs, const float* src, long long src_fstride, int src_pitch,
                                int w, int h, float* dst, long long dst_fstride, int dst_pitch, int dw, int dh,
                                int nframes, hipStream_t s) {
  if (!src_ptrs == !src ||
      !strip_pyrdown_supported(src, src_fstride, src_pitch, w, h, dst, dst_fstride, dst_pitch, dw, dh, nframes))
    return hipErrorInvalidValue;
  StripArgs a{};
  a.src_ptrs = src_ptrs;
  a.src = src;
  a.src_fstride = src_fstride;
  a.src_pitch = src_pitch;
  a.w = w;
  a.h = h;
  a.d1 = dst;
  a.d1_fstride = dst_fstride;
  a.pitch1 = dst_pitch;
  a.dw = dw;
  a.dh = dh;
  dim3 g;
  a = strip_grid(a, nframes, g);
  hipLaunchKernelGGL((k_strip_pyrdown<0>), g, dim3(256), 0, s, a, SepKernel{});
  return hipGetLastError();
}

bool smooth_pyrdown_supported(const SepKernel& k, int src_pitch, const float* pyr, long long fstride, int pitch0,
                              int w, int h, long long off1, int pitch1, int dw, int dh, int nframes) {
  return nframes > 0 && k.nx == k.ny && (k.nx & 1) && k.nx >= 3 && k.nx <= 9 && w >= 8 && h >= 8 &&
         src_pitch % 4 == 0 && src_pitch >= w && pitch0 % 4 == 0 && fstride % 4 == 0 && ((uintptr_t)pyr & 15) == 0 &&
         off1 % 2 == 0 && pitch1 % 2 == 0 && pitch1 >= dw && dw >= 1 && dh >= 1 && 2 * dw <= w + 1 &&
         2 * dh <= h + 1 && strip_grid_fits(w, h, nframes);
}

hipError_t launch_smooth_pyrdown(const float* const* src_ptrs, int src_pitch, float* pyr, long long fstride,
                                 int pitch0, int w, int h, long long off1, int pitch1, int dw, int dh, int nframes,
                                 const SepKernel& k, hipStream_t s) {
  if (!smooth_pyrdown_supported(k, src_pitch, pyr, fstride, pitch0, w, h, off1, pitch1, dw, dh, nframes))
    return hipErrorInvalidValue;
  StripArgs a{};
  a.src_ptrs = src_ptrs;
  a.src_pitch = src_pitch;
  a.w = w;
  a.h = h;
  a.d0 = pyr;
  a.d0_fstride = fstride;
  a.pitch0 = pitch0;
  a.d1 = pyr + off1;
  a.d1_fstride = fstride;
  a.pitch1 = pitch1;
  a.dw = dw;
  a.dh = dh;
  dim3 g;
  a = strip_grid(a, nframes, g);
  switch (k.nx / 2) {
    case 1: hipLaunchKernelGGL((k_strip_pyrdown<1>), g, dim3(256), 0, s, a, k); break;
    case 2: hipLaunchKernelGGL((k_strip_pyrdown<2>), g, dim3(256), 0, s, a, k); break;
    case 3: hipLaunchKernelGGL((k_strip_pyrdown<3>), g, dim3(256), 0, s, a, k); break;
    default: hipLaunchKernelGGL((k_strip_pyrdown<4>), g, dim3(256), 0, s, a, k); break;
  }
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
__global__ void k_copy_frames(const float* const* __restrict__ src_ptrs, int src_pitch, float* __restrict__ dst,
                              long long dst_fstride, int dst_pitch, int w, int h) {
  const int f = blockIdx.z;
  const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y;
  if (x < w) dst[f * dst_fstride + (size_t)y * dst_pitch + x] = src_ptrs[f][(size_t)y * src_pitch + x];
}

hipError_t launch_copy_frames(const float* const* src_ptrs, int src_pitch, float* dst, long long dst_fstride,
                              int dst_pitch, int w, int h, int nframes, hipStream_t s) {
  dim3 grid((w + 255) / 256, h, nframes);
  hipLaunchKernelGGL(k_copy_frames, grid, dim3(256), 0, s, src_ptrs, src_pitch, dst, dst_fstride, dst_pitch, w, h);
  return hipGetLastError();
}

// cv::resize INTER_NEAREST of a u8 mask: sx = min(floor(x * sw/dw), sw-1)
// (also used as a plain copy when the sizes are equal)
__global__ void k_mask_prepare(const uint8_t* const* __restrict__ src_ptrs, int src_pitch, uint8_t* __restrict__ dst,
                               long long dst_fstride, int dst_pitch, int sw, int sh, int dw, int dh) {
  const int f = blockIdx.z;
  const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y;
  if (x >= dw) return;
  const double ifx = (double)sw / dw, ify = (double)sh / dh;
  const int sx = min((int)floor(x * ifx), sw - 1), sy = min((int)floor(y * ify), sh - 1);
  dst[f * dst_fstride + (size_t)y * dst_pitch + x] = src_ptrs[f][(size_t)sy * src_pitch + sx];
}

hipError_t launch_mask_prepare(const uint8_t* const* src_ptrs, int src_pitch, uint8_t* dst, long long dst_fstride,
                               int dst_pitch, int sw, int sh, int dw, int dh, int nframes, hipStream_t s) {
  dim3 grid((dw + 255) / 256, dh, nframes);
  hipLaunchKernelGGL(k_mask_prepare, grid, dim3(256), 0, s, src_ptrs, src_pitch, dst, dst_fstride, dst_pitch, sw, sh,
                     dw, dh);
  return hipGetLastError();
}

// cv::erode with a 5x5 rectangle (min over the window)
__global__ void k_erode5(const uint8_t* __restrict__ src, long long src_fstride, int src_pitch,
                         uint8_t* __restrict__ dst, long long dst_fstride, int dst_pitch, int w, int h, int border,
                         uint8_t cval) {
  const int f = blockIdx.z;
  const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y;
  if (x >= w) return;
  const uint8_t* S = src + f * src_fstride;
  int m = 255;
  for (int j = -2; j <= 2; ++j) {
    const int yy = border_interp(y + j, h, border);
    for (int i = -2; i <= 2; ++i) {
      const int xx = border_interp(x + i, w, border);
      const int v = (xx < 0 || yy < 0) ? cval : S[(size_t)yy * src_pitch + xx];
      m = min(m, v);
    }
  }
  dst[f * dst_fstride + (size_t)y * dst_pitch + x] = (uint8_t)m;
}

hipError_t launch_erode5(const uint8_t* src, long long src_fstride, int src_pitch, uint8_t* dst,
                         long long dst_fstride, int dst_pitch, int w, int h, int nframes, int border, uint8_t cval,
                         hipStream_t s) {
  dim3 grid((w + 255) / 256, h, nframes);
  hipLaunchKernelGGL(k_erode5, grid, dim3(256), 0, s, src, src_fstride, src_pitch, dst, dst_fstride, dst_pitch, w, h,
                     border, cval);
  return hipGetLastError();
}

// number of zero pixels per frame (out[f], must be zeroed by the caller)
__global__ void k_count_zero_u8(const uint8_t* __restrict__ src, long long fstride, int pitch, int w, int h,
                                unsigned* __restrict__ out) {
  const int f = blockIdx.z;
  const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y;
  unsigned z = (x < w && src[f * fstride + (size_t)y * pitch + x] == 0) ? 1u : 0u;
  z = (unsigned)__popcll(__ballot(z));
  if ((threadIdx.x & (WAVE - 1)) == 0 && z) atomicAdd(&out[f], z);
}

hipError_t launch_count_zero_u8(const uint8_t* src, long long fstride, int pitch, int w, int h, int nframes,
                                unsigned* out, hipStream_t s) {
  dim3 grid((w + 255) / 256, h, nframes);
  hipLaunchKernelGGL(k_count_zero_u8, grid, dim3(256), 0, s, src, fstride, pitch, w, h, out);
  return hipGetLastError();
}

// sum of an image in double (one block per row, fixed order: deterministic)
__global__ void k_row_sums(const float* __restrict__ src, int pitch, int w, double* __restrict__ rows) {
  __shared__ double red[4];
  const int y = blockIdx.x;
  double s = 0;
  for (int x = threadIdx.x; x < w; x += blockDim.x) s += (double)src[(size_t)y * pitch + x];
  s = wave_sum(s);
  if ((threadIdx.x & (WAVE - 1)) == 0) red[threadIdx.x / WAVE] = s;
  __syncthreads();
  if (threadIdx.x == 0) rows[y] = red[0] + red[1] + red[2] + red[3];
}

__global__ void k_sum_rows(const double* __restrict__ rows, int h, double* __restrict__ out) {
  if (threadIdx.x == 0) {
    double s = 0;
    for (int y = 0; y < h; ++y) s += rows[y];
    out[0] = s;
  }
}

hipError_t launch_sum_f32(const float* src, int pitch, int w, int h, double* out, hipStream_t s) {
  // out must hold h + 1 doubles: row sums then the total in out[h]
  hipLaunchKernelGGL(k_row_sums, dim3(h), dim3(256), 0, s, src, pitch, w, out);
  hipLaunchKernelGGL(k_sum_rows, dim3(1), dim3(64), 0, s, out, h, out + h);
  return hipGetLastError();
}

// Per-frame sharpness weight: mean of the squared 3x3 laplacian
// [0 1 0; 1 -4 1; 0 1 0], BORDER_REPLICATE.  Canonical per pixel
// L = fma(-4, c, (up + down) + (left + right)), L^2 exact in double.  One block per (row strip,
// frame) writes a double partial, summed in strip order afterwards.  Frames whose rows are
// 16-byte aligned runs of a multiple of 4 pixels: each thread owns 4 adjacent columns, walks the
// strip with the rows above, at and below in registers (one 16-byte load per row) and keeps one
// double sum per column; other frames: one column per thread.
__global__ void __launch_bounds__(256) k_sharpness(const float* const* __restrict__ src_ptrs, int pitch, int w, int h,
                                                   int rows_per_block, double* __restrict__ partial) {
  __shared__ double red[4];
  const int f = blockIdx.y;
  const float* __restrict__ S = src_ptrs[f];
  const int y0 = blockIdx.x * rows_per_block, y1 = min(h, y0 + rows_per_block);
  double acc = 0;
  if ((((uintptr_t)S & 15u) == 0) && (pitch & 3) == 0 && (w & 3) == 0) {
    double a4[4] = {0, 0, 0, 0};
    for (int x = 4 * (int)threadIdx.x; x < w; x += 4 * (int)blockDim.x) {
      const int xl = max(x - 1, 0), xr = min(x + 4, w - 1);
      float4 up = *(const float4*)(S + (size_t)max(y0 - 1, 0) * pitch + x);
      float4 c = *(const float4*)(S + (size_t)y0 * pitch + x);
      for (int y = y0; y < y1; ++y) {
        const float* cr = S + (size_t)y * pitch;
        const float4 dn = *(const float4*)(S + (size_t)min(y + 1, h - 1) * pitch + x);
        const float l = cr[xl], r = cr[xr];
        const float L0 = fmaf(-4.f, c.x, (up.x + dn.x) + (l + c.y));
        const float L1 = fmaf(-4.f, c.y, (up.y + dn.y) + (c.x + c.z));
        const float L2 = fmaf(-4.f, c.z, (up.z + dn.z) + (c.y + c.w));
        const float L3 = fmaf(-4.f, c.w, (up.w + dn.w) + (c.z + r));
        a4[0] = fma((double)L0, (double)L0, a4[0]);
        a4[1] = fma((double)L1, (double)L1, a4[1]);
        a4[2] = fma((double)L2, (double)L2, a4[2]);
        a4[3] = fma((double)L3, (double)L3, a4[3]);
        up = c;
        c = dn;
      }
    }
    acc = (a4[0] + a4[1]) + (a4[2] + a4[3]);
  } else {
    for (int y = y0; y < y1; ++y) {
      const float* up = S + (size_t)max(y - 1, 0) * pitch;
      const float* c = S + (size_t)y * pitch;
      const float* dn = S + (size_t)min(y + 1, h - 1) * pitch;
      for (int x = threadIdx.x; x < w; x += blockDim.x) {
        const float l = c[max(x - 1, 0)], r = c[min(x + 1, w - 1)];
        const float L = fmaf(-4.f, c[x], (up[x] + dn[x]) + (l + r));
        acc = fma((double)L, (double)L, acc);
      }
    }
  }
  acc = wave_sum(acc);
  if ((threadIdx.x & (WAVE - 1)) == 0) red[threadIdx.x / WAVE] = acc;
  __syncthreads();
  if (threadIdx.x == 0) partial[(size_t)f * gridDim.x + blockIdx.x] = red[0] + red[1] + red[2] + red[3];
}

hipError_t launch_sharpness(const float* const* src_ptrs, int pitch, int w, int h, int nframes, double* partial,
                            hipStream_t s) {
  // partial must hold nframes * ceil(h / 16) doubles
  const int rpb = 16;
  dim3 grid((h + rpb - 1) / rpb, nframes);
  hipLaunchKernelGGL(k_sharpness, grid, dim3(256), 0, s, src_ptrs, pitch, w, h, rpb, partial);
  return hipGetLastError();
}

// create_ecc_image's channel extraction of BGR f32 frames (extract_channel,
// core/proc/extract_channel.cc:693-697): channel < 0 is cv::cvtColor(COLOR_BGR2GRAY) as
// (b * 0.114f + g * 0.587f) + r * 0.299f, else cv::extractChannel(channel); blockIdx.z = frame
__global__ void __launch_bounds__(256) k_ecc_channel(const float* const* __restrict__ src, int pitch, int w, int channel,
                                                     float* __restrict__ dst, long long dst_fs, int dst_pitch) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y, f = blockIdx.z;
  if (x >= w) return;
  const float* p = src[f] + (size_t)y * pitch + 3 * (size_t)x;
  dst[f * dst_fs + (size_t)y * dst_pitch + x] =
      channel < 0 ? (p[0] * 0.114f + p[1] * 0.587f) + p[2] * 0.299f : p[channel];
}

hipError_t launch_ecc_channel(const float* const* src, int pitch, int w, int h, int n, int channel, float* dst,
                              long long dst_fs, int dst_pitch, hipStream_t s) {
  hipLaunchKernelGGL(k_ecc_channel, dim3((w + 255) / 256, h, n), dim3(256), 0, s, src, pitch, w, channel, dst, dst_fs,
                     dst_pitch);
  return hipGetLastError();
}

// zero-fill of device memory by a kernel (a memset node is not wanted in captured graphs)
__global__ void __launch_bounds__(256) k_zero_u32(uint32_t* __restrict__ p, size_t n) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) p[i] = 0u;
}

hipError_t launch_zero(void* p, size_t bytes, hipStream_t s) {
  if (bytes % 4) return hipErrorInvalidValue;
  const size_t n = bytes / 4;
  const unsigned g = (unsigned)std::min<size_t>(4096, (n + 255) / 256);
  if (n) hipLaunchKernelGGL(k_zero_u32, dim3(g ? g : 1), dim3(256), 0, s, (uint32_t*)p, n);
  return hipGetLastError();
}

// The frames' weights from the partials of k_sharpness: per frame the partials summed in
// block order, then divided by the pixel count (the same double operations in the same order
// as the host reduction of ssg_frame_sharpness).  One thread per frame.
__global__ void __launch_bounds__(64) k_sharpness_finish(const double* __restrict__ partial, int nb, int nframes,
                                                         double area, float* __restrict__ out) {
  const int f = blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= nframes) return;
  double t = 0;
  for (int b = 0; b < nb; ++b) t += partial[(size_t)f * nb + b];
  out[f] = (float)(t / area);
}

hipError_t launch_sharpness_finish(const double* partial, int nb, int nframes, double area, float* out, hipStream_t s) {
  hipLaunchKernelGGL(k_sharpness_finish, dim3((nframes + 63) / 64), dim3(64), 0, s, partial, nb, nframes, area, out);
  return hipGetLastError();
}

// the last step of ecc_normalize (ecc2.cc:392-396): dst = src - up, then 0 where the mask is
// not 255 (setTo(0, ~mask)); blockIdx.z = frame
__global__ void __launch_bounds__(256) k_norm_sub(const float* const* __restrict__ src, int spitch,
                                                  const float* __restrict__ up, long long ufs,
                                                  const uint8_t* const* __restrict__ masks, int mpitch,
                                                  float* __restrict__ dst, int dpitch, long long dfs, int w) {
  const int x = blockIdx.x * 256 + threadIdx.x, y = blockIdx.y, f = blockIdx.z;
  if (x >= w) return;
  float v = src[f][(size_t)y * spitch + x] - up[f * ufs + (size_t)y * w + x];
  if (masks && masks[f] && masks[f][(size_t)y * mpitch + x] != 255) v = 0.f;
  dst[f * dfs + (size_t)y * dpitch + x] = v;
}

hipError_t launch_norm_sub(const float* const* src, int spitch, const float* up, long long ufs,
                           const uint8_t* const* masks, int mpitch, float* dst, int dpitch, long long dfs, int w, int h,
                           int nframes, hipStream_t s) {
  hipLaunchKernelGGL(k_norm_sub, dim3((w + 255) / 256, h, nframes), dim3(256), 0, s, src, spitch, up, ufs, masks,
                     mpitch, dst, dpitch, dfs, w);
  return hipGetLastError();
}

__global__ void k_u16_to_f32(const uint16_t* __restrict__ src, int src_pitch, float* __restrict__ dst,
                             int dst_pitch, int w, float scale) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y;
  if (x < w) dst[(size_t)y * dst_pitch + x] = (float)src[(size_t)y * src_pitch + x] * scale;
}

hipError_t launch_u16_to_f32(const uint16_t* src, int src_pitch, float* dst, int dst_pitch, int w, int h,
                             float scale, hipStream_t s) {
  dim3 grid((w + 255) / 256, h);
  hipLaunchKernelGGL(k_u16_to_f32, grid, dim3(256), 0, s, src, src_pitch, dst, dst_pitch, w, scale);
  return hipGetLastError();
}

}  // namespace ssg

// ssg_bayer.hip -- c_bayer_average on gfx950 (core/average/c_frame_accumulation.cc:988-1334).
//
// The reference adds a raw Bayer frame into a 3-channel float accumulator and
// counter.  With a remap set, destination pixel (x, y) takes the source point
// p = rmap(x, y), truncates it (src = (int)p), and gives the four taps of the
// source cell the side weights (src+1-p) and (p-src) -- computed in float, as
// the reference's int/float expression is -- multiplied in double with the
// pixel weight; each tap goes to the channel of its own Bayer position and the
// float accumulator takes the double sum rounded back to float (:1044-1081).
// Without a remap, every pixel adds its own value to its own channel
// (:996-1039).  These kernels do exactly that per pixel, frames in order, so
// results match the reference arithmetic bit for bit; a per-tap count of
// non-zero weights gives the integer per-cell sample counts.
//
// Layout: planes [c][h*w] for c = B, G, R (BAYER_B = 0, BAYER_G = 1,
// BAYER_R = 2, c_frame_accumulation.h:230-234).
#include "ssg_engine.hpp"

namespace ssg {

namespace {

constexpr int BY_TW = 64, BY_RPT = 4, BY_TH = BY_RPT * 4;  // batched tile: 64 x 16, 4 rows per thread

__device__ __forceinline__ int bayer_ch(unsigned pattern, int x, int y) {
  return (int)((pattern >> (2 * (((y & 1) << 1) | (x & 1)))) & 3u);
}

template <bool U16>
__device__ __forceinline__ float raw_value(const void* src, size_t idx, float scale) {
  if constexpr (U16) return (float)((const uint16_t*)src)[idx] * scale;
  else return ((const float*)src)[idx];
}

// the four taps of one destination pixel (reference :1049-1081), accumulated in
// registers; acc = (float)((double)acc + v * s)
template <bool U16>
__device__ __forceinline__ void bayer_gather(const void* src, int pitch, int w, int h, float scale, unsigned pattern,
                                             float px, float py, double wt, float acc[3], float cntr[3],
                                             unsigned cnt[3]) {
  const int sx = (int)px, sy = (int)py;
  if (!(sx >= 0 && sx < w - 1 && sy >= 0 && sy < h - 1)) return;
  const double ax = (double)((float)(sx + 1) - px), ay = (double)((float)(sy + 1) - py);
  const double bx = (double)(px - (float)sx), by = (double)(py - (float)sy);
  const double s[4] = {ax * ay * wt, bx * ay * wt, ax * by * wt, bx * by * wt};
  const size_t o = (size_t)sy * pitch + sx;
  const float v[4] = {raw_value<U16>(src, o, scale), raw_value<U16>(src, o + 1, scale),
                      raw_value<U16>(src, o + pitch, scale), raw_value<U16>(src, o + pitch + 1, scale)};
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const int c = bayer_ch(pattern, sx + (k & 1), sy + (k >> 1));
    // select the channel without a dynamic register index
#pragma unroll
    for (int cc = 0; cc < 3; ++cc) {
      if (cc == c) {
        acc[cc] = (float)((double)acc[cc] + (double)v[k] * s[k]);
        cntr[cc] = (float)((double)cntr[cc] + s[k]);
        cnt[cc] += s[k] != 0.0;
      }
    }
  }
}

// ---- one add(src, weights) ------------------------------------------------------------------------
template <bool U16, bool MAP, int WMODE>  // WMODE 0 none, 1 u8 mask, 2 f32 weights
__global__ void __launch_bounds__(256) k_bayer_add(BayerState st, const void* src, int pitch, float scale,
                                                   const float* rmap, int rmap_pitch, const uint8_t* mask,
                                                   int mask_pitch, const float* wts, int wts_pitch) {
  const int x = blockIdx.x * 64 + (threadIdx.x & 63), y = blockIdx.y * 4 + (threadIdx.x >> 6);
  if (x >= st.w || y >= st.h) return;
  if (WMODE == 1 && !mask[(size_t)y * mask_pitch + x]) return;
  const size_t o = (size_t)y * st.w + x;
  if constexpr (!MAP) {
    const int cc = bayer_ch(st.pattern, x, y);
    const float v = raw_value<U16>(src, (size_t)y * pitch + x, scale);
    const size_t q = (size_t)cc * st.plane + o;
    if (WMODE == 2) {
      const float wt = wts[(size_t)y * wts_pitch + x];
      st.acc[q] = st.acc[q] + v * wt;
      st.cntr[q] = st.cntr[q] + wt;
      st.cnt[q] += wt != 0.f;
    } else {
      st.acc[q] = st.acc[q] + v;
      st.cntr[q] = st.cntr[q] + 1.f;
      st.cnt[q] += 1u;
    }
  } else {
    float acc[3], cntr[3];
    unsigned cnt[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      acc[c] = st.acc[(size_t)c * st.plane + o];
      cntr[c] = st.cntr[(size_t)c * st.plane + o];
      cnt[c] = st.cnt[(size_t)c * st.plane + o];
    }
    const float* mp = rmap + (size_t)y * rmap_pitch + 2 * x;
    const double wt = WMODE == 2 ? (double)wts[(size_t)y * wts_pitch + x] : 1.0;
    bayer_gather<U16>(src, pitch, st.w, st.h, scale, st.pattern, mp[0], mp[1], wt, acc, cntr, cnt);
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      st.acc[(size_t)c * st.plane + o] = acc[c];
      st.cntr[(size_t)c * st.plane + o] = cntr[c];
      st.cnt[(size_t)c * st.plane + o] = cnt[c];
    }
  }
}

// ---- batched: registered frames, fused map + base_remap mask + gather ---------------------------
// The map of frame k is the create_remap field of its transform (warp_xy, the
// same arithmetic as ssg_registration_create_remap); its mask is base_remap's
// (c_frame_registration.cc:1319-1338): remap of an all-255 mask (LINEAR,
// CONSTANT 0) >= 255, eroded 5x5 with BORDER_CONSTANT 255.  Each thread holds
// its pixels' accumulators in registers over all frames, in frame order.
template <int MOTION, bool U16>
__global__ void __launch_bounds__(256) k_bayer_batch(BayerBatchArgs a) {
  __shared__ uint8_t mk[BY_TH + 4][BY_TW + 4];
  __shared__ uint8_t rmk[BY_TH + 4][BY_TW];
  const BayerState& st = a.st;
  const int tid = threadIdx.x, lane = tid & (WAVE - 1), wv = tid / WAVE;
  const int x0 = blockIdx.x * BY_TW, y0 = blockIdx.y * BY_TH;
  const int x = x0 + lane;
  float acc[BY_RPT][3], cntr[BY_RPT][3];
  unsigned cnt[BY_RPT][3];
#pragma unroll
  for (int r = 0; r < BY_RPT; ++r) {
    const int y = y0 + wv * BY_RPT + r;
    const bool in = x < st.w && y < st.h;
    const size_t o = (size_t)y * st.w + x;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      acc[r][c] = in ? st.acc[(size_t)c * st.plane + o] : 0.f;
      cntr[r][c] = in ? st.cntr[(size_t)c * st.plane + o] : 0.f;
      cnt[r][c] = in ? st.cnt[(size_t)c * st.plane + o] : 0u;
    }
  }
  for (int k = 0; k < a.nframes; ++k) {
    if (!a.valid[k]) continue;
    MapCoef mc;
#pragma unroll
    for (int i = 0; i < 8; ++i) mc.c[i] = a.maps[k].c[i];
    const void* src = a.frames[k];
    // uniform test: the tile and its erosion halo map strictly inside the frame
    const int ex0 = max(x0 - 2, 0), ey0 = max(y0 - 2, 0);
    const int ex1 = min(x0 + BY_TW + 1, st.w - 1), ey1 = min(y0 + BY_TH + 1, st.h - 1);
    float mnx = 1e30f, mxx = -1e30f, mny = 1e30f, mxy = -1e30f;
    {
      const int cx[4] = {ex0, ex1, ex0, ex1}, cy[4] = {ey0, ey0, ey1, ey1};
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        float sx, sy;
        warp_xy<MOTION>(mc, cx[i], cy[i], sx, sy);
        mnx = fminf(mnx, sx); mxx = fmaxf(mxx, sx); mny = fminf(mny, sy); mxy = fmaxf(mxy, sy);
        if (sx != sx || sy != sy) mnx = -1e30f;  // NaN positions take the bordered path
      }
    }
    const bool interior = mnx >= 0.25f && mny >= 0.25f && mxx <= (float)st.w - 1.25f && mxy <= (float)st.h - 1.25f;
    if (!interior) {
      for (int i = tid; i < (BY_TH + 4) * (BY_TW + 4); i += 256) {
        const int ly = i / (BY_TW + 4), lx = i - ly * (BY_TW + 4);
        const int px = x0 - 2 + lx, py = y0 - 2 + ly;
        uint8_t v = 1;  // outside the destination: the erosion border value 255
        if ((unsigned)px < (unsigned)st.w && (unsigned)py < (unsigned)st.h) {
          float sx, sy;
          warp_xy<MOTION>(mc, px, py, sx, sy);
          v = remap_mask_u8(nullptr, 0, st.w, st.h, quant32(sx), quant32(sy)) >= 255;
        }
        mk[ly][lx] = v;
      }
      __syncthreads();
      for (int i = tid; i < (BY_TH + 4) * BY_TW; i += 256) {
        const int ly = i / BY_TW, lx = i - ly * BY_TW;
        rmk[ly][lx] = mk[ly][lx] & mk[ly][lx + 1] & mk[ly][lx + 2] & mk[ly][lx + 3] & mk[ly][lx + 4];
      }
      __syncthreads();
    }
#pragma unroll
    for (int r = 0; r < BY_RPT; ++r) {
      const int ly = wv * BY_RPT + r, y = y0 + ly;
      if (x >= st.w || y >= st.h) continue;
      if (!interior && !(rmk[ly][lane] & rmk[ly + 1][lane] & rmk[ly + 2][lane] & rmk[ly + 3][lane] & rmk[ly + 4][lane]))
        continue;
      float sx, sy;
      warp_xy<MOTION>(mc, x, y, sx, sy);
      bayer_gather<U16>(src, a.src_pitch, st.w, st.h, a.scale, st.pattern, sx, sy, 1.0, acc[r], cntr[r], cnt[r]);
    }
    if (!interior) __syncthreads();
  }
#pragma unroll
  for (int r = 0; r < BY_RPT; ++r) {
    const int y = y0 + wv * BY_RPT + r;
    if (x >= st.w || y >= st.h) continue;
    const size_t o = (size_t)y * st.w + x;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      st.acc[(size_t)c * st.plane + o] = acc[r][c];
      st.cntr[(size_t)c * st.plane + o] = cntr[r][c];
      st.cnt[(size_t)c * st.plane + o] = cnt[r][c];
    }
  }
}

// compute() (:1205-1238): avg = acc / cntr per channel (float division, 0 where
// cntr <= 0), interleaved BGR; mask = any channel's cntr > 0
__global__ void __launch_bounds__(256) k_bayer_compute(BayerState st, float* avg, int avg_pitch, uint8_t* mask,
                                                       int mask_pitch) {
  const int x = blockIdx.x * 64 + (threadIdx.x & 63), y = blockIdx.y * 4 + (threadIdx.x >> 6);
  if (x >= st.w || y >= st.h) return;
  const size_t o = (size_t)y * st.w + x;
  bool any = false;
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    const float n = st.cntr[(size_t)c * st.plane + o];
    if (avg) avg[(size_t)y * avg_pitch + 3 * x + c] = n > 0.f ? st.acc[(size_t)c * st.plane + o] / n : 0.f;
    any |= n > 0.f;
  }
  if (mask) mask[(size_t)y * mask_pitch + x] = any ? 255 : 0;
}

// get_acc_counters() (:1240-1250): the counter scaled by (1, 0.5, 1), interleaved
__global__ void __launch_bounds__(256) k_bayer_counters(BayerState st, float* out, int out_pitch) {
  const int x = blockIdx.x * 64 + (threadIdx.x & 63), y = blockIdx.y * 4 + (threadIdx.x >> 6);
  if (x >= st.w || y >= st.h) return;
  const size_t o = (size_t)y * st.w + x;
  out[(size_t)y * out_pitch + 3 * x + 0] = st.cntr[o];
  out[(size_t)y * out_pitch + 3 * x + 1] = st.cntr[(size_t)st.plane + o] * 0.5f;
  out[(size_t)y * out_pitch + 3 * x + 2] = st.cntr[(size_t)2 * st.plane + o];
}

}  // namespace

unsigned bayer_pattern_bits(int colorid) {
  // generate_bayer_pattern_mask (:1267-1334): cell positions (0,0) (0,1) (1,0) (1,1)
  static const int tab[4][4] = {{2, 1, 1, 0}, {1, 2, 0, 1}, {1, 0, 2, 1}, {0, 1, 1, 2}};
  if (colorid < 8 || colorid > 11) return 0;
  unsigned bits = 0;
  for (int i = 0; i < 4; ++i) bits |= (unsigned)tab[colorid - 8][i] << (2 * i);
  return bits;
}

hipError_t launch_bayer_add(const BayerState& st, const void* src, int src_pitch, bool u16, float scale,
                            const float* rmap, int rmap_pitch, const uint8_t* mask, int mask_pitch,
                            const float* wts, int wts_pitch, hipStream_t s) {
  dim3 grid((st.w + 63) / 64, (st.h + 3) / 4);
  const int wm = wts ? 2 : (mask ? 1 : 0);
#define SSG_BA(U, M, W) \
  hipLaunchKernelGGL((k_bayer_add<U, M, W>), grid, dim3(256), 0, s, st, src, src_pitch, scale, rmap, rmap_pitch, \
                     mask, mask_pitch, wts, wts_pitch)
#define SSG_BA_W(U, M)            \
  do {                            \
    if (wm == 2) SSG_BA(U, M, 2); \
    else if (wm == 1) SSG_BA(U, M, 1); \
    else SSG_BA(U, M, 0);         \
  } while (0)
  if (u16) {
    if (rmap) SSG_BA_W(true, true);
    else SSG_BA_W(true, false);
  } else {
    if (rmap) SSG_BA_W(false, true);
    else SSG_BA_W(false, false);
  }
#undef SSG_BA_W
#undef SSG_BA
  return hipGetLastError();
}

hipError_t launch_bayer_batch(int motion, bool u16, const BayerBatchArgs& a, hipStream_t s) {
  dim3 grid((a.st.w + BY_TW - 1) / BY_TW, (a.st.h + BY_TH - 1) / BY_TH);
#define SSG_BB(M)                                                                     \
  do {                                                                                \
    if (u16) hipLaunchKernelGGL((k_bayer_batch<M, true>), grid, dim3(256), 0, s, a);  \
    else hipLaunchKernelGGL((k_bayer_batch<M, false>), grid, dim3(256), 0, s, a);     \
  } while (0)
  switch (motion) {
    case MOTION_TRANSLATION: SSG_BB(MOTION_TRANSLATION); break;
    case MOTION_EUCLIDEAN:
    case MOTION_SCALED_EUCLIDEAN: SSG_BB(MOTION_EUCLIDEAN); break;
    case MOTION_AFFINE: SSG_BB(MOTION_AFFINE); break;
    default: return hipErrorInvalidValue;
  }
#undef SSG_BB
  return hipGetLastError();
}

hipError_t launch_bayer_compute(const BayerState& st, float* avg, int avg_pitch, uint8_t* mask, int mask_pitch,
                                hipStream_t s) {
  dim3 grid((st.w + 63) / 64, (st.h + 3) / 4);
  hipLaunchKernelGGL(k_bayer_compute, grid, dim3(256), 0, s, st, avg, avg_pitch, mask, mask_pitch);
  return hipGetLastError();
}

hipError_t launch_bayer_counters(const BayerState& st, float* out, int out_pitch, hipStream_t s) {
  dim3 grid((st.w + 63) / 64, (st.h + 3) / 4);
  hipLaunchKernelGGL(k_bayer_counters, grid, dim3(256), 0, s, st, out, out_pitch);
  return hipGetLastError();
}

}  // namespace ssg

// ssg_accum.hip -- fused remap + mask + weighted accumulation for gfx950, and
// the stand-alone remap used by base_remap.
//
// Accumulator layout (one plane = W*H floats, no padding):
//   planes [0, cn)        sum hi    (sum of w * I, compensated pair)
//   planes [cn, 2cn)      sum lo
//   plane  2cn            weight hi (sum of w)
//   plane  2cn + 1        weight lo
// plus a u32 plane of per-pixel sample counts.  The sum form replaces the
// reference's incremental mean (c_frame_accumulation.cc:115-121); both give
// A = sum(w I) / sum(w).  The pairs are kept with TwoSum so 2000-frame sums
// lose no precision, and compute() divides in FP64.
#include "ssg_stencil.hpp"

namespace ssg {

// BicubicTab_i for the u8 mask remap (values computed on the host)
static __constant__ short c_cubic_itab[32][32][16];

void set_cubic_table(const float tab[32][4], hipStream_t s) {
  (void)hipMemcpyToSymbolAsync(HIP_SYMBOL(c_cubic_tab), tab, sizeof(float) * 32 * 4, 0, hipMemcpyHostToDevice, s);
}

hipError_t set_cubic_itab(const short* tab, hipStream_t s) {
  return hipMemcpyToSymbolAsync(HIP_SYMBOL(c_cubic_itab), tab, sizeof(short) * 32 * 32 * 16, 0,
                                hipMemcpyHostToDevice, s);
}

__device__ __forceinline__ void two_sum(float a, float b, float& s, float& e) {
  s = a + b;
  const float bb = s - a;
  e = (a - (s - bb)) + (b - bb);
}

// ---- per-channel samplers (bilinear: lin_fast / lin_general in ssg_stencil.hpp) ----------------
// remapBicubic: interior rows as fma chains added row by row; border pixels
// as sum = cval + sum over in-image taps of (S - cval) * w (imgwarp.cpp).
template <int CN>
__device__ __forceinline__ void cub_fast(const float* __restrict__ img, int pitch, int X, int Y, float* out) {
  const int bx = (X >> 5) - 1, by = (Y >> 5) - 1;
  const float* cx = c_cubic_tab[X & 31];
  const float* cy = c_cubic_tab[Y & 31];
  const float* r = img + (size_t)by * pitch + bx * CN;
#pragma unroll
  for (int c = 0; c < CN; ++c) {
    float sum = 0.f;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const float* q = r + (size_t)j * pitch + c;
      float t = q[0] * (cy[j] * cx[0]);
      t = fmaf(q[CN], cy[j] * cx[1], t);
      t = fmaf(q[2 * CN], cy[j] * cx[2], t);
      t = fmaf(q[3 * CN], cy[j] * cx[3], t);
      sum = j == 0 ? t : sum + t;
    }
    out[c] = sum;
  }
}

template <int CN>
__device__ __forceinline__ bool cub_general(const float* __restrict__ img, int pitch, int w, int h, int X, int Y,
                                            int border, float cval, float* out) {
  const int bx = sat_short(X >> 5) - 1, by = sat_short(Y >> 5) - 1;
  if ((unsigned)bx < (unsigned)max(w - 3, 0) && (unsigned)by < (unsigned)max(h - 3, 0)) {
    cub_fast<CN>(img, pitch, X, Y, out);
    return true;
  }
  if (border == BORDER_TRANSPARENT && ((unsigned)(bx + 1) >= (unsigned)w || (unsigned)(by + 1) >= (unsigned)h))
    return false;
  if (border == BORDER_CONSTANT && (bx >= w || bx + 4 <= 0 || by >= h || by + 4 <= 0)) {
#pragma unroll
    for (int c = 0; c < CN; ++c) out[c] = cval;
    return true;
  }
  const float* cx = c_cubic_tab[X & 31];
  const float* cy = c_cubic_tab[Y & 31];
  int xs[4], ys[4];
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    xs[i] = border_interp(bx + i, w, border);
    ys[i] = border_interp(by + i, h, border);
  }
#pragma unroll
  for (int c = 0; c < CN; ++c) {
    float sum = cval;
    for (int j = 0; j < 4; ++j) {
      if (ys[j] < 0) continue;
      const float* r = img + (size_t)ys[j] * pitch + c;
      for (int i = 0; i < 4; ++i)
        if (xs[i] >= 0) sum = fmaf(r[xs[i] * CN] - cval, cy[j] * cx[i], sum);
    }
    out[c] = sum;
  }
  return true;
}

// cv::remap of a constant image of value c with BORDER_CONSTANT 0 at (X, Y): the
// pipeline remaps its per-frame weight image that way (c_image_stacking_pipeline.cc:1653-1660)
template <int INTERP>
__device__ __forceinline__ float const_sample(float c, int w, int h, int X, int Y) {
  if (INTERP == INTER_CUBIC) {
    const int bx = sat_short(X >> 5) - 1, by = sat_short(Y >> 5) - 1;
    const float* cx = c_cubic_tab[X & 31];
    const float* cy = c_cubic_tab[Y & 31];
    if ((unsigned)bx < (unsigned)max(w - 3, 0) && (unsigned)by < (unsigned)max(h - 3, 0)) {
      float sum = 0.f;
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        float t = c * (cy[j] * cx[0]);
        t = fmaf(c, cy[j] * cx[1], t);
        t = fmaf(c, cy[j] * cx[2], t);
        t = fmaf(c, cy[j] * cx[3], t);
        sum = j == 0 ? t : sum + t;
      }
      return sum;
    }
    if (bx >= w || bx + 4 <= 0 || by >= h || by + 4 <= 0) return 0.f;
    float sum = 0.f;
    for (int j = 0; j < 4; ++j) {
      if ((unsigned)(by + j) >= (unsigned)h) continue;
      for (int i = 0; i < 4; ++i)
        if ((unsigned)(bx + i) < (unsigned)w) sum = fmaf(c - 0.f, cy[j] * cx[i], sum);
    }
    return sum;
  }
  const int ix = sat_short(X >> 5), iy = sat_short(Y >> 5);
  float w0, w1, w2, w3;
  bilinear_w(X & 31, Y & 31, w0, w1, w2, w3);
  if ((unsigned)ix < (unsigned)(w - 1) && (unsigned)iy < (unsigned)(h - 1)) return bilinear_sum(c, c, c, c, w0, w1, w2, w3);
  if (ix >= w || ix + 1 < 0 || iy >= h || iy + 1 < 0) return 0.f;
  const bool x0 = (unsigned)ix < (unsigned)w, x1 = (unsigned)(ix + 1) < (unsigned)w;
  const bool y0 = (unsigned)iy < (unsigned)h, y1 = (unsigned)(iy + 1) < (unsigned)h;
  return bilinear_sum(x0 && y0 ? c : 0.f, x1 && y0 ? c : 0.f, x0 && y1 ? c : 0.f, x1 && y1 ? c : 0.f, w0, w1, w2, w3);
}

// remap(mask or 255, INTER_CUBIC, BORDER_CONSTANT 0) at (X, Y), 15-bit fixed point
__device__ __forceinline__ int remap_mask_u8_cubic(const uint8_t* __restrict__ m, int pitch, int w, int h, int X,
                                                   int Y) {
  const int bx = sat_short(X >> 5) - 1, by = sat_short(Y >> 5) - 1;
  const short* wt = c_cubic_itab[Y & 31][X & 31];
  int sum = 0;
  if ((unsigned)bx < (unsigned)max(w - 3, 0) && (unsigned)by < (unsigned)max(h - 3, 0)) {
    for (int j = 0; j < 4; ++j)
      for (int i = 0; i < 4; ++i) sum += (m ? m[(size_t)(by + j) * pitch + bx + i] : 255) * wt[j * 4 + i];
  } else {
    if (bx >= w || bx + 4 <= 0 || by >= h || by + 4 <= 0) return 0;
    for (int j = 0; j < 4; ++j) {
      const int yy = by + j;
      if ((unsigned)yy >= (unsigned)h) continue;
      for (int i = 0; i < 4; ++i) {
        const int xx = bx + i;
        if ((unsigned)xx < (unsigned)w) sum += (m ? m[(size_t)yy * pitch + xx] : 255) * wt[j * 4 + i];
      }
    }
  }
  return min(max((sum + (1 << 14)) >> 15, 0), 255);
}

template <int INTERP>
__device__ __forceinline__ int mask_value(const uint8_t* m, int pitch, int w, int h, int X, int Y) {
  if (INTERP == INTER_CUBIC) return remap_mask_u8_cubic(m, pitch, w, h, X, Y);
  return remap_mask_u8(m, pitch, w, h, X, Y);
}

// ---------------------------------------------------------------------------------------------
// Fused accumulate: for each frame of the chunk, the base_remap mask
// (remap(255) >= 255, erode 5x5 with BORDER_CONSTANT 255), the frame sample
// (cv::remap with the option interpolation and border) and the weighted
// update, all from registers; one write of the chunk partials at the end.
constexpr int AC_TW = 64, AC_WAVES = 4;
constexpr int AC_GROUP = 16;  // frames summed plainly before folding into the compensated pair

// Whether every source position of the destination rectangle [ex0, ex1] x [ey0, ey1] under the
// parametric map lies in [lo, hx] x [lo, hy]: the positions are monotone in x and in y, so the
// corners bound them; a NaN coefficient fails the test.
template <int MOTION>
__device__ __forceinline__ bool tile_inside(const MapCoef& mc, int ex0, int ey0, int ex1, int ey1, float lo, float hx,
                                            float hy) {
  float mnx = 1e30f, mxx = -1e30f, mny = 1e30f, mxy = -1e30f;
  const int cx[4] = {ex0, ex1, ex0, ex1}, cy[4] = {ey0, ey0, ey1, ey1};
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    float sx, sy;
    warp_xy<MOTION>(mc, cx[i], cy[i], sx, sy);
    mnx = fminf(mnx, sx); mxx = fmaxf(mxx, sx); mny = fminf(mny, sy); mxy = fmaxf(mxy, sy);
    if (sx != sx || sy != sy) mnx = -1e30f;
  }
  return mnx >= lo && mny >= lo && mxx <= hx && mxy <= hy;
}

// WMAP: per-frame weight images (frames' own grid, f32) remapped with the
// frame's interpolation and BORDER_CONSTANT 0; else a per-frame scalar weight,
// remapped as a constant image.  Either is then gated by the mask (weights *
// mask / 255, c_image_stacking_pipeline.cc:1704-1712, the mask being 0 or 255).
template <int MOTION, int INTERP, int CN, bool WMAP, int RPT>
__global__ void __launch_bounds__(256) k_accumulate(AccumArgs a) {
  constexpr int TH = RPT * AC_WAVES;  // tile rows: RPT rows per thread, AC_WAVES waves
  // MOTION_MAP: the source positions come from a dense per-frame field, staged in LDS for the
  // tile and its 2-px erosion halo (the whole tile takes the bordered path)
  constexpr bool DENSE = MOTION == MOTION_MAP;
  // parametric motion, bilinear sampling and scalar weights: the path of the stack batch
  constexpr bool FAST = !DENSE && !WMAP && INTERP == INTER_LINEAR;
  // the bordered path's warped mask and its erosion
  __shared__ uint8_t mk[TH + 4][AC_TW + 4];
  __shared__ uint8_t rmk[TH + 4][AC_TW];
  __shared__ float2 smap[DENSE ? (TH + 4) * (AC_TW + 4) : 1];
  __shared__ float sred[DENSE ? AC_WAVES * 4 : 1];
  const int tid = threadIdx.x, lane = tid & (WAVE - 1), wv = tid / WAVE;
  // XCD-aware tile order: blocks are dispatched round-robin over the 8 XCDs, so dispatch slot b
  // runs on XCD b % 8; each XCD takes a contiguous range of row-major tiles, whose neighbours
  // (left/right and above/below) then fetch the source lines they share through that XCD's L2
  int tx, ty, chunk;
  {
    const int G = gridDim.x * gridDim.y * gridDim.z;
    const int b = blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z);
    const int q = G >> 3, rm = G & 7, j = b & 7;
    const int t = j * q + min(j, rm) + (b >> 3);
    tx = t % gridDim.x;
    const int rest = t / gridDim.x;
    ty = rest % gridDim.y;
    chunk = rest / gridDim.y;
  }
  const int x0 = tx * AC_TW, y0 = ty * TH;
  const int kb = chunk * a.per_chunk, kend = min(kb + a.per_chunk, a.nframes);
  const int x = x0 + lane, xc = min(x, a.dw - 1);  // lanes past the right edge sample column dw - 1
  float shi[RPT][CN], slo[RPT][CN], sg[RPT][CN];
  float whi[RPT], wlo[RPT], wg[RPT];
  unsigned cnt[RPT];
#pragma unroll
  for (int r = 0; r < RPT; ++r) {
#pragma unroll
    for (int c = 0; c < CN; ++c) shi[r][c] = slo[r][c] = sg[r][c] = 0.f;
    whi[r] = wlo[r] = wg[r] = 0.f;
    cnt[r] = 0;
  }
  int ingroup = 0;
  // every AC_GROUP frames the plain group sums fold into the compensated pairs
  auto fold = [&]() {
    ingroup = 0;
#pragma unroll
    for (int r = 0; r < RPT; ++r) {
#pragma unroll
      for (int c = 0; c < CN; ++c) {
        float s_, e_;
        two_sum(shi[r][c], sg[r][c], s_, e_);
        shi[r][c] = s_;
        slo[r][c] += e_;
        sg[r][c] = 0.f;
      }
      float s_, e_;
      two_sum(whi[r], wg[r], s_, e_);
      whi[r] = s_;
      wlo[r] += e_;
      wg[r] = 0.f;
    }
  };
  const float lo_m = INTERP == INTER_CUBIC ? 1.25f : 0.25f;
  const float hi_x = (float)a.sw - (INTERP == INTER_CUBIC ? 3.25f : 1.25f);
  const float hi_y = (float)a.sh - (INTERP == INTER_CUBIC ? 3.25f : 1.25f);
  // INTER_LINEAR tap weights of the 32 x 32 sub-pixel positions (cv::remap's table)
  __shared__ float4 wtab[FAST ? 1024 : 1];
  if constexpr (FAST) {
    for (int i = tid; i < 1024; i += 256) {
      float w0, w1, w2, w3;
      bilinear_w(i & 31, i >> 5, w0, w1, w2, w3);
      wtab[i] = make_float4(w0, w1, w2, w3);
    }
    __syncthreads();
  }
  // tile + erosion halo, clipped to the destination
  const int ex0 = max(x0 - 2, 0), ey0 = max(y0 - 2, 0);
  const int ex1 = min(x0 + AC_TW + 1, a.dw - 1), ey1 = min(y0 + TH + 1, a.dh - 1);
  const uint32_t pb = (uint32_t)a.src_pitch * 4u;
  // source byte offset of tap (ix, iy) from the ubfe fields of quant32_bits (ix + 2^17, iy + 2^17)
  const uint32_t obase = 0u - 0x20000u * (4u * CN + pb);
  unsigned nint = 0;  // frames added through the interior path: they count for every pixel
  for (int g0 = kb; g0 < kend; g0 += WAVE) {
    // lane l classifies frame g0 + l: valid, and (parametric motions) whether the tile and its
    // erosion halo sample inside the fast region, so one ballot covers 64 frames
    unsigned long long vmask, imask;
    {
      const int kl = g0 + lane;
      bool v = false, in = false;
      if (kl < kend && a.valid[kl]) {
        v = true;
        if constexpr (!DENSE) in = tile_inside<MOTION>(a.maps[kl], ex0, ey0, ex1, ey1, lo_m, hi_x, hi_y);
      }
      vmask = __ballot(v);
      imask = __ballot(in);
    }
  while (vmask) {
    const int kj = __builtin_ctzll(vmask);
    vmask &= vmask - 1;
    const int k = g0 + kj;
    MapCoef mc;
    if constexpr (!DENSE) {
#pragma unroll
      for (int i = 0; i < 8; ++i) mc.c[i] = a.maps[k].c[i];
    }
    const float2* __restrict__ rm = DENSE ? a.rmaps[k] : nullptr;
    const float wk = a.weights[k];
    const float* __restrict__ img = a.frames[k];
    bool interior = (imask >> kj) & 1;
    if constexpr (DENSE) {
      // stage the field of the tile + halo and reduce its extent over the block
      float mnx = 1e30f, mxx = -1e30f, mny = 1e30f, mxy = -1e30f;
      for (int i = tid; i < (TH + 4) * (AC_TW + 4); i += 256) {
        const int ly = i / (AC_TW + 4), lx = i - ly * (AC_TW + 4);
        const int px = x0 - 2 + lx, py = y0 - 2 + ly;
        if ((unsigned)px < (unsigned)a.dw && (unsigned)py < (unsigned)a.dh) {
          const float2 m = rm[(size_t)py * a.rmap_pitch + px];
          smap[i] = m;
          mnx = fminf(mnx, m.x); mxx = fmaxf(mxx, m.x); mny = fminf(mny, m.y); mxy = fmaxf(mxy, m.y);
          if (m.x != m.x || m.y != m.y) mnx = -1e30f;  // a NaN position takes the bordered path
        }
      }
#pragma unroll
      for (int o = 32; o > 0; o >>= 1) {
        mnx = fminf(mnx, __shfl_xor(mnx, o));
        mxx = fmaxf(mxx, __shfl_xor(mxx, o));
        mny = fminf(mny, __shfl_xor(mny, o));
        mxy = fmaxf(mxy, __shfl_xor(mxy, o));
      }
      if (lane == 0) {
        sred[wv * 4 + 0] = mnx;
        sred[wv * 4 + 1] = mxx;
        sred[wv * 4 + 2] = mny;
        sred[wv * 4 + 3] = mxy;
      }
      __syncthreads();
#pragma unroll
      for (int q = 0; q < AC_WAVES; ++q) {
        mnx = fminf(mnx, sred[q * 4 + 0]);
        mxx = fmaxf(mxx, sred[q * 4 + 1]);
        mny = fminf(mny, sred[q * 4 + 2]);
        mxy = fmaxf(mxy, sred[q * 4 + 3]);
      }
      interior = mnx >= lo_m && mny >= lo_m && mxx <= hi_x && mxy <= hi_y;
    }
    // Interior tiles, parametric motion, bilinear sampling, scalar weights: every lane's source
    // positions first, then all loads (raw buffer loads; lanes outside the destination sample a
    // clamped pixel, whose position lies inside the tile's hull), then the weighted updates.  Two
    // interior frames go together, so both frames' samples are in flight at once; per pixel the
    // updates stay in frame order and the compensated fold keeps its 16-frame boundaries, so the
    // sums are those of the one-frame path.  The positions quantise with quant32_bits, the tap
    // weights come from the table.
    if constexpr (FAST) {
      if (interior) {
        int k2 = -1;
        if (ingroup + 2 <= AC_GROUP && vmask) {
          const int j2 = __builtin_ctzll(vmask);
          if ((imask >> j2) & 1) {
            k2 = g0 + j2;
            vmask &= vmask - 1;
          }
        }
        const bool pair = k2 >= 0;
        MapCoef mc2;
#pragma unroll
        for (int i = 0; i < 8; ++i) mc2.c[i] = a.maps[pair ? k2 : k].c[i];
        uint32_t o1[RPT], q1[RPT], o2[RPT], q2[RPT];
#pragma unroll
        for (int r = 0; r < RPT; ++r) {
          const int yc = min(y0 + wv * RPT + r, a.dh - 1);
          float sx, sy;
          warp_xy<MOTION>(mc, xc, yc, sx, sy);
          uint32_t tx = quant32_bits(sx), ty = quant32_bits(sy);
          o1[r] = __umul24(__builtin_amdgcn_ubfe(ty, 5, 18), pb) + (__builtin_amdgcn_ubfe(tx, 5, 18) * (4u * CN) + obase);
          q1[r] = ((ty & 31u) << 5) | (tx & 31u);
          if (pair) {
            warp_xy<MOTION>(mc2, xc, yc, sx, sy);
            tx = quant32_bits(sx);
            ty = quant32_bits(sy);
            o2[r] = __umul24(__builtin_amdgcn_ubfe(ty, 5, 18), pb) + (__builtin_amdgcn_ubfe(tx, 5, 18) * (4u * CN) + obase);
            q2[r] = ((ty & 31u) << 5) | (tx & 31u);
          }
        }
        const rsrc_t r1 = make_rsrc(img, (uint32_t)a.sh * pb);
        const rsrc_t r2 = make_rsrc(pair ? a.frames[k2] : img, (uint32_t)a.sh * pb);
        float t1[RPT][4][CN], t2[RPT][4][CN];
#pragma unroll
        for (int r = 0; r < RPT; ++r) {
#pragma unroll
          for (int c = 0; c < CN; ++c) {
            t1[r][0][c] = bload_f32(r1, o1[r] + 4u * c, 0u);
            t1[r][1][c] = bload_f32(r1, o1[r] + 4u * (CN + c), 0u);
            t1[r][2][c] = bload_f32(r1, o1[r] + 4u * c, pb);
            t1[r][3][c] = bload_f32(r1, o1[r] + 4u * (CN + c), pb);
          }
        }
        if (pair) {
#pragma unroll
          for (int r = 0; r < RPT; ++r) {
#pragma unroll
            for (int c = 0; c < CN; ++c) {
              t2[r][0][c] = bload_f32(r2, o2[r] + 4u * c, 0u);
              t2[r][1][c] = bload_f32(r2, o2[r] + 4u * (CN + c), 0u);
              t2[r][2][c] = bload_f32(r2, o2[r] + 4u * c, pb);
              t2[r][3][c] = bload_f32(r2, o2[r] + 4u * (CN + c), pb);
            }
          }
        }
        const float wk2 = pair ? a.weights[k2] : 0.f;
        // lanes and rows outside the destination accumulate their clamped pixel too (never
        // written), which keeps the updates free of branches
#pragma unroll
        for (int r = 0; r < RPT; ++r) {
          float4 w = wtab[q1[r]];
          // the scalar weight remapped as a constant image (const_sample inside the image)
          const float wp = bilinear_sum(wk, wk, wk, wk, w.x, w.y, w.z, w.w);
#pragma unroll
          for (int c = 0; c < CN; ++c)
            sg[r][c] = fmaf(wp, bilinear_sum(t1[r][0][c], t1[r][1][c], t1[r][2][c], t1[r][3][c], w.x, w.y, w.z, w.w),
                            sg[r][c]);
          wg[r] += wp;
          if (pair) {
            w = wtab[q2[r]];
            const float wp2 = bilinear_sum(wk2, wk2, wk2, wk2, w.x, w.y, w.z, w.w);
#pragma unroll
            for (int c = 0; c < CN; ++c)
              sg[r][c] = fmaf(wp2, bilinear_sum(t2[r][0][c], t2[r][1][c], t2[r][2][c], t2[r][3][c], w.x, w.y, w.z, w.w),
                              sg[r][c]);
            wg[r] += wp2;
          }
        }
        const int done = pair ? 2 : 1;
        nint += done;
        ingroup += done;
        if (ingroup == AC_GROUP) fold();
        continue;
      }
    }
    // the interior path of the dense fields and the weight images: one frame at a time
    if constexpr (INTERP == INTER_LINEAR && !FAST) {
      if (interior) {
        int X1[RPT], Y1[RPT];
#pragma unroll
        for (int r = 0; r < RPT; ++r) {
          const int yc = min(y0 + wv * RPT + r, a.dh - 1);
          float sx, sy;
          if constexpr (DENSE) {
            const float2 q = smap[(yc - y0 + 2) * (AC_TW + 4) + (xc - x0) + 2];
            sx = q.x;
            sy = q.y;
          } else {
            warp_xy<MOTION>(mc, xc, yc, sx, sy);
          }
          X1[r] = quant32(sx);
          Y1[r] = quant32(sy);
        }
        const rsrc_t r1 = make_rsrc(img, (uint32_t)a.sh * pb);
        float t1[RPT][4][CN];
#pragma unroll
        for (int r = 0; r < RPT; ++r) {
          const uint32_t o = (uint32_t)(Y1[r] >> 5) * pb + (uint32_t)(X1[r] >> 5) * (4u * CN);
#pragma unroll
          for (int c = 0; c < CN; ++c) {
            t1[r][0][c] = bload_f32(r1, o + 4u * c, 0u);
            t1[r][1][c] = bload_f32(r1, o + 4u * (CN + c), 0u);
            t1[r][2][c] = bload_f32(r1, o + pb + 4u * c, 0u);
            t1[r][3][c] = bload_f32(r1, o + pb + 4u * (CN + c), 0u);
          }
        }
        const float* __restrict__ wimg = WMAP ? a.wmaps[k] : nullptr;
#pragma unroll
        for (int r = 0; r < RPT; ++r) {
          const int y = y0 + wv * RPT + r;
          if (x >= a.dw || y >= a.dh) continue;
          float w0, w1, w2, w3;
          bilinear_w(X1[r] & 31, Y1[r] & 31, w0, w1, w2, w3);
          float wp;
          if (WMAP) {
            float wv1[1];
            lin_fast<1>(wimg, a.wmap_pitch, X1[r], Y1[r], wv1);
            wp = wv1[0];
          } else {
            wp = const_sample<INTERP>(wk, a.sw, a.sh, X1[r], Y1[r]);
          }
#pragma unroll
          for (int c = 0; c < CN; ++c)
            sg[r][c] = fmaf(wp, bilinear_sum(t1[r][0][c], t1[r][1][c], t1[r][2][c], t1[r][3][c], w0, w1, w2, w3),
                            sg[r][c]);
          wg[r] += wp;
          cnt[r] += 1;
        }
        if (DENSE) __syncthreads();  // the next frame rewrites the staged field
        if (++ingroup == AC_GROUP) fold();
        continue;
      }
    }
    if (!interior) {
      // warped mask for the tile + 2-px halo; outside the destination counts as 255
      for (int i = tid; i < (TH + 4) * (AC_TW + 4); i += 256) {
        const int ly = i / (AC_TW + 4), lx = i - ly * (AC_TW + 4);
        const int px = x0 - 2 + lx, py = y0 - 2 + ly;
        uint8_t v = 1;
        if ((unsigned)px < (unsigned)a.dw && (unsigned)py < (unsigned)a.dh) {
          float sx, sy;
          if constexpr (DENSE) {
            const float2 m = smap[i];
            sx = m.x;
            sy = m.y;
          } else {
            warp_xy<MOTION>(mc, px, py, sx, sy);
          }
          v = mask_value<INTERP>(nullptr, 0, a.sw, a.sh, quant32(sx), quant32(sy)) >= 255;
        }
        mk[ly][lx] = v;
      }
      __syncthreads();
      for (int i = tid; i < (TH + 4) * AC_TW; i += 256) {
        const int ly = i / AC_TW, lx = i - ly * AC_TW;
        rmk[ly][lx] = mk[ly][lx] & mk[ly][lx + 1] & mk[ly][lx + 2] & mk[ly][lx + 3] & mk[ly][lx + 4];
      }
      __syncthreads();
    }
#pragma unroll
    for (int r = 0; r < RPT; ++r) {
      const int ly = wv * RPT + r, y = y0 + ly;
      if (x >= a.dw || y >= a.dh) continue;
      bool m = true;
      if (!interior) m = rmk[ly][lane] & rmk[ly + 1][lane] & rmk[ly + 2][lane] & rmk[ly + 3][lane] & rmk[ly + 4][lane];
      if (!m) continue;
      float sx, sy;
      if constexpr (DENSE) {
        const float2 mv = smap[(ly + 2) * (AC_TW + 4) + lane + 2];
        sx = mv.x;
        sy = mv.y;
      } else {
        warp_xy<MOTION>(mc, xc, min(y, a.dh - 1), sx, sy);  // x < dw, y < dh here
      }
      const int X = quant32(sx), Y = quant32(sy);
      float v[CN];
      bool ok;
      if (interior) {
        if (INTERP == INTER_CUBIC) cub_fast<CN>(img, a.src_pitch, X, Y, v);
        else lin_fast<CN>(img, a.src_pitch, X, Y, v);
        ok = true;
      } else {
        ok = INTERP == INTER_CUBIC ? cub_general<CN>(img, a.src_pitch, a.sw, a.sh, X, Y, a.border, a.cval, v)
                                   : lin_general<CN>(img, a.src_pitch, a.sw, a.sh, X, Y, a.border, a.cval, v);
      }
      if (!ok) continue;
      float wp;
      if (WMAP) {
        const float* __restrict__ wimg = a.wmaps[k];
        float wv[1];
        if (interior) {
          if (INTERP == INTER_CUBIC) cub_fast<1>(wimg, a.wmap_pitch, X, Y, wv);
          else lin_fast<1>(wimg, a.wmap_pitch, X, Y, wv);
        } else if (INTERP == INTER_CUBIC) {
          cub_general<1>(wimg, a.wmap_pitch, a.sw, a.sh, X, Y, BORDER_CONSTANT, 0.f, wv);
        } else {
          lin_general<1>(wimg, a.wmap_pitch, a.sw, a.sh, X, Y, BORDER_CONSTANT, 0.f, wv);
        }
        wp = wv[0];
      } else {
        wp = const_sample<INTERP>(wk, a.sw, a.sh, X, Y);
      }
#pragma unroll
      for (int c = 0; c < CN; ++c) sg[r][c] = fmaf(wp, v[c], sg[r][c]);
      wg[r] += wp;
      cnt[r] += 1;
    }
    // the next frame rewrites the LDS tiles (masks, staged field)
    if (DENSE || !interior) __syncthreads();
    if (++ingroup == AC_GROUP) fold();
  }
  }
#pragma unroll
  for (int r = 0; r < RPT; ++r) {
    const int y = y0 + wv * RPT + r;
    if (x >= a.dw || y >= a.dh) continue;
    const size_t o = (size_t)y * a.dw + x;
    float* P = a.part + (size_t)chunk * (2 * CN + 2) * a.plane;
#pragma unroll
    for (int c = 0; c < CN; ++c) {
      float s, e;
      two_sum(shi[r][c], sg[r][c], s, e);
      P[(size_t)c * a.plane + o] = s;
      P[(size_t)(CN + c) * a.plane + o] = slo[r][c] + e;
    }
    float s, e;
    two_sum(whi[r], wg[r], s, e);
    P[(size_t)(2 * CN) * a.plane + o] = s;
    P[(size_t)(2 * CN + 1) * a.plane + o] = wlo[r] + e;
    a.part_cnt[(size_t)chunk * a.plane + o] = cnt[r] + nint;
  }
}

hipError_t launch_accumulate(int motion, const AccumArgs& a, hipStream_t s) {
  const int nch = a.nchunks;
  // the interior path forms tap offsets with 24-bit multiplies and positions with quant32_bits
  if ((long long)a.src_pitch * 4 >= (1LL << 24) || a.sw >= (1 << 17) || a.sh >= (1 << 17)) return hipErrorInvalidValue;
  // tile height: a.rows_per_thread rows per thread (2 or 4) times AC_WAVES waves
  const int rpt = a.rows_per_thread == 2 ? 2 : 4;
  dim3 grid((a.dw + AC_TW - 1) / AC_TW, (a.dh + rpt * AC_WAVES - 1) / (rpt * AC_WAVES), nch);
  const bool cubic = a.interp == INTER_CUBIC;
#define SSG_ACC(M, I, C)                                                                                  \
  do {                                                                                                    \
    if (rpt == 2) {                                                                                       \
      if (a.wmaps) hipLaunchKernelGGL((k_accumulate<M, I, C, true, 2>), grid, dim3(256), 0, s, a);        \
      else hipLaunchKernelGGL((k_accumulate<M, I, C, false, 2>), grid, dim3(256), 0, s, a);               \
    } else {                                                                                              \
      if (a.wmaps) hipLaunchKernelGGL((k_accumulate<M, I, C, true, 4>), grid, dim3(256), 0, s, a);        \
      else hipLaunchKernelGGL((k_accumulate<M, I, C, false, 4>), grid, dim3(256), 0, s, a);               \
    }                                                                                                     \
  } while (0)
#define SSG_ACC_IC(M)                                   \
  do {                                                  \
    if (cubic) {                                        \
      if (a.cn == 3) SSG_ACC(M, INTER_CUBIC, 3);        \
      else SSG_ACC(M, INTER_CUBIC, 1);                  \
    } else {                                            \
      if (a.cn == 3) SSG_ACC(M, INTER_LINEAR, 3);       \
      else SSG_ACC(M, INTER_LINEAR, 1);                 \
    }                                                   \
  } while (0)
  switch (motion) {
    case MOTION_TRANSLATION: SSG_ACC_IC(MOTION_TRANSLATION); break;
    case MOTION_EUCLIDEAN:
    case MOTION_SCALED_EUCLIDEAN: SSG_ACC_IC(MOTION_EUCLIDEAN); break;
    case MOTION_AFFINE: SSG_ACC_IC(MOTION_AFFINE); break;
    case MOTION_MAP: SSG_ACC_IC(MOTION_MAP); break;
    default: return hipErrorInvalidValue;
  }
#undef SSG_ACC_IC
#undef SSG_ACC
  return hipGetLastError();
}

// merge chunk partials into the accumulator in chunk order (deterministic)
__global__ void k_accum_merge(const float* __restrict__ part, const unsigned* __restrict__ part_cnt, int nchunks,
                              long long plane, int cn, float* __restrict__ acc, unsigned* __restrict__ cnt) {
  const long long o = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (o >= plane) return;
  const int np = 2 * cn + 2;
  for (int c = 0; c < cn + 1; ++c) {
    // value c < cn: sum channel c; c == cn: weight
    const int ih = c < cn ? c : 2 * cn, il = c < cn ? cn + c : 2 * cn + 1;
    float hi = acc[(size_t)ih * plane + o], lo = acc[(size_t)il * plane + o];
    for (int k = 0; k < nchunks; ++k) {
      const float* P = part + (size_t)k * np * plane;
      float s, e;
      two_sum(hi, P[(size_t)ih * plane + o], s, e);
      hi = s;
      lo += e + P[(size_t)il * plane + o];
    }
    float s, e;
    two_sum(hi, lo, s, e);
    acc[(size_t)ih * plane + o] = s;
    acc[(size_t)il * plane + o] = e;
  }
  unsigned n = cnt[o];
  for (int k = 0; k < nchunks; ++k) n += part_cnt[(size_t)k * plane + o];
  cnt[o] = n;
}

hipError_t launch_accum_merge(const float* part, const unsigned* part_cnt, int nchunks, long long plane, int cn,
                              float* acc, float* accw, unsigned* cnt, hipStream_t s) {
  (void)accw;
  hipLaunchKernelGGL(k_accum_merge, dim3((unsigned)((plane + 255) / 256)), dim3(256), 0, s, part, part_cnt, nchunks,
                     plane, cn, acc, cnt);
  return hipGetLastError();
}

// c_weigthed_average::add of an aligned frame (c_frame_accumulation.cc:20-129):
// weight 1 (no weights), mask != 0 -> 1, or a float weight > 0.
__global__ void k_acc_add(const float* __restrict__ src, int src_pitch, int w, int h, int cn,
                          const uint8_t* __restrict__ mask, int mask_pitch, const float* __restrict__ wmap,
                          int wmap_pitch, float* __restrict__ acc, unsigned* __restrict__ cnt, long long plane) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y;
  if (x >= w) return;
  float wv = 1.f;
  if (wmap) {
    wv = wmap[(size_t)y * wmap_pitch + x];
    if (!(wv > 0.f)) return;
  } else if (mask) {
    if (!mask[(size_t)y * mask_pitch + x]) return;
  }
  const size_t o = (size_t)y * w + x;
  for (int c = 0; c < cn; ++c) {
    const float v = src[(size_t)y * src_pitch + x * cn + c] * wv;
    float s, e;
    two_sum(acc[(size_t)c * plane + o], v, s, e);
    acc[(size_t)c * plane + o] = s;
    acc[(size_t)(cn + c) * plane + o] += e;
  }
  float s, e;
  two_sum(acc[(size_t)(2 * cn) * plane + o], wv, s, e);
  acc[(size_t)(2 * cn) * plane + o] = s;
  acc[(size_t)(2 * cn + 1) * plane + o] += e;
  cnt[o] += 1;
}

hipError_t launch_acc_add(const float* src, int src_pitch, int w, int h, int cn, const uint8_t* mask, int mask_pitch,
                          const float* wmap, int wmap_pitch, float* acc, float* accw, unsigned* cnt,
                          long long plane, hipStream_t s) {
  (void)accw;
  dim3 grid((w + 255) / 256, h);
  hipLaunchKernelGGL(k_acc_add, grid, dim3(256), 0, s, src, src_pitch, w, h, cn, mask, mask_pitch, wmap, wmap_pitch,
                     acc, cnt, plane);
  return hipGetLastError();
}

// c_weigthed_average::compute: avg = sum / weight in FP64, mask = weight > 0
__global__ void k_acc_compute(const float* __restrict__ acc, int w, int h, int cn, long long plane,
                              float* __restrict__ avg, int avg_pitch, uint8_t* __restrict__ mask, int mask_pitch,
                              double dscale) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y;
  if (x >= w) return;
  const size_t o = (size_t)y * w + x;
  const double W = (double)acc[(size_t)(2 * cn) * plane + o] + (double)acc[(size_t)(2 * cn + 1) * plane + o];
  if (avg) {
    for (int c = 0; c < cn; ++c) {
      const double S = (double)acc[(size_t)c * plane + o] + (double)acc[(size_t)(cn + c) * plane + o];
      avg[(size_t)y * avg_pitch + x * cn + c] = W > 0 ? (float)(S / W * dscale) : 0.f;
    }
  }
  if (mask) mask[(size_t)y * mask_pitch + x] = W > 0 ? 255 : 0;
}

// c_weigthed_average::reinitialize(src, accw) (c_frame_accumulation.cc:171-180) in the sum form:
// the weighted sum becomes src * accw as an exact compensated pair (hi = the rounded product,
// lo = its rounding residue), so compute() divides back to src exactly; the weight planes take
// accw; the per-pixel counts restart at 1 where accw > 0.
__global__ void k_acc_reinit(const float* __restrict__ src, int src_pitch, const float* __restrict__ accw,
                             int accw_pitch, int w, int h, int cn, long long plane, float* __restrict__ acc,
                             unsigned* __restrict__ cnt) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y;
  if (x >= w) return;
  const size_t o = (size_t)y * w + x;
  const float W = accw[(size_t)y * accw_pitch + x];
  for (int c = 0; c < cn; ++c) {
    const float v = src[(size_t)y * src_pitch + x * cn + c];
    const float hi = v * W;
    acc[(size_t)c * plane + o] = hi;
    acc[(size_t)(cn + c) * plane + o] = fmaf(v, W, -hi);
  }
  acc[(size_t)(2 * cn) * plane + o] = W;
  acc[(size_t)(2 * cn + 1) * plane + o] = 0.f;
  cnt[o] = W > 0.f ? 1u : 0u;
}

hipError_t launch_acc_reinit(const float* src, int src_pitch, const float* accw, int accw_pitch, int w, int h, int cn,
                             long long plane, float* acc, unsigned* cnt, hipStream_t s) {
  dim3 grid((w + 255) / 256, h);
  hipLaunchKernelGGL(k_acc_reinit, grid, dim3(256), 0, s, src, src_pitch, accw, accw_pitch, w, h, cn, plane, acc, cnt);
  return hipGetLastError();
}

hipError_t launch_acc_compute(const float* acc, const float* accw, int w, int h, int cn, long long plane,
                              float* avg, int avg_pitch, uint8_t* mask, int mask_pitch, double dscale,
                              hipStream_t s) {
  (void)accw;
  dim3 grid((w + 255) / 256, h);
  hipLaunchKernelGGL(k_acc_compute, grid, dim3(256), 0, s, acc, w, h, cn, plane, avg, avg_pitch, mask, mask_pitch,
                     dscale);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------------------------
// Stand-alone remap (cv::remap) with an analytic map (MOTION >= 0) or a CV_32FC2 field (MOTION < 0)
template <int MOTION>
__device__ __forceinline__ void map_at(const MapCoef& mc, const float* map, int map_pitch, int x, int y, float& sx,
                                       float& sy) {
  if constexpr (MOTION < 0) {
    sx = map[(size_t)y * map_pitch + 2 * x];
    sy = map[(size_t)y * map_pitch + 2 * x + 1];
  } else {
    warp_xy<MOTION>(mc, x, y, sx, sy);
  }
}

template <int MOTION, int CN>
__global__ void k_remap(const float* __restrict__ src, int src_pitch, int sw, int sh, MapCoef mc,
                        const float* __restrict__ map, int map_pitch, float* __restrict__ dst, int dst_pitch, int dw,
                        int dh, int interp, int border, float cval) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y;
  if (x >= dw) return;
  float sx, sy;
  map_at<MOTION>(mc, map, map_pitch, x, y, sx, sy);
  float v[CN];
  bool ok;
  if (interp == INTER_NEAREST) {
    const int ix = sat_short(__float2int_rn(sx)), iy = sat_short(__float2int_rn(sy));
    const int xx = border_interp(ix, sw, border), yy = border_interp(iy, sh, border);
    ok = !(border == BORDER_TRANSPARENT && (xx < 0 || yy < 0));
    for (int c = 0; c < CN; ++c) v[c] = (xx < 0 || yy < 0) ? cval : src[(size_t)yy * src_pitch + xx * CN + c];
  } else if (interp == INTER_CUBIC) {
    ok = cub_general<CN>(src, src_pitch, sw, sh, quant32(sx), quant32(sy), border, cval, v);
  } else {
    ok = lin_general<CN>(src, src_pitch, sw, sh, quant32(sx), quant32(sy), border, cval, v);
  }
  if (!ok) return;
  for (int c = 0; c < CN; ++c) dst[(size_t)y * dst_pitch + x * CN + c] = v[c];
}

template <int MOTION>
static hipError_t launch_remap_m(const float* src, int src_pitch, int sw, int sh, int cn, const MapCoef& mc,
                                 const float* map, int map_pitch, float* dst, int dst_pitch, int dw, int dh,
                                 int interp, int border, float cval, hipStream_t s) {
  dim3 grid((dw + 255) / 256, dh);
  if (cn == 3)
    hipLaunchKernelGGL((k_remap<MOTION, 3>), grid, dim3(256), 0, s, src, src_pitch, sw, sh, mc, map, map_pitch, dst,
                       dst_pitch, dw, dh, interp, border, cval);
  else
    hipLaunchKernelGGL((k_remap<MOTION, 1>), grid, dim3(256), 0, s, src, src_pitch, sw, sh, mc, map, map_pitch, dst,
                       dst_pitch, dw, dh, interp, border, cval);
  return hipGetLastError();
}

hipError_t launch_remap(int motion, const float* src, int src_pitch, int sw, int sh, int cn, const MapCoef& mc,
                        const float* map, int map_pitch, float* dst, int dst_pitch, int dw, int dh, int interp,
                        int border, float cval, hipStream_t s) {
  switch (motion) {
    case -1: return launch_remap_m<-1>(src, src_pitch, sw, sh, cn, mc, map, map_pitch, dst, dst_pitch, dw, dh, interp, border, cval, s);
    case MOTION_TRANSLATION: return launch_remap_m<MOTION_TRANSLATION>(src, src_pitch, sw, sh, cn, mc, map, map_pitch, dst, dst_pitch, dw, dh, interp, border, cval, s);
    case MOTION_EUCLIDEAN:
    case MOTION_SCALED_EUCLIDEAN: return launch_remap_m<MOTION_EUCLIDEAN>(src, src_pitch, sw, sh, cn, mc, map, map_pitch, dst, dst_pitch, dw, dh, interp, border, cval, s);
    case MOTION_AFFINE: return launch_remap_m<MOTION_AFFINE>(src, src_pitch, sw, sh, cn, mc, map, map_pitch, dst, dst_pitch, dw, dh, interp, border, cval, s);
  }
  return hipErrorInvalidValue;
}

// base_remap mask, first half: remap(mask or 255, interp, CONSTANT 0) >= 255 -> 0/255
template <int MOTION>
__global__ void k_remap_mask(const uint8_t* __restrict__ src, int src_pitch, int sw, int sh, MapCoef mc,
                             const float* __restrict__ map, int map_pitch, uint8_t* __restrict__ dst, int dst_pitch,
                             int dw, int dh, int interp) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y;
  if (x >= dw) return;
  float sx, sy;
  map_at<MOTION>(mc, map, map_pitch, x, y, sx, sy);
  const int X = quant32(sx), Y = quant32(sy);
  const int v = interp == INTER_CUBIC ? remap_mask_u8_cubic(src, src_pitch, sw, sh, X, Y)
                                      : remap_mask_u8(src, src_pitch, sw, sh, X, Y);
  dst[(size_t)y * dst_pitch + x] = v >= 255 ? 255 : 0;
}

hipError_t launch_remap_mask(int motion, const uint8_t* src_mask, int src_pitch, int sw, int sh, const MapCoef& mc,
                             const float* map, int map_pitch, uint8_t* dst, int dst_pitch, int dw, int dh,
                             hipStream_t s, int interp) {
  dim3 grid((dw + 255) / 256, dh);
#define SSG_RM(M) hipLaunchKernelGGL((k_remap_mask<M>), grid, dim3(256), 0, s, src_mask, src_pitch, sw, sh, mc, map, map_pitch, dst, dst_pitch, dw, dh, interp)
  switch (motion) {
    case -1: SSG_RM(-1); break;
    case MOTION_TRANSLATION: SSG_RM(MOTION_TRANSLATION); break;
    case MOTION_EUCLIDEAN:
    case MOTION_SCALED_EUCLIDEAN: SSG_RM(MOTION_EUCLIDEAN); break;
    case MOTION_AFFINE: SSG_RM(MOTION_AFFINE); break;
    default: return hipErrorInvalidValue;
  }
#undef SSG_RM
  return hipGetLastError();
}

// c_image_transform::create_remap: the CV_32FC2 field of an analytic map
template <int MOTION>
__global__ void k_create_map(MapCoef mc, float* __restrict__ map, int map_pitch, int w, int h) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y;
  if (x >= w) return;
  float sx, sy;
  warp_xy<MOTION>(mc, x, y, sx, sy);
  map[(size_t)y * map_pitch + 2 * x] = sx;
  map[(size_t)y * map_pitch + 2 * x + 1] = sy;
}

hipError_t launch_create_map(int motion, const MapCoef& mc, float* map, int map_pitch, int w, int h, hipStream_t s) {
  dim3 grid((w + 255) / 256, h);
  switch (motion) {
    case MOTION_TRANSLATION: hipLaunchKernelGGL((k_create_map<MOTION_TRANSLATION>), grid, dim3(256), 0, s, mc, map, map_pitch, w, h); break;
    case MOTION_EUCLIDEAN:
    case MOTION_SCALED_EUCLIDEAN: hipLaunchKernelGGL((k_create_map<MOTION_EUCLIDEAN>), grid, dim3(256), 0, s, mc, map, map_pitch, w, h); break;
    case MOTION_AFFINE: hipLaunchKernelGGL((k_create_map<MOTION_AFFINE>), grid, dim3(256), 0, s, mc, map, map_pitch, w, h); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

}  // namespace ssg

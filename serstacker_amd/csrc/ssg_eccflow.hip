// ssg_eccflow.hip -- dense ECC flow (c_eccflow, core/proc/image_registration/ecc2.cc:2103-2865)
// for gfx950: INTER_AREA cell means, the 3x3 cell blur with the per-cell 2x2 solve, the
// warped residual products, and the INTER_CUBIC flow resizes.
//
// Every kernel is batched over frames (blockIdx.z) with frame-strided buffers, so one launch
// covers a level of every frame in flight.  The arithmetic follows the oracle's canonical
// order (oracle/eccflow.py): OpenCV's resize loops summed in their source order, the
// reference's own per-pixel expressions evaluated as written with no contraction.
#include "ssg_engine.hpp"

namespace ssg {

// ---------------------------------------------------------------------------------------------
// cv::resize INTER_AREA (imgproc resize.cpp): resizeAreaFast_ for whole factors (cells read
// row-major, sum += ((a+b)+c)+d per group of four, then the tail, times 1.f/area) and
// ResizeArea_Invoker otherwise (buf = sum_k S[si] alpha over the row's x entries; the cell
// sum starts at beta*buf of its first row and adds beta*buf for each further row).
// ---------------------------------------------------------------------------------------------
template <int CN>
__global__ void __launch_bounds__(256) k_area(const float* __restrict__ src, long long sfs, int spitch,
                                              float* __restrict__ dst, long long dfs, int dpitch, AreaGeom g) {
  const int dx = blockIdx.x * 64 + threadIdx.x, dy = blockIdx.y * 4 + threadIdx.y;
  if (dx >= g.dw || dy >= g.dh) return;
  const float* S = src + blockIdx.z * sfs;
  float acc[CN];
#pragma unroll
  for (int c = 0; c < CN; ++c) acc[c] = 0.f;
  if (g.fast) {
    const int area = g.isx * g.isy, gend = area & ~3;
    const float* base = S + (size_t)dy * g.isy * spitch + (size_t)dx * g.isx * CN;
    float grp[CN];
#pragma unroll
    for (int c = 0; c < CN; ++c) grp[c] = 0.f;
    int k = 0;
    for (int r = 0; r < g.isy; ++r) {
      const float* row = base + (size_t)r * spitch;
      for (int q = 0; q < g.isx; ++q, ++k) {
#pragma unroll
        for (int c = 0; c < CN; ++c) {
          const float v = row[q * CN + c];
          if (k < gend) {
            grp[c] = (k & 3) ? grp[c] + v : v;
            if ((k & 3) == 3) acc[c] = acc[c] + grp[c];
          } else {
            acc[c] = acc[c] + v;
          }
        }
      }
    }
    const float scale = 1.f / (float)area;
#pragma unroll
    for (int c = 0; c < CN; ++c) acc[c] = acc[c] * scale;
  } else {
    const int j0 = g.yoff[dy], j1 = g.yoff[dy + 1], k0 = g.xoff[dx], k1 = g.xoff[dx + 1];
    for (int j = j0; j < j1; ++j) {
      const float beta = g.yal[j];
      const float* row = S + (size_t)g.ysi[j] * spitch;
      float buf[CN];
#pragma unroll
      for (int c = 0; c < CN; ++c) buf[c] = 0.f;
      for (int k = k0; k < k1; ++k) {
        const float a = g.xal[k];
        const float* p = row + (size_t)g.xsi[k] * CN;
#pragma unroll
        for (int c = 0; c < CN; ++c) buf[c] = buf[c] + p[c] * a;
      }
#pragma unroll
      for (int c = 0; c < CN; ++c) acc[c] = (j == j0) ? beta * buf[c] : acc[c] + beta * buf[c];
    }
  }
  float* D = dst + blockIdx.z * dfs + (size_t)dy * dpitch + (size_t)dx * CN;
#pragma unroll
  for (int c = 0; c < CN; ++c) D[c] = acc[c];
}

hipError_t launch_area(int cn, const float* src, long long sfs, int spitch, float* dst, long long dfs, int dpitch,
                       const AreaGeom& g, int nframes, hipStream_t s) {
  dim3 grid((g.dw + 63) / 64, (g.dh + 3) / 4, nframes), block(64, 4);
  if (cn == 1)
    hipLaunchKernelGGL(k_area<1>, grid, block, 0, s, src, sfs, spitch, dst, dfs, dpitch, g);
  else
    hipLaunchKernelGGL(k_area<2>, grid, block, 0, s, src, sfs, spitch, dst, dfs, dpitch, g);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------------------------
// 3x3 [1/4 1/2 1/4] blur with BORDER_REPLICATE at a cell (sepFilter2D, canonical order of the
// symmetric row and column filters: r = fma(S[-1]+S[1], 1/4, S[0]/2);
// out = fma(r[-1]+r[1], 1/4, fma(r[0], 1/2, 0)))
// ---------------------------------------------------------------------------------------------
template <int CN>
__device__ __forceinline__ float blur3_at(const float* __restrict__ S, int pitch, int w, int h, int x, int y, int c) {
  const int xl = max(x - 1, 0), xr = min(x + 1, w - 1);
  const int yu = max(y - 1, 0), yd = min(y + 1, h - 1);
  auto rowf = [&](int yy) {
    const float* R = S + (size_t)yy * pitch;
    return fmaf(R[xl * CN + c] + R[xr * CN + c], 0.25f, R[x * CN + c] * 0.5f);
  };
  const float r0 = rowf(yu), r1 = rowf(y), r2 = rowf(yd);
  return fmaf(r0 + r2, 0.25f, fmaf(r1, 0.5f, 0.f));
}

// Reference products: P2 = (Ix Ix, Ix Iy), P1 = Iy Iy (cv::multiply, float)
__global__ void k_flow_products(const float* __restrict__ Ix, const float* __restrict__ Iy, int n,
                                float2* __restrict__ P2, float* __restrict__ P1) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float gx = Ix[i], gy = Iy[i];
  P2[i] = make_float2(gx * gx, gx * gy);
  P1[i] = gy * gy;
}

hipError_t launch_flow_products(const float* Ix, const float* Iy, int n, float2* P2, float* P1, hipStream_t s) {
  hipLaunchKernelGGL(k_flow_products, dim3((n + 255) / 256), dim3(256), 0, s, Ix, Iy, n, P2, P1);
  return hipGetLastError();
}

// D = (a00, a01, a11, update_multiplier / (|a00 a11 - a01^2| + reg)) from the area means of
// the products, blurred (ecc2.cc:2623-2663)
__global__ void k_flow_D(const float* __restrict__ A2, const float* __restrict__ A1, int cw, int ch, float um,
                         float reg, float4* __restrict__ D) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y;
  if (x >= cw) return;
  const float a00 = blur3_at<2>(A2, cw * 2, cw, ch, x, y, 0);
  const float a01 = blur3_at<2>(A2, cw * 2, cw, ch, x, y, 1);
  const float a11 = blur3_at<1>(A1, cw, cw, ch, x, y, 0);
  const float p = a00 * a11, q = a01 * a01;
  const float det = fabsf(p - q);
  D[(size_t)y * cw + x] = make_float4(a00, a01, a11, um / (det + reg));
}

hipError_t launch_flow_D(const float* A2, const float* A1, int cw, int ch, float um, float reg, float4* D,
                         hipStream_t s) {
  hipLaunchKernelGGL(k_flow_D, dim3((cw + 63) / 64, ch), dim3(64), 0, s, A2, A1, cw, ch, um, reg, D);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------------------------
// Residual products of one iteration (ecc2.cc:2247-2343): the current level warped by
// identity + uv (INTER_LINEAR, BORDER_REPLICATE, 1/32-px positions), the mask
// remap(current mask, INTER_NEAREST, BORDER_CONSTANT) & reference mask, It = ref - W,
// Itxy = (It Ix, It Iy), zero where the mask is zero.
// ---------------------------------------------------------------------------------------------

__global__ void __launch_bounds__(256) k_flow_residual(FlowLevelArgs a, const float2* __restrict__ uv, long long ufs,
                                                       float2* __restrict__ itxy, long long ifs) {
  const int x = blockIdx.x * 64 + threadIdx.x, y = blockIdx.y * 4 + threadIdx.y;
  if (x >= a.w || y >= a.h) return;
  const int f = blockIdx.z;
  const size_t p = (size_t)y * a.w + x;
  const float2 u = uv[f * ufs + p];
  const float mx = u.x + (float)x, my = u.y + (float)y;
  const float* C = a.cur + f * a.cfs;
  const float W = sample_linear_replicate(C, a.w, a.w, a.h, quant32(mx), quant32(my));
  int m = 255;
  if (a.cmask) {
    const int ix = round_coord(mx), iy = round_coord(my);
    m = ((unsigned)ix < (unsigned)a.w && (unsigned)iy < (unsigned)a.h)
            ? (int)a.cmask[f * a.mfs + (size_t)iy * a.w + ix]
            : 0;
  }
  if (a.rmask) m &= (int)a.rmask[p];
  float2 o = make_float2(0.f, 0.f);
  if (m) {
    const float It = a.ref[p] - W;
    o = make_float2(It * a.Ix[p], It * a.Iy[p]);
  }
  itxy[f * ifs + p] = o;
}

hipError_t launch_flow_residual(const FlowLevelArgs& a, const float2* uv, long long ufs, float2* itxy, long long ifs,
                                int nframes, hipStream_t s) {
  dim3 grid((a.w + 63) / 64, (a.h + 3) / 4, nframes), block(64, 4);
  hipLaunchKernelGGL(k_flow_residual, grid, block, 0, s, a, uv, ufs, itxy, ifs);
  return hipGetLastError();
}

// Per-cell update (ecc2.cc:2357-2390): b = blurred cell means of Itxy,
// uv = idet (a11 b0 - a01 b1, a00 b1 - a01 b0)
__global__ void k_flow_solve(const float* __restrict__ cells, long long cfs, const float4* __restrict__ D, int cw,
                             int ch, float2* __restrict__ cuv, long long ufs) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y;
  if (x >= cw) return;
  const int f = blockIdx.z;
  const float* S = cells + f * cfs;
  const float b0 = blur3_at<2>(S, cw * 2, cw, ch, x, y, 0);
  const float b1 = blur3_at<2>(S, cw * 2, cw, ch, x, y, 1);
  const float4 d = D[(size_t)y * cw + x];
  const float t0 = d.z * b0, t1 = d.y * b1;
  const float t2 = d.x * b1, t3 = d.y * b0;
  cuv[f * ufs + (size_t)y * cw + x] = make_float2(d.w * (t0 - t1), d.w * (t2 - t3));
}

hipError_t launch_flow_solve(const float* cells, long long cfs, const float4* D, int cw, int ch, float2* cuv,
                             long long ufs, int nframes, hipStream_t s) {
  hipLaunchKernelGGL(k_flow_solve, dim3((cw + 63) / 64, ch, nframes), dim3(64), 0, s, cells, cfs, D, cw, ch, cuv,
                     ufs);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------------------------
// cv::resize INTER_CUBIC of a 2-channel flow (resizeGeneric_ with HResizeCubic/VResizeCubic):
// per output pixel the four clamped source rows are filtered horizontally,
// h = ((S0 a0 + S1 a1) + S2 a2) + S3 a3, then vertically in the same order.  The taps and
// float coefficients of each axis come from host tables (CubicAxis).  Modes:
//   CUBIC_ADD   : out = base + cubic(src)                         (cv::add(uv, cuv, uv))
//   CUBIC_PLAIN : out = cubic(src)                                (the level upscale, after k_flow_scale)
//   CUBIC_INIT  : out = (float)((double)cubic(map - identity) * ratio)  (initial flow)
// ---------------------------------------------------------------------------------------------
template <int MODE>
__global__ void __launch_bounds__(256) k_flow_cubic(FlowCubicArgs a) {
  const int x = blockIdx.x * 64 + threadIdx.x, y = blockIdx.y * 4 + threadIdx.y;
  if (x >= a.dw || y >= a.dh) return;
  const int f = blockIdx.z;
  const float2* S = MODE == CUBIC_INIT ? a.maps[f] : a.src + f * a.sfs;
  const int spitch = MODE == CUBIC_INIT ? a.map_pitch : a.sw;
  const int sx = a.xs[x], sy = a.ys[y];
  const float4 al = a.xa[x], be = a.ya[y];
  int cx[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) cx[k] = min(max(sx - 1 + k, 0), a.sw - 1);
  float hx[4], hy[4];
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    const int ry = min(max(sy - 1 + j, 0), a.sh - 1);
    const float2* R = S + (size_t)ry * spitch;
    float2 v[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      float2 t = R[cx[k]];
      if (MODE == CUBIC_INIT) {
        t.x = t.x - (float)cx[k];
        t.y = t.y - (float)ry;
      }
      v[k] = t;
    }
    hx[j] = ((v[0].x * al.x + v[1].x * al.y) + v[2].x * al.z) + v[3].x * al.w;
    hy[j] = ((v[0].y * al.x + v[1].y * al.y) + v[2].y * al.z) + v[3].y * al.w;
  }
  float ox = ((hx[0] * be.x + hx[1] * be.y) + hx[2] * be.z) + hx[3] * be.w;
  float oy = ((hy[0] * be.x + hy[1] * be.y) + hy[2] * be.z) + hy[3] * be.w;
  const size_t p = f * a.dfs + (size_t)y * a.dw + x;
  if (MODE == CUBIC_ADD) {
    const float2 b = a.dst[p];
    ox = b.x + ox;
    oy = b.y + oy;
  } else if (MODE == CUBIC_INIT) {
    ox = (float)((double)ox * a.rx);
    oy = (float)((double)oy * a.ry);
  }
  a.dst[p] = make_float2(ox, oy);
}

hipError_t launch_flow_cubic(int mode, const FlowCubicArgs& a, int nframes, hipStream_t s) {
  dim3 grid((a.dw + 63) / 64, (a.dh + 3) / 4, nframes), block(64, 4);
  switch (mode) {
    case CUBIC_ADD: hipLaunchKernelGGL(k_flow_cubic<CUBIC_ADD>, grid, block, 0, s, a); break;
    case CUBIC_PLAIN: hipLaunchKernelGGL(k_flow_cubic<CUBIC_PLAIN>, grid, block, 0, s, a); break;
    default: hipLaunchKernelGGL(k_flow_cubic<CUBIC_INIT>, grid, block, 0, s, a); break;
  }
  return hipGetLastError();
}

// ---------------------------------------------------------------------------------------------
// Fused iteration for k x k cells, k = 4..32 a power of two dividing the level (the whole-factor
// INTER_AREA and an exact 1/k cubic scale).  One wave owns 64 columns (lane = column) of one
// cell row (k pixel rows):
//   ADD    : uv += cubic(cuv): the lane filters the five cell rows j-2..j+2 horizontally once
//            (rows kj..kj+k-1 take their four vertical taps from them), then per pixel row
//            forms the vertical sum and adds it (cv::add(uv, cuv) of ecc2.cc:2830);
//   RESID  : the residual products of k_flow_residual at the updated uv, and their cell sums in
//            resizeAreaFast_ order -- each group of four columns of a row summed
//            ((a+b)+c)+d across lanes, then the groups added to the cell sum row by row, left
//            to right, at the cell's first lane;
//   TO_MAP : (with ADD, without RESID) the updated flow goes out as identity + uv into the
//            caller's maps (ecc_flow_to_remap) instead of back into uv.
// ---------------------------------------------------------------------------------------------
template <bool ADD, bool RESID, bool TO_MAP>
__global__ void __launch_bounds__(256) k_flow_iter(FlowIterArgs a) {
  const int lane = threadIdx.x & (WAVE - 1);
  const int j = blockIdx.y * 4 + (threadIdx.x >> 6);
  if (j >= a.ch) return;  // whole waves leave; no block barrier below
  const int w = a.L.w, k = a.k, f = blockIdx.z;
  const int x = blockIdx.x * WAVE + lane;
  const bool colok = x < w;
  const int xc = colok ? x : w - 1;
  const int y0 = j * k;
  float2* U = a.uv + f * a.ufs;
  // horizontal cubic of cell rows j-2..j+2 at this column
  float hx[5], hy[5];
  if (ADD) {
    const int sx = a.xs[xc];
    const float4 al = a.xa[xc];
    int cx[4];
#pragma unroll
    for (int t = 0; t < 4; ++t) cx[t] = min(max(sx - 1 + t, 0), a.cw - 1);
    const float2* C = a.cuv + f * a.cfs;
#pragma unroll
    for (int rr = 0; rr < 5; ++rr) {
      const int row = min(max(j - 2 + rr, 0), a.ch - 1);
      const float2* R = C + (size_t)row * a.cw;
      const float2 v0 = R[cx[0]], v1 = R[cx[1]], v2 = R[cx[2]], v3 = R[cx[3]];
      hx[rr] = ((v0.x * al.x + v1.x * al.y) + v2.x * al.z) + v3.x * al.w;
      hy[rr] = ((v0.y * al.x + v1.y * al.y) + v2.y * al.z) + v3.y * al.w;
    }
  }
  const float* Cur = a.L.cur + f * a.L.cfs;
  float accx = 0.f, accy = 0.f;
  const int gper = k >> 2;
  const bool base = (lane & (k - 1)) == 0;
  for (int r = 0; r < k; ++r) {
    const int y = y0 + r;
    const size_t p = (size_t)y * w + xc;
    float2 u = U[p];
    if (ADD) {
      // rows ys[y]-1 .. ys[y]+2 are hx[o..o+3] with o = ys[y] - (j - 1) in {0, 1}
      const bool o = a.ys[y] != j - 1;
      const float4 be = a.ya[y];
      const float h0x = o ? hx[1] : hx[0], h1x = o ? hx[2] : hx[1], h2x = o ? hx[3] : hx[2], h3x = o ? hx[4] : hx[3];
      const float h0y = o ? hy[1] : hy[0], h1y = o ? hy[2] : hy[1], h2y = o ? hy[3] : hy[2], h3y = o ? hy[4] : hy[3];
      const float vx = ((h0x * be.x + h1x * be.y) + h2x * be.z) + h3x * be.w;
      const float vy = ((h0y * be.x + h1y * be.y) + h2y * be.z) + h3y * be.w;
      u.x = u.x + vx;
      u.y = u.y + vy;
      if (colok) {
        if (TO_MAP) a.maps[f][(size_t)y * a.map_pitch + x] = make_float2(u.x + (float)x, u.y + (float)y);
        else U[p] = u;
      }
    }
    if (RESID) {
      const float mx = u.x + (float)xc, my = u.y + (float)y;
      const float Wv = sample_linear_replicate(Cur, w, w, a.L.h, quant32(mx), quant32(my));
      int m = colok ? 255 : 0;
      if (a.L.cmask) {
        const int ix = round_coord(mx), iy = round_coord(my);
        m = ((unsigned)ix < (unsigned)w && (unsigned)iy < (unsigned)a.L.h)
                ? (m & (int)a.L.cmask[f * a.L.mfs + (size_t)iy * w + ix])
                : 0;
      }
      if (a.L.rmask) m &= (int)a.L.rmask[p];
      float p0 = 0.f, p1 = 0.f;
      if (m) {
        const float It = a.L.ref[p] - Wv;
        p0 = It * a.L.Ix[p];
        p1 = It * a.L.Iy[p];
      }
      // ((v_l + v_l+1) + v_l+2) + v_l+3 at lanes l = 0 mod 4
      float gx = p0 + __shfl_down(p0, 1);
      float gy = p1 + __shfl_down(p1, 1);
      gx = gx + __shfl_down(p0, 2);
      gy = gy + __shfl_down(p1, 2);
      gx = gx + __shfl_down(p0, 3);
      gy = gy + __shfl_down(p1, 3);
      for (int q = 0; q < gper; ++q) {
        const float tx = __shfl(gx, lane + 4 * q), ty = __shfl(gy, lane + 4 * q);
        accx = accx + tx;
        accy = accy + ty;
      }
    }
  }
  if (RESID && base && colok) {
    const float scale = 1.f / (float)(k * k);
    float* D = a.cells + f * a.cellfs + (size_t)j * 2 * a.cw + 2 * (x / k);
    D[0] = accx * scale;
    D[1] = accy * scale;
  }
}

hipError_t launch_flow_iter(bool add, bool resid, bool to_map, const FlowIterArgs& a, int nframes, hipStream_t s) {
  dim3 grid((a.L.w + WAVE - 1) / WAVE, (a.ch + 3) / 4, nframes), block(256);
  if (add && resid) hipLaunchKernelGGL((k_flow_iter<true, true, false>), grid, block, 0, s, a);
  else if (resid) hipLaunchKernelGGL((k_flow_iter<false, true, false>), grid, block, 0, s, a);
  else if (add && to_map) hipLaunchKernelGGL((k_flow_iter<true, false, true>), grid, block, 0, s, a);
  else if (add) hipLaunchKernelGGL((k_flow_iter<true, false, false>), grid, block, 0, s, a);
  else return hipErrorInvalidValue;
  return hipGetLastError();
}

// cv::multiply(uv, Scalar(rx, ry), uv) in place: each product taken in double, rounded to float
__global__ void k_flow_scale(float2* __restrict__ uv, long long ufs, int w, int h, double rx, double ry) {
  const int x = blockIdx.x * 64 + threadIdx.x, y = blockIdx.y * 4 + threadIdx.y;
  if (x >= w || y >= h) return;
  float2* p = uv + blockIdx.z * ufs + (size_t)y * w + x;
  const float2 u = *p;
  *p = make_float2((float)((double)u.x * rx), (float)((double)u.y * ry));
}

hipError_t launch_flow_scale(float2* uv, long long ufs, int w, int h, double rx, double ry, int nframes, hipStream_t s) {
  dim3 grid((w + 63) / 64, (h + 3) / 4, nframes), block(64, 4);
  hipLaunchKernelGGL(k_flow_scale, grid, block, 0, s, uv, ufs, w, h, rx, ry);
  return hipGetLastError();
}

// rmap = uv + identity (ecc_flow_to_remap, ecc2.cc:452-481), into the caller's maps
__global__ void k_flow_to_remap(const float2* __restrict__ uv, long long ufs, int w, int h, float2* const* maps,
                                int pitch) {
  const int x = blockIdx.x * 64 + threadIdx.x, y = blockIdx.y * 4 + threadIdx.y;
  if (x >= w || y >= h) return;
  const int f = blockIdx.z;
  const float2 u = uv[f * ufs + (size_t)y * w + x];
  maps[f][(size_t)y * pitch + x] = make_float2(u.x + (float)x, u.y + (float)y);
}

hipError_t launch_flow_to_remap(const float2* uv, long long ufs, int w, int h, float2* const* maps, int pitch,
                                int nframes, hipStream_t s) {
  dim3 grid((w + 63) / 64, (h + 3) / 4, nframes), block(64, 4);
  hipLaunchKernelGGL(k_flow_to_remap, grid, block, 0, s, uv, ufs, w, h, maps, pitch);
  return hipGetLastError();
}

// cv::resize INTER_NEAREST of frame-strided u8 masks: sx = min(floor(dx sw/dw), sw-1)
__global__ void k_nearest_u8(const uint8_t* __restrict__ src, long long sfs, int sw, int sh, uint8_t* __restrict__ dst,
                             long long dfs, int dw, int dh) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y;
  if (x >= dw) return;
  const double ifx = (double)sw / dw, ify = (double)sh / dh;
  const int sx = min((int)floor(x * ifx), sw - 1), sy = min((int)floor(y * ify), sh - 1);
  dst[blockIdx.z * dfs + (size_t)y * dw + x] = src[blockIdx.z * sfs + (size_t)sy * sw + sx];
}

hipError_t launch_nearest_u8(const uint8_t* src, long long sfs, int sw, int sh, uint8_t* dst, long long dfs, int dw,
                             int dh, int nframes, hipStream_t s) {
  hipLaunchKernelGGL(k_nearest_u8, dim3((dw + 255) / 256, dh, nframes), dim3(256), 0, s, src, sfs, sw, sh, dst, dfs,
                     dw, dh);
  return hipGetLastError();
}

}  // namespace ssg

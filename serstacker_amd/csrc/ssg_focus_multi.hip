// ssg_focus_multi.hip -- focus stacking of several independent sequences at once (gfx950).
//
// The lpg-weighted laplacian-pyramid composition (c_laplacian_pyramid_focus_stacking with
// compute_energy replaced by W = GaussianBlur(lpg(k, p)); core/proc/lpg.cc:63-127,277-344,
// core/proc/laplacian_pyramid.cc:44-127, core/average/c_frame_accumulation.cc:645-683) for a
// batch of sequences of different sizes.  Every step of a chunk is ONE launch over all sequences:
// its blocks are split into jobs (sequence, level, kind) by prefix block counts, so the launch
// count per chunk does not grow with the number of sequences and small levels of one sequence
// share the machine with large levels of another.
//
// Steps of a chunk of frames (m per sequence):
//   FM_PREP   the translation-aligned frame (cv::remap INTER_LINEAR, BORDER_REFLECT101) split
//             into level-0 planes, and its channel average;
//   FM_LPG    the lpg measure of the channel average;
//   FM_BLUR   W0 = the gaussian smoothing of the measure (sepFilter2D, BORDER_REFLECT101), one
//             LDS tile per 64x32 block with the whole filter halo;
//   FM_DOWN_* pyrDown of the image planes and of the weight plane to the next level;
//   FM_FUSE   per pixel of level i, frame after frame: L = G_i - pyrUp(G_i+1) (read through
//             L2 from the next level), acc_i = acc_i + W_i * L, wsum_i = wsum_i + W_i -- the
//             laplacian level is never written to HBM; FM_TOP the same with L = G_top.
// Each formula is the oracle's canonical one (ssg_stencil.hpp), and every pixel adds its frames
// in order, so the sums equal those of one add() per frame, bit for bit.
#include <type_traits>

#include "ssg_stencil.hpp"

namespace ssg {

namespace {

constexpr int FB_TW = 64, FB_TH = 32, FB_MAXR = (SEP_MAX_TAPS - 1) / 2;
constexpr int FD_TW = 64, FD_TH = 32, FD_SW = 2 * FD_TW + 4, FD_SH = 2 * FD_TH + 4;

__device__ __forceinline__ int find_job(const FmJobs& J, int b) {
  int j = 0;
  while (j + 1 < J.n && b >= J.j[j].end) ++j;
  return j;
}
__device__ __forceinline__ int job_begin(const FmJobs& J, int j) { return j ? J.j[j - 1].end : 0; }

__device__ __forceinline__ long long plane_n(const FmSeq& q, int i) { return (long long)q.lw[i] * q.lh[i]; }

// ---- FM_PREP / FM_LPG: one thread per level-0 pixel of one frame --------------------------------
__global__ void __launch_bounds__(256) k_fm_prep(FmJobs J, const FmSeq* __restrict__ seqs) {
  const int j = find_job(J, blockIdx.x);
  const FmSeq& q = seqs[J.j[j].seq];
  const long long n0 = (long long)q.w * q.h;
  const int bpf = (int)((n0 + 255) / 256);
  const int lb = blockIdx.x - job_begin(J, j);
  const int k = lb / bpf;
  const long long i = (long long)(lb - k * bpf) * 256 + threadIdx.x;
  if (i >= n0) return;
  const int y = (int)(i / q.w), x = (int)(i - (long long)y * q.w);
  const float* __restrict__ src = q.frames[k];
  float v[3];
  if (q.shift) {
    // the translation map of the registration: source (x + t0, y + t1), quantised to 1/32
    const float sx = (float)x + q.shift[2 * k], sy = (float)y + q.shift[2 * k + 1];
    const int X = quant32(sx), Y = quant32(sy);
    if (q.cn == 3) lin_general<3>(src, q.frame_pitch, q.w, q.h, X, Y, BORDER_REFLECT101, 0.f, v);
    else lin_general<1>(src, q.frame_pitch, q.w, q.h, X, Y, BORDER_REFLECT101, 0.f, v);
  } else {
    for (int c = 0; c < q.cn; ++c) v[c] = src[(size_t)y * q.frame_pitch + (size_t)x * q.cn + c];
  }
  float* G0 = q.G + (size_t)k * q.cn * n0;
  for (int c = 0; c < q.cn; ++c) G0[(size_t)c * n0 + i] = v[c];
  q.gray[(size_t)k * n0 + i] = q.cn == 3 ? gray3(v[0], v[1], v[2]) : v[0];
}

__global__ void __launch_bounds__(256) k_fm_lpg(FmJobs J, const FmSeq* __restrict__ seqs) {
  const int j = find_job(J, blockIdx.x);
  const FmSeq& q = seqs[J.j[j].seq];
  const long long n0 = (long long)q.w * q.h;
  const int bpf = (int)((n0 + 255) / 256);
  const int lb = blockIdx.x - job_begin(J, j);
  const int k = lb / bpf;
  const long long i = (long long)(lb - k * bpf) * 256 + threadIdx.x;
  if (i >= n0) return;
  const int y = (int)(i / q.w), x = (int)(i - (long long)y * q.w);
  q.M[(size_t)k * n0 + i] = lpg_at(q.gray + (size_t)k * n0, q.w, q.w, q.h, x, y, q.alpha, q.beta, q.eps, q.p);
}

// ---- FM_BLUR: W0 = sepFilter2D(M, K, K, BORDER_REFLECT101), 64x32 tiles with the whole halo -----
__global__ void __launch_bounds__(256) k_fm_blur(FmJobs J, const FmSeq* __restrict__ seqs) {
  __shared__ float sm[FB_TH + 2 * FB_MAXR][FB_TW + 2 * FB_MAXR];
  __shared__ float rb[FB_TH + 2 * FB_MAXR][FB_TW];
  const int j = find_job(J, blockIdx.x);
  const FmSeq& q = seqs[J.j[j].seq];
  const SepKernel& K = q.K;
  const long long n0 = (long long)q.w * q.h;
  const int tx = (q.w + FB_TW - 1) / FB_TW, ty = (q.h + FB_TH - 1) / FB_TH;
  const int lb = blockIdx.x - job_begin(J, j);
  const int k = lb / (tx * ty), t = lb - k * (tx * ty);
  const int x0 = (t % tx) * FB_TW, y0 = (t / tx) * FB_TH;
  const float* __restrict__ S = q.M + (size_t)k * n0;
  float* __restrict__ D = q.W + (size_t)k * n0;  // level 0 of the weight pyramid
  const int r = K.nx / 2;
  const int TW = FB_TW + 2 * r, TH = FB_TH + 2 * r;
  for (int i = threadIdx.x; i < TW * TH; i += 256) {
    const int ly = i / TW, lx = i - ly * TW;
    const int bx = border_interp(x0 - r + lx, q.w, BORDER_REFLECT101);
    const int by = border_interp(y0 - r + ly, q.h, BORDER_REFLECT101);
    sm[ly][lx] = S[(size_t)by * q.w + bx];
  }
  __syncthreads();
  for (int i = threadIdx.x; i < TH * FB_TW; i += 256) {
    const int ly = i / FB_TW, lx = i - ly * FB_TW;
    rb[ly][lx] = row_filter(&sm[ly][lx + r], K);
  }
  __syncthreads();
  const int lx = threadIdx.x & (FB_TW - 1);
  for (int ly = threadIdx.x / FB_TW; ly < FB_TH; ly += 256 / FB_TW) {
    const int gx = x0 + lx, gy = y0 + ly;
    if (gx < q.w && gy < q.h) D[(size_t)gy * q.w + gx] = col_filter(&rb[ly + r][lx], FB_TW, K);
  }
}

// The same smoothing for one radius R known at compile time: the row pass computes four
// neighbouring outputs from one register window of 4 + 2R samples, the column pass slides a
// (2R+1)-row register window down an 8-row strip -- each output keeps the tap order of
// row_filter / col_filter (generic row chain, symmetric column form), so the values are those of
// k_fm_blur, with a fraction of its LDS reads.
template <int R>
__global__ void __launch_bounds__(256) k_fm_blur_r(FmJobs J, const FmSeq* __restrict__ seqs) {
  constexpr int TW = FB_TW + 2 * R, TH = FB_TH + 2 * R, N = 2 * R + 1;
  __shared__ float sm[TH][TW];
  __shared__ float rb[TH][FB_TW];
  const int j = find_job(J, blockIdx.x);
  const FmSeq& q = seqs[J.j[j].seq];
  const SepKernel& K = q.K;
  const long long n0 = (long long)q.w * q.h;
  const int tx = (q.w + FB_TW - 1) / FB_TW, ty = (q.h + FB_TH - 1) / FB_TH;
  const int lb = blockIdx.x - job_begin(J, j);
  const int k = lb / (tx * ty), t = lb - k * (tx * ty);
  const int x0 = (t % tx) * FB_TW, y0 = (t / tx) * FB_TH;
  const float* __restrict__ S = q.M + (size_t)k * n0;
  float* __restrict__ D = q.W + (size_t)k * n0;
  float kx[N], ky[R + 1];
#pragma unroll
  for (int i = 0; i < N; ++i) kx[i] = K.kx[i];
#pragma unroll
  for (int i = 0; i <= R; ++i) ky[i] = K.ky[R + i];
  const bool inside = x0 - R >= 0 && y0 - R >= 0 && x0 + FB_TW + R <= q.w && y0 + FB_TH + R <= q.h;
  for (int i = threadIdx.x; i < TW * TH; i += 256) {
    const int ly = i / TW, lx = i - ly * TW;
    int bx = x0 - R + lx, by = y0 - R + ly;
    if (!inside) {
      bx = border_interp(bx, q.w, BORDER_REFLECT101);
      by = border_interp(by, q.h, BORDER_REFLECT101);
    }
    sm[ly][lx] = S[(size_t)by * q.w + bx];
  }
  __syncthreads();
  // row pass: item = (row, group of four columns)
  for (int it = threadIdx.x; it < TH * (FB_TW / 4); it += 256) {
    const int ly = it / (FB_TW / 4), lx = (it - ly * (FB_TW / 4)) * 4;
    float v[N + 3];
#pragma unroll
    for (int i = 0; i < N + 3; ++i) v[i] = sm[ly][lx + i];
#pragma unroll
    for (int o = 0; o < 4; ++o) {
      float a = 0.f;
#pragma unroll
      for (int i = 0; i < N; ++i) a = fmaf(v[o + i], kx[i], a);
      rb[ly][lx + o] = a;
    }
  }
  __syncthreads();
  // column pass: column lx, rows r0 .. r0 + 7
  const int lx = threadIdx.x & (FB_TW - 1), r0 = (threadIdx.x >> 6) * (FB_TH / 4);
  float win[N];
#pragma unroll
  for (int i = 0; i < N - 1; ++i) win[i + 1] = rb[r0 + i][lx];
#pragma unroll
  for (int dy = 0; dy < FB_TH / 4; ++dy) {
#pragma unroll
    for (int i = 0; i < N - 1; ++i) win[i] = win[i + 1];
    win[N - 1] = rb[r0 + dy + 2 * R][lx];
    float a = fmaf(win[R], ky[0], K.delta);
#pragma unroll
    for (int i = 1; i <= R; ++i) a = fmaf(win[R - i] + win[R + i], ky[i], a);
    const int gx = x0 + lx, gy = y0 + r0 + dy;
    if (gx < q.w && gy < q.h) D[(size_t)gy * q.w + gx] = a;
  }
}

// ---- FM_DOWN_G / FM_DOWN_W: pyrDown of level i to level i+1, planes [frame][channel] ----------
// The canonical form of k_pyrdown: r = fma(s-1 + s1, 4, fma(s0, 6, s-2 + s2)) per row, then the
// same along the column, times 1/256; BORDER_REFLECT101.
__global__ void __launch_bounds__(256) k_fm_down(FmJobs J, const FmSeq* __restrict__ seqs) {
  __shared__ float sm[FD_SH][FD_SW + 1];
  __shared__ float rb[FD_SH][FD_TW];
  const int j = find_job(J, blockIdx.x);
  const FmJob jb = J.j[j];
  const FmSeq& q = seqs[jb.seq];
  const int i = jb.level;
  const int sw = q.lw[i], sh = q.lh[i], dw = q.lw[i + 1], dh = q.lh[i + 1];
  const int tx = (dw + FD_TW - 1) / FD_TW, ty = (dh + FD_TH - 1) / FD_TH;
  const int lb = blockIdx.x - job_begin(J, j);
  const int p = lb / (tx * ty), t = lb - p * (tx * ty);
  const bool img = jb.kind == FM_DOWN_G;
  const int planes = img ? q.m * q.cn : q.m;
  const float* __restrict__ S = (img ? q.G : q.W) + (size_t)planes * q.loff[i] + (size_t)p * plane_n(q, i);
  float* __restrict__ D = (img ? q.G : q.W) + (size_t)planes * q.loff[i + 1] + (size_t)p * plane_n(q, i + 1);
  const int x0 = (t % tx) * FD_TW, y0 = (t / tx) * FD_TH;
  const int sx0 = 2 * x0 - 2, sy0 = 2 * y0 - 2;
  {
    // eight loads in flight per thread; tiles whose source box lies inside the level skip the
    // border mapping (the same samples)
    constexpr int NS = FD_SW * FD_SH, PER = (NS + 255) / 256, FB = 8;
    auto load_tile = [&](auto inside_c) {
      constexpr bool INSIDE = decltype(inside_c)::value;
#pragma nounroll
      for (int b0 = 0; b0 < PER; b0 += FB) {
        int off[FB], ly[FB], lx[FB];
#pragma unroll
        for (int jj = 0; jj < FB; ++jj) {
          const int i0 = threadIdx.x + (b0 + jj) * 256;
          const bool ok = b0 + jj < PER && i0 < NS;
          const int e = ok ? i0 : 0;
          ly[jj] = ok ? e / FD_SW : -1;
          lx[jj] = e - (e / FD_SW) * FD_SW;
          const int yy = sy0 + e / FD_SW, xx = sx0 + lx[jj];
          off[jj] = INSIDE ? yy * sw + xx
                           : border_interp(yy, sh, BORDER_REFLECT101) * sw + border_interp(xx, sw, BORDER_REFLECT101);
        }
        float v[FB];
#pragma unroll
        for (int jj = 0; jj < FB; ++jj) v[jj] = S[off[jj]];
#pragma unroll
        for (int jj = 0; jj < FB; ++jj)
          if (ly[jj] >= 0) sm[ly[jj]][lx[jj]] = v[jj];
      }
    };
    if (sx0 >= 0 && sy0 >= 0 && sx0 + FD_SW <= sw && sy0 + FD_SH <= sh) load_tile(std::true_type{});
    else load_tile(std::false_type{});
  }
  __syncthreads();
  for (int e = threadIdx.x; e < FD_SH * FD_TW; e += 256) {
    const int ly = e / FD_TW, lx = e - ly * FD_TW;
    const float* c = &sm[ly][2 * lx + 2];
    rb[ly][lx] = fmaf(c[-1] + c[1], 4.f, fmaf(c[0], 6.f, c[-2] + c[2]));
  }
  __syncthreads();
  const int lx = threadIdx.x & (FD_TW - 1);
  for (int ly = threadIdx.x / FD_TW; ly < FD_TH; ly += 256 / FD_TW) {
    const int gx = x0 + lx, gy = y0 + ly;
    if (gx < dw && gy < dh) {
      const int rr = 2 * ly + 2;
      const float v = fmaf(rb[rr - 1][lx] + rb[rr + 1][lx], 4.f, fmaf(rb[rr][lx], 6.f, rb[rr - 2][lx] + rb[rr + 2][lx]));
      D[(size_t)gy * dw + gx] = v * (1.f / 256.f);
    }
  }
}

// ---- FM_FUSE / FM_TOP: the weighted sums of one level over the chunk's frames -------------------
// A block owns a 64x16 tile of level i (four rows per thread, the sums of the chunk in registers).
// Per frame and channel the horizontally upsampled rows of the next level that the tile's pyrUp
// needs are formed once into LDS (hup of the clamped coarse rows), then every output takes its
// two or three rows from there: pyrup_at's values without its repeated gathers.
constexpr int FF_TW = 64, FF_TH = 16, FF_ROWS = FF_TH / 2 + 4;

__global__ void __launch_bounds__(256) k_fm_fuse(FmJobs J, const FmSeq* __restrict__ seqs) {
  __shared__ float hb[FF_ROWS][FF_TW];
  const int j = find_job(J, blockIdx.x);
  const FmJob jb = J.j[j];
  const FmSeq& q = seqs[jb.seq];
  const int i = jb.level;
  const bool top = jb.kind == FM_TOP;
  const int w = q.lw[i], h = q.lh[i];
  const long long n = (long long)w * h;
  const int tx = (w + FF_TW - 1) / FF_TW;
  const int lb = blockIdx.x - job_begin(J, j);
  const int x0 = (lb % tx) * FF_TW, y0 = (lb / tx) * FF_TH;
  const int lx = threadIdx.x & (FF_TW - 1), r0 = (threadIdx.x >> 6) * (FF_TH / 4);
  const int x = x0 + lx;
  float* __restrict__ A = q.acc + (size_t)q.cn * q.loff[i];
  float* __restrict__ WS = q.wsum + q.loff[i];
  const float* __restrict__ Gi = q.G + (size_t)q.m * q.cn * q.loff[i];
  const float* __restrict__ Wi = q.W + (size_t)q.m * q.loff[i];
  const int sw = top ? 1 : q.lw[i + 1], sh = top ? 1 : q.lh[i + 1];
  const float* __restrict__ Gn = top ? nullptr : q.G + (size_t)q.m * q.cn * q.loff[i + 1];
  const long long nn = (long long)sw * sh;
  // the coarse rows the tile reads: pyrup_at's ye (the odd extra row 2 sh maps to 2 sh - 2)
  auto ye_of = [&](int yo) { return (h > 2 * sh && yo == 2 * sh) ? 2 * sh - 2 : yo; };
  const int cy0 = (min(ye_of(y0), y0) >> 1) - 1;
  float acc[3][FF_TH / 4], ws[FF_TH / 4];
  bool in[FF_TH / 4];
  long long e[FF_TH / 4];
#pragma unroll
  for (int r = 0; r < FF_TH / 4; ++r) {
    const int y = y0 + r0 + r;
    in[r] = x < w && y < h;
    e[r] = in[r] ? (long long)y * w + x : 0;
    for (int c = 0; c < 3; ++c) acc[c][r] = (in[r] && c < q.cn) ? A[(size_t)c * n + e[r]] : 0.f;
    ws[r] = in[r] ? WS[e[r]] : 0.f;
  }
  for (int k = 0; k < q.m; ++k) {
    float wk[FF_TH / 4];
#pragma unroll
    for (int r = 0; r < FF_TH / 4; ++r) wk[r] = in[r] ? Wi[(size_t)k * n + e[r]] : 0.f;
    for (int c = 0; c < q.cn; ++c) {
      const int pl = k * q.cn + c;
      float up[FF_TH / 4];
      if (!top) {
        const float* __restrict__ Sn = Gn + (size_t)pl * nn;
        __syncthreads();  // hb of the previous plane is consumed
        for (int t = threadIdx.x; t < FF_ROWS * FF_TW; t += 256) {
          const int jr = t / FF_TW, cx = t - jr * FF_TW;
          int rr = cy0 + jr;  // srow: REFLECT101 above, replicated below
          if (rr < 0) rr = sh > 1 ? 1 : 0;
          if (rr >= sh) rr = sh - 1;
          hb[jr][cx] = x0 + cx < w ? hup(Sn + (size_t)rr * sw, sw, x0 + cx) : 0.f;
        }
        __syncthreads();
#pragma unroll
        for (int r = 0; r < FF_TH / 4; ++r) {
          const int ye = ye_of(y0 + r0 + r), yy = (ye >> 1) - cy0;
          const float scale = 1.f / 64.f;
          up[r] = (ye & 1) ? (hb[yy][lx] + hb[yy + 1][lx]) * 4.f * scale
                           : fmaf(hb[yy][lx], 6.f, hb[yy - 1][lx] + hb[yy + 1][lx]) * scale;
        }
      }
#pragma unroll
      for (int r = 0; r < FF_TH / 4; ++r) {
        if (!in[r]) continue;
        float L = Gi[(size_t)pl * n + e[r]];
        if (!top) L = L - up[r];
        acc[c][r] = acc[c][r] + wk[r] * L;
      }
    }
#pragma unroll
    for (int r = 0; r < FF_TH / 4; ++r) ws[r] = ws[r] + wk[r];
  }
#pragma unroll
  for (int r = 0; r < FF_TH / 4; ++r) {
    if (!in[r]) continue;
    for (int c = 0; c < q.cn; ++c) A[(size_t)c * n + e[r]] = acc[c][r];
    WS[e[r]] = ws[r];
  }
}

}  // namespace

long long fm_job_blocks(const FmSeq& q, int kind, int level) {
  const long long n0 = (long long)q.w * q.h;
  switch (kind) {
    case FM_PREP:
    case FM_LPG: return (long long)q.m * ((n0 + 255) / 256);
    case FM_BLUR: return (long long)q.m * ((q.w + FB_TW - 1) / FB_TW) * ((q.h + FB_TH - 1) / FB_TH);
    case FM_DOWN_G:
    case FM_DOWN_W: {
      const long long tiles = (long long)((q.lw[level + 1] + FD_TW - 1) / FD_TW) * ((q.lh[level + 1] + FD_TH - 1) / FD_TH);
      return tiles * (kind == FM_DOWN_G ? (long long)q.m * q.cn : q.m);
    }
    default: return (long long)((q.lw[level] + FF_TW - 1) / FF_TW) * ((q.lh[level] + FF_TH - 1) / FF_TH);
  }
}

hipError_t launch_fm_step(int kind, const FmJobs& J, const FmSeq* seqs, int blur_radius, hipStream_t s) {
  if (J.n <= 0) return hipSuccess;
  if (J.n > FM_MAXJOBS) return hipErrorInvalidValue;
  const dim3 grid((unsigned)J.j[J.n - 1].end);
  switch (kind) {
    case FM_PREP: hipLaunchKernelGGL(k_fm_prep, grid, dim3(256), 0, s, J, seqs); break;
    case FM_LPG: hipLaunchKernelGGL(k_fm_lpg, grid, dim3(256), 0, s, J, seqs); break;
    case FM_BLUR: {
      // one radius for every job of the launch (fm_blur_radius): the register-blocked form
      switch (blur_radius) {
#define SSG_FMB(RR) case RR: hipLaunchKernelGGL(k_fm_blur_r<RR>, grid, dim3(256), 0, s, J, seqs); break;
        SSG_FMB(1) SSG_FMB(2) SSG_FMB(3) SSG_FMB(4) SSG_FMB(5) SSG_FMB(6) SSG_FMB(7) SSG_FMB(8)
        SSG_FMB(9) SSG_FMB(10) SSG_FMB(11) SSG_FMB(12) SSG_FMB(13) SSG_FMB(14) SSG_FMB(15) SSG_FMB(16)
        SSG_FMB(17) SSG_FMB(18) SSG_FMB(19) SSG_FMB(20) SSG_FMB(21) SSG_FMB(22) SSG_FMB(23) SSG_FMB(24)
#undef SSG_FMB
        default: hipLaunchKernelGGL(k_fm_blur, grid, dim3(256), 0, s, J, seqs); break;
      }
      break;
    }
    case FM_DOWN_G:
    case FM_DOWN_W: hipLaunchKernelGGL(k_fm_down, grid, dim3(256), 0, s, J, seqs); break;
    default: hipLaunchKernelGGL(k_fm_fuse, grid, dim3(256), 0, s, J, seqs); break;
  }
  return hipGetLastError();
}

}  // namespace ssg

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
  for (int e = threadIdx.x; e < FD_SW * FD_SH; e += 256) {
    const int ly = e / FD_SW, lx = e - ly * FD_SW;
    const int bx = border_interp(sx0 + lx, sw, BORDER_REFLECT101), by = border_interp(sy0 + ly, sh, BORDER_REFLECT101);
    sm[ly][lx] = S[(size_t)by * sw + bx];
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
__global__ void __launch_bounds__(256) k_fm_fuse(FmJobs J, const FmSeq* __restrict__ seqs) {
  const int j = find_job(J, blockIdx.x);
  const FmJob jb = J.j[j];
  const FmSeq& q = seqs[jb.seq];
  const int i = jb.level;
  const bool top = jb.kind == FM_TOP;
  const long long n = plane_n(q, i);
  const long long e = (long long)(blockIdx.x - job_begin(J, j)) * 256 + threadIdx.x;
  if (e >= n) return;
  const int w = q.lw[i];
  const int y = (int)(e / w), x = (int)(e - (long long)y * w);
  float* __restrict__ A = q.acc + (size_t)q.cn * q.loff[i];
  float* __restrict__ WS = q.wsum + q.loff[i];
  const float* __restrict__ Gi = q.G + (size_t)q.m * q.cn * q.loff[i];
  const float* __restrict__ Wi = q.W + (size_t)q.m * q.loff[i];
  const float* __restrict__ Gn = top ? nullptr : q.G + (size_t)q.m * q.cn * q.loff[i + 1];
  const long long nn = top ? 0 : plane_n(q, i + 1);
  const int nw = top ? 0 : q.lw[i + 1], nh = top ? 0 : q.lh[i + 1];
  float a[3];
  for (int c = 0; c < q.cn; ++c) a[c] = A[(size_t)c * n + e];
  float ws = WS[e];
  for (int k = 0; k < q.m; ++k) {
    const float wk = Wi[(size_t)k * n + e];
    for (int c = 0; c < q.cn; ++c) {
      const int pl = k * q.cn + c;
      float L = Gi[(size_t)pl * n + e];
      if (!top) L = L - pyrup_at(Gn + (size_t)pl * nn, nw, nw, nh, q.lh[i], x, y);
      a[c] = a[c] + wk * L;
    }
    ws = ws + wk;
  }
  for (int c = 0; c < q.cn; ++c) A[(size_t)c * n + e] = a[c];
  WS[e] = ws;
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
    default: return ((long long)q.lw[level] * q.lh[level] + 255) / 256;
  }
}

hipError_t launch_fm_step(int kind, const FmJobs& J, const FmSeq* seqs, hipStream_t s) {
  if (J.n <= 0) return hipSuccess;
  if (J.n > FM_MAXJOBS) return hipErrorInvalidValue;
  const dim3 grid((unsigned)J.j[J.n - 1].end);
  switch (kind) {
    case FM_PREP: hipLaunchKernelGGL(k_fm_prep, grid, dim3(256), 0, s, J, seqs); break;
    case FM_LPG: hipLaunchKernelGGL(k_fm_lpg, grid, dim3(256), 0, s, J, seqs); break;
    case FM_BLUR: hipLaunchKernelGGL(k_fm_blur, grid, dim3(256), 0, s, J, seqs); break;
    case FM_DOWN_G:
    case FM_DOWN_W: hipLaunchKernelGGL(k_fm_down, grid, dim3(256), 0, s, J, seqs); break;
    default: hipLaunchKernelGGL(k_fm_fuse, grid, dim3(256), 0, s, J, seqs); break;
  }
  return hipGetLastError();
}

}  // namespace ssg

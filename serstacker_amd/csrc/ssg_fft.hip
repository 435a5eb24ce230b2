// ssg_fft.hip -- FFT phase-correlation translation estimate on gfx950.
//
// SerStacker has no phase correlation on this path (nearest code:
// core/proc/fft.cc:17-160); this is the pre-alignment the build defines:
// optional 2x2 binning, mean removal and separable Hann apodisation, zero
// padding to powers of two, 2-D FFT, normalised cross-power F conj(G) / |F conj(G)|
// with a gaussian weight (correlation peak of sigma 1 px), inverse FFT, integer
// peak and a 3-point gaussian refinement per axis in FP64.
//
// The 2-D FFT is row FFTs in LDS (radix-2 Stockham, natural order in and out,
// one row per block) with tiled transposes between the two directions:
//   spatial [N][M] --rows--> [N][M] --T--> [M][N] --rows--> spectrum [M][N]
// The inverse runs the same steps back.  Every array holds a batch of frames.
#include "ssg_engine.hpp"

namespace ssg {

namespace {

constexpr int FFT_THREADS = 256;

// one row of n = 2^log2n complex values per block; tw[k] = exp(-2 pi i k / n), k < n/2
template <bool INV>
__global__ void __launch_bounds__(FFT_THREADS) k_fft_rows(float2* data, int n, int log2n, const float2* tw,
                                                          float scale) {
  extern __shared__ float2 sbuf[];
  float2* a = sbuf;
  float2* b = sbuf + n;
  float2* row = data + (size_t)blockIdx.x * n;
  for (int i = threadIdx.x; i < n; i += FFT_THREADS) a[i] = row[i];
  __syncthreads();
  const int half = n >> 1;
  for (int s = 0, ns = 1; s < log2n; ++s, ns <<= 1) {
    for (int j = threadIdx.x; j < half; j += FFT_THREADS) {
      const int k = j & (ns - 1);  // j % ns
      const float2 w = tw[k << (log2n - 1 - s)];  // exp(-2 pi i k / (2 ns))
      const float wy = INV ? -w.y : w.y;
      const float2 u = a[j], v0 = a[j + half];
      const float2 v = make_float2(fmaf(v0.x, w.x, -(v0.y * wy)), fmaf(v0.x, wy, v0.y * w.x));
      const int d = ((j - k) << 1) + k;  // (j / ns) * 2 ns + j % ns
      b[d] = make_float2(u.x + v.x, u.y + v.y);
      b[d + ns] = make_float2(u.x - v.x, u.y - v.y);
    }
    __syncthreads();
    float2* t = a;
    a = b;
    b = t;
  }
  for (int i = threadIdx.x; i < n; i += FFT_THREADS) row[i] = make_float2(a[i].x * scale, a[i].y * scale);
}

// [rows][cols] -> [cols][rows] per frame (blockIdx.z), 32x32 tiles through LDS
__global__ void __launch_bounds__(256) k_transpose(const float2* in, float2* out, int rows, int cols) {
  __shared__ float2 t[32][33];
  const size_t fo = (size_t)blockIdx.z * rows * cols;
  const int x0 = blockIdx.x * 32, y0 = blockIdx.y * 32;
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
  for (int r = ty; r < 32; r += 8) {
    const int y = y0 + r, x = x0 + tx;
    if (y < rows && x < cols) t[r][tx] = in[fo + (size_t)y * cols + x];
  }
  __syncthreads();
  for (int r = ty; r < 32; r += 8) {
    const int y = x0 + r, x = y0 + tx;  // out row = in column
    if (y < cols && x < rows) out[fo + (size_t)y * rows + x] = t[tx][r];
  }
}

template <bool U16>
__device__ __forceinline__ float binned(const void* src, int pitch, int binning, int x, int y) {
  auto at = [&](int xx, int yy) {
    const size_t o = (size_t)yy * pitch + xx;
    return U16 ? (float)((const uint16_t*)src)[o] : ((const float*)src)[o];
  };
  if (binning == 2) {
    const int sx = 2 * x, sy = 2 * y;
    return ((at(sx, sy) + at(sx + 1, sy)) + (at(sx, sy + 1) + at(sx + 1, sy + 1))) * 0.25f;
  }
  return at(x, y);
}

// mean of the binned image of frame blockIdx.x (double accumulation)
constexpr int PC_PARTS = PC_PARTS_MAX;  // blocks per frame of the two-stage reductions

// stage 1: double sums of the binned frame over row bands, one partial per block (PC_PARTS per frame)
template <bool U16>
__global__ void __launch_bounds__(256) k_pc_mean_part(const void* const* frames, int pitch, int bw, int bh,
                                                      int binning, double* part) {
  __shared__ double red[256];
  const void* src = frames[blockIdx.y];
  const int y0 = (int)((long long)bh * blockIdx.x / PC_PARTS), y1 = (int)((long long)bh * (blockIdx.x + 1) / PC_PARTS);
  double s = 0;
  for (int y = y0 + (threadIdx.x >> 6); y < y1; y += 4)
    for (int x = threadIdx.x & 63; x < bw; x += 64) s += (double)binned<U16>(src, pitch, binning, x, y);
  red[threadIdx.x] = s;
  __syncthreads();
  for (int k = 128; k > 0; k >>= 1) {
    if (threadIdx.x < k) red[threadIdx.x] += red[threadIdx.x + k];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[(size_t)blockIdx.y * PC_PARTS + blockIdx.x] = red[0];
}

// stage 2: the partials of a frame in order, then the mean
__global__ void k_pc_mean_fin(const double* part, int bw, int bh, float* mean) {
  if (threadIdx.x != 0) return;
  double s = 0;
  for (int i = 0; i < PC_PARTS; ++i) s += part[(size_t)blockIdx.x * PC_PARTS + i];
  mean[blockIdx.x] = (float)(s / ((double)bw * bh));
}

// binning, mean removal, apodisation and zero padding into the complex buffer [N][M]
// of frame blockIdx.z
template <bool U16>
__global__ void __launch_bounds__(256) k_pc_load(const void* const* frames, int pitch, int bw, int bh, int binning,
                                                 const float* mean, const float* wx, const float* wy, float2* out,
                                                 int M, int N) {
  const int x = blockIdx.x * 64 + (threadIdx.x & 63), y = blockIdx.y * 4 + (threadIdx.x >> 6);
  if (x >= M || y >= N) return;
  float v = 0.f;
  if (x < bw && y < bh)
    v = (binned<U16>(frames[blockIdx.z], pitch, binning, x, y) - mean[blockIdx.z]) * (wx[x] * wy[y]);
  out[(size_t)blockIdx.z * M * N + (size_t)y * M + x] = make_float2(v, 0.f);
}

// R = F conj(G) / |F conj(G)| against the reference spectrum (same layout, [M][N]),
// times the gaussian weight exp(-2 pi^2 s^2 (kx^2 + ky^2)) that makes the
// correlation peak a gaussian of sigma s = 1 px
__global__ void __launch_bounds__(256) k_pc_cross(float2* spec, const float2* ref, int M, int N) {
  const long long per_frame = (long long)M * N;
  const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
  if (i >= per_frame) return;
  float2* s = spec + (size_t)blockIdx.y * per_frame + i;
  const float2 f = *s, g = ref[i];
  const float re = fmaf(f.x, g.x, f.y * g.y), im = fmaf(f.y, g.x, -(f.x * g.y));
  const float mag = sqrtf(fmaf(re, re, im * im));
  const int ix = (int)(i / N), iy = (int)(i - (long long)ix * N);
  const float kx = (float)(ix < M / 2 ? ix : ix - M) / (float)M, ky = (float)(iy < N / 2 ? iy : iy - N) / (float)N;
  const float wgt = expf(-2.f * 9.8696044f * fmaf(kx, kx, ky * ky));  // 2 pi^2 s^2, s = 1
  *s = mag > 0.f ? make_float2(re / mag * wgt, im / mag * wgt) : make_float2(0.f, 0.f);
}

// stage 1: argmax of the real part over a band of the frame (first index on ties)
__global__ void __launch_bounds__(256) k_pc_peak_part(const float2* r, int M, int N, float* pv, int* pi) {
  __shared__ float sv[256];
  __shared__ int si[256];
  const float2* p = r + (size_t)blockIdx.y * M * N;
  const int total = M * N;
  const int i0 = (int)((long long)total * blockIdx.x / PC_PARTS), i1 = (int)((long long)total * (blockIdx.x + 1) / PC_PARTS);
  float best = -3.4e38f;
  int bi = i0;
  for (int i = i0 + threadIdx.x; i < i1; i += 256) {
    const float v = p[i].x;
    if (v > best) {
      best = v;
      bi = i;
    }
  }
  sv[threadIdx.x] = best;
  si[threadIdx.x] = bi;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s) {
      const float v = sv[threadIdx.x + s];
      const int j = si[threadIdx.x + s];
      if (v > sv[threadIdx.x] || (v == sv[threadIdx.x] && j < si[threadIdx.x])) {
        sv[threadIdx.x] = v;
        si[threadIdx.x] = j;
      }
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    pv[(size_t)blockIdx.y * PC_PARTS + blockIdx.x] = sv[0];
    pi[(size_t)blockIdx.y * PC_PARTS + blockIdx.x] = si[0];
  }
}

// stage 2, per frame: the band maxima (first index on ties), then the 3-point fits in FP64
__global__ void k_pc_peak(const float2* r, int M, int N, int binning, const float* pv, const int* pi,
                          double* out) {
  const float2* p = r + (size_t)blockIdx.x * M * N;
  if (threadIdx.x == 0) {
    float best = pv[(size_t)blockIdx.x * PC_PARTS];
    int bi = pi[(size_t)blockIdx.x * PC_PARTS];
    for (int b = 1; b < PC_PARTS; ++b) {
      const float v = pv[(size_t)blockIdx.x * PC_PARTS + b];
      const int j = pi[(size_t)blockIdx.x * PC_PARTS + b];
      if (v > best || (v == best && j < bi)) {
        best = v;
        bi = j;
      }
    }
    const int i = bi, py = i / M, px = i - py * M;
    const double r0 = p[i].x;
    const double rxm = p[(size_t)py * M + (px + M - 1) % M].x, rxp = p[(size_t)py * M + (px + 1) % M].x;
    const double rym = p[(size_t)((py + N - 1) % N) * M + px].x, ryp = p[(size_t)((py + 1) % N) * M + px].x;
    // gaussian fit: a parabola through the logarithms (a plain parabola if a value is not positive)
    auto fit = [](double a, double b, double c) {
      if (a > 0 && b > 0 && c > 0) {
        a = log(a);
        b = log(b);
        c = log(c);
      }
      const double d = a - 2.0 * b + c;
      return d == 0.0 ? 0.0 : 0.5 * (a - c) / d;
    };
    const double cx = fit(rxm, r0, rxp), cy = fit(rym, r0, ryp);
    const int sx = px > M / 2 ? px - M : px, sy = py > N / 2 ? py - N : py;
    out[3 * blockIdx.x + 0] = ((double)sx + cx) * binning;
    out[3 * blockIdx.x + 1] = ((double)sy + cy) * binning;
    out[3 * blockIdx.x + 2] = r0;
  }
}


int ilog2(int n) {
  int l = 0;
  while ((1 << l) < n) ++l;
  return l;
}

}  // namespace

// one direction of a batched 2-D FFT.  forward: spatial buf [N][M] -> spectrum in
// tmp [M][N]; inverse: spectrum buf [M][N] -> spatial tmp [N][M] (scaled 1/(M N)).
hipError_t launch_fft2(float2* buf, float2* tmp, int M, int N, int nframes, const float2* tw_m, const float2* tw_n,
                       bool inverse, hipStream_t s) {
  const int lm = ilog2(M), ln = ilog2(N);
  const dim3 tgrid_mn((M + 31) / 32, (N + 31) / 32, nframes), tgrid_nm((N + 31) / 32, (M + 31) / 32, nframes);
  if (!inverse) {
    hipLaunchKernelGGL(k_fft_rows<false>, dim3(N * nframes), dim3(FFT_THREADS), 2 * M * sizeof(float2), s, buf, M, lm,
                       tw_m, 1.f);
    hipLaunchKernelGGL(k_transpose, tgrid_mn, dim3(256), 0, s, buf, tmp, N, M);
    hipLaunchKernelGGL(k_fft_rows<false>, dim3(M * nframes), dim3(FFT_THREADS), 2 * N * sizeof(float2), s, tmp, N, ln,
                       tw_n, 1.f);
  } else {
    hipLaunchKernelGGL(k_fft_rows<true>, dim3(M * nframes), dim3(FFT_THREADS), 2 * N * sizeof(float2), s, buf, N, ln,
                       tw_n, 1.f / (float)N);
    hipLaunchKernelGGL(k_transpose, tgrid_nm, dim3(256), 0, s, buf, tmp, M, N);
    hipLaunchKernelGGL(k_fft_rows<true>, dim3(N * nframes), dim3(FFT_THREADS), 2 * M * sizeof(float2), s, tmp, M, lm,
                       tw_m, 1.f / (float)M);
  }
  return hipGetLastError();
}

hipError_t launch_pc_load(const void* const* frames, bool u16, int pitch, int bw, int bh, int binning, float* mean,
                          double* part, const float* wx, const float* wy, float2* out, int M, int N, int nframes,
                          hipStream_t s) {
  dim3 grid((M + 63) / 64, (N + 3) / 4, nframes);
  if (u16) {
    hipLaunchKernelGGL(k_pc_mean_part<true>, dim3(PC_PARTS, nframes), dim3(256), 0, s, frames, pitch, bw, bh, binning,
                       part);
    hipLaunchKernelGGL(k_pc_mean_fin, dim3(nframes), dim3(64), 0, s, part, bw, bh, mean);
    hipLaunchKernelGGL(k_pc_load<true>, grid, dim3(256), 0, s, frames, pitch, bw, bh, binning, mean, wx, wy, out, M, N);
  } else {
    hipLaunchKernelGGL(k_pc_mean_part<false>, dim3(PC_PARTS, nframes), dim3(256), 0, s, frames, pitch, bw, bh, binning,
                       part);
    hipLaunchKernelGGL(k_pc_mean_fin, dim3(nframes), dim3(64), 0, s, part, bw, bh, mean);
    hipLaunchKernelGGL(k_pc_load<false>, grid, dim3(256), 0, s, frames, pitch, bw, bh, binning, mean, wx, wy, out, M, N);
  }
  return hipGetLastError();
}

hipError_t launch_pc_cross(float2* spec, const float2* ref, int M, int N, int nframes, hipStream_t s) {
  dim3 grid((unsigned)(((long long)M * N + 255) / 256), nframes);
  hipLaunchKernelGGL(k_pc_cross, grid, dim3(256), 0, s, spec, ref, M, N);
  return hipGetLastError();
}

hipError_t launch_pc_peak(const float2* r, int M, int N, int binning, int nframes, float* pv, int* pi, double* out,
                          hipStream_t s) {
  hipLaunchKernelGGL(k_pc_peak_part, dim3(PC_PARTS, nframes), dim3(256), 0, s, r, M, N, pv, pi);
  hipLaunchKernelGGL(k_pc_peak, dim3(nframes), dim3(64), 0, s, r, M, N, binning, pv, pi, out);
  return hipGetLastError();
}

}  // namespace ssg

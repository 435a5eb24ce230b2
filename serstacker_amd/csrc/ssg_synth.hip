// ssg_synth.hip -- seeded synthetic planetary-disk frames (benchmark and test
// inputs).  Counter-based Philox4x32-10 (Salmon et al., SC'11) keyed by
// (seed, frame, stream) makes every frame reproducible independently of the
// batch it is rendered in.
//
// Scene: a disk of radius R at the image centre with limb darkening
// I = 1 - u (1 - mu), mu = sqrt(1 - r^2), textured by (1 + a * t) where t is
// uniform noise U(-1,1) filtered by a gaussian (sigma = texture_sigma) and
// normalised to unit peak; dark background.  Frame k samples the scene at
// T_k(x) = C + R(theta) [[s, h], [0, s]] (x - C) + t (bicubic texture
// lookup), is blurred by a gaussian of its own sigma, and gets Poisson photon
// noise (peak electrons at I = 1, Knuth's method below 30 e-, a normal
// approximation above) plus gaussian read noise.  The reference frame
// (index -1) uses T = identity and the mean blur.
#include <math.h>
#include <string.h>

#include <vector>

#include "../../include/ssg.h"
#include "ssg_engine.hpp"

namespace ssg {

struct Philox {
  uint32_t c[4];
};

__host__ __device__ inline void philox4x32_10(uint32_t ctr[4], uint32_t k0, uint32_t k1) {
  const uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u, W0 = 0x9E3779B9u, W1 = 0xBB67AE85u;
  for (int r = 0; r < 10; ++r) {
    const uint64_t p0 = (uint64_t)M0 * ctr[0], p1 = (uint64_t)M1 * ctr[2];
    const uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0, hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
    const uint32_t n0 = hi1 ^ ctr[1] ^ k0, n2 = hi0 ^ ctr[3] ^ k1;
    ctr[0] = n0;
    ctr[1] = lo1;
    ctr[2] = n2;
    ctr[3] = lo0;
    k0 += W0;
    k1 += W1;
  }
}

__host__ __device__ inline float u01(uint32_t x) { return ((x >> 8) + 0.5f) * (1.f / 16777216.f); }

// ---- texture --------------------------------------------------------------------------------
__global__ void k_tex_noise(float* tex, int w, int h, uint32_t k0, uint32_t k1) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y;
  if (x >= w) return;
  uint32_t c[4] = {(uint32_t)x, (uint32_t)y, 0x7e77u, 0};
  philox4x32_10(c, k0, k1);
  tex[(size_t)y * w + x] = 2.f * u01(c[0]) - 1.f;
}

// separable gaussian (per image sigma), BORDER_REPLICATE; one pass per axis
__global__ void k_gauss_axis(const float* __restrict__ src, float* __restrict__ dst, int w, int h, long long fstride,
                             const float* __restrict__ sigmas, float sigma_const, int axis) {
  const int f = blockIdx.z;
  const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y;
  if (x >= w) return;
  const float sg = sigmas ? sigmas[f] : sigma_const;
  const float* S = src + f * fstride;
  float* D = dst + f * fstride;
  if (!(sg > 0.f)) {
    D[(size_t)y * w + x] = S[(size_t)y * w + x];
    return;
  }
  const int rad = min(24, (int)ceilf(3.f * sg));
  const float a = -0.5f / (sg * sg);
  float acc = 0.f, ws = 0.f;
  for (int i = -rad; i <= rad; ++i) {
    const float k = __expf(a * (float)(i * i));
    const int xx = axis == 0 ? min(max(x + i, 0), w - 1) : x;
    const int yy = axis == 1 ? min(max(y + i, 0), h - 1) : y;
    acc = fmaf(k, S[(size_t)yy * w + xx], acc);
    ws += k;
  }
  D[(size_t)y * w + x] = acc / ws;
}

__global__ void k_absmax(const float* __restrict__ v, long long n, unsigned* out) {
  float m = 0.f;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
    m = fmaxf(m, fabsf(v[i]));
  for (int o = 32; o > 0; o >>= 1) m = fmaxf(m, __shfl_xor(m, o, 64));
  if ((threadIdx.x & 63) == 0) atomicMax(out, __float_as_uint(m));
}

// ---- render ----------------------------------------------------------------------------------
__device__ float cubic_w(float t, int i) {
  const float A = -0.75f;
  const float x = fabsf(t - (float)(i - 1));
  if (x <= 1.f) return ((A + 2.f) * x - (A + 3.f)) * x * x + 1.f;
  if (x < 2.f) return ((A * x - 5.f * A) * x + 8.f * A) * x - 4.f * A;
  return 0.f;
}

__global__ void k_render(const float* __restrict__ tex, int tw, int th, float margin, float tex_norm,
                         const SynthFrame* __restrict__ fr, int w, int h, double cx, double cy, double radius,
                         double limb, double amp, float* __restrict__ out, long long fstride) {
  const int f = blockIdx.z;
  const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y;
  if (x >= w) return;
  const double* A = fr[f].A;
  const double u = A[0] * x + A[1] * y + A[2], v = A[3] * x + A[4] * y + A[5];
  const double r = sqrt((u - cx) * (u - cx) + (v - cy) * (v - cy)) / radius;
  float I = 0.f;
  if (r < 1.0) {
    const double mu = sqrt(1.0 - r * r);
    I = (float)(1.0 - limb * (1.0 - mu));
    // bicubic texture sample
    const float tu = (float)u + margin, tv = (float)v + margin;
    const int iu = (int)floorf(tu), iv = (int)floorf(tv);
    const float fu = tu - iu, fv = tv - iv;
    float t = 0.f;
    for (int j = 0; j < 4; ++j) {
      const int yy = min(max(iv - 1 + j, 0), th - 1);
      float row = 0.f;
      for (int i = 0; i < 4; ++i) {
        const int xx = min(max(iu - 1 + i, 0), tw - 1);
        row = fmaf(cubic_w(fu, i), tex[(size_t)yy * tw + xx], row);
      }
      t = fmaf(cubic_w(fv, j), row, t);
    }
    I *= 1.f + (float)amp * t * tex_norm;
  }
  out[f * fstride + (size_t)y * w + x] = I;
}

__device__ float poisson(float lam, uint32_t c[4], uint32_t k0, uint32_t k1) {
  if (!(lam > 0.f)) return 0.f;
  if (lam < 30.f) {
    const float L = __expf(-lam);
    float p = 1.f;
    int k = 0;
    int used = 4;
    uint32_t r[4] = {0, 0, 0, 0};
    for (int it = 0; it < 256; ++it) {  // bounded: P(k > 256) is nil for lam < 30
      if (used == 4) {
        c[3] += 1;
        r[0] = c[0]; r[1] = c[1]; r[2] = c[2]; r[3] = c[3];
        philox4x32_10(r, k0, k1);
        used = 0;
      }
      p *= u01(r[used++]);
      if (p <= L) break;
      ++k;
    }
    return (float)k;
  }
  uint32_t r[4] = {c[0], c[1], c[2], c[3] + 1000u};
  philox4x32_10(r, k0, k1);
  const float z = sqrtf(-2.f * __logf(u01(r[0]))) * __cosf(6.2831853f * u01(r[1]));
  return fmaxf(0.f, rintf(lam + sqrtf(lam) * z));
}

__global__ void k_noise_out(const float* __restrict__ img, long long fstride, int w, int h,
                            const SynthFrame* __restrict__ fr, double peak, double read_noise, double out_scale,
                            uint32_t k0, void* const* __restrict__ out_ptrs, int out_pitch_bytes, int out_u16) {
  const int f = blockIdx.z;
  const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y;
  if (x >= w) return;
  const float I = img[f * fstride + (size_t)y * w + x];
  float dn = I * (float)peak;
  if (peak > 0) {
    uint32_t c[4] = {(uint32_t)x, (uint32_t)y, fr[f].key, 0u};
    dn = poisson(I * (float)peak, c, k0, 0x5eedu);
    if (read_noise > 0) {
      uint32_t r[4] = {(uint32_t)x, (uint32_t)y, fr[f].key, 0xfffff000u};
      philox4x32_10(r, k0, 0x7ead);
      const float z = sqrtf(-2.f * __logf(u01(r[0]))) * __cosf(6.2831853f * u01(r[1]));
      dn += (float)read_noise * z;
    }
  }
  char* row = (char*)out_ptrs[f] + (size_t)y * out_pitch_bytes;
  if (out_u16) {
    const float v = fminf(fmaxf(rintf(dn), 0.f), 65535.f);
    ((uint16_t*)row)[x] = (uint16_t)v;
  } else {
    ((float*)row)[x] = (float)(dn * out_scale);
  }
}

}  // namespace ssg

using namespace ssg;

namespace {

// T = C + R(theta) [[s, h], [0, s]] (x - C) + t, returned as scene = A (x, y, 1)
void frame_transform(const ssg_synth_params* p, int index, double A[6], float* blur, uint32_t* key) {
  const double cx = 0.5 * (p->width - 1), cy = 0.5 * (p->height - 1);
  double tx = 0, ty = 0, th = 0, s = 1, hsh = 0;
  double sg = 0.5 * (p->blur_sigma_min + p->blur_sigma_max);
  if (index >= 0) {
    uint32_t c[4] = {(uint32_t)index, 0xf00du, 0, 0};
    philox4x32_10(c, (uint32_t)p->seed, (uint32_t)(p->seed >> 32));
    uint32_t d[4] = {(uint32_t)index, 0xbeefu, 0, 0};
    philox4x32_10(d, (uint32_t)p->seed, (uint32_t)(p->seed >> 32));
    auto U = [](uint32_t v, double a) { return (2.0 * u01(v) - 1.0) * a; };
    tx = U(c[0], p->max_shift);
    ty = U(c[1], p->max_shift);
    th = U(c[2], p->max_rotation_deg) * M_PI / 180.0;
    s = 1.0 + U(c[3], p->max_scale_dev);
    hsh = U(d[0], p->max_shear);
    sg = p->blur_sigma_min + u01(d[1]) * (p->blur_sigma_max - p->blur_sigma_min);
  }
  const double ca = cos(th), sa = sin(th);
  // M = R [[s, h], [0, s]]
  const double m00 = ca * s, m01 = ca * hsh - sa * s, m10 = sa * s, m11 = sa * hsh + ca * s;
  A[0] = m00;
  A[1] = m01;
  A[2] = cx + tx - (m00 * cx + m01 * cy);
  A[3] = m10;
  A[4] = m11;
  A[5] = cy + ty - (m10 * cx + m11 * cy);
  *blur = (float)sg;
  *key = (uint32_t)(index + 1);
}

struct SynthCache {
  ssg_synth_params p{};
  float* tex = nullptr;
  int tw = 0, th = 0;
  float norm = 1;
  bool valid = false;
};
SynthCache g_synth;

}  // namespace

extern "C" int ssg_synth_render(const ssg_synth_params* p, int first, int n, const ssg_image* frames, double* truth) {
  if (!p || (n > 0 && !frames)) return SSG_E_INVALID_ARG;
  if (p->width <= 0 || p->height <= 0) return SSG_E_INVALID_ARG;
  for (int i = 0; i < n; ++i) {
    const ssg_image& im = frames[i];
    if (!im.on_device || im.width != p->width || im.height != p->height) return SSG_E_INVALID_ARG;
    if (im.type != SSG_32FC1 && im.type != SSG_16UC1) return SSG_E_INVALID_ARG;
    if (im.type != frames[0].type || im.step != frames[0].step) return SSG_E_INVALID_ARG;
  }
  hipStream_t s = 0;
  const int margin = 64;
  // texture (cached per parameter set)
  if (!g_synth.valid || memcmp(&g_synth.p, p, sizeof(*p)) != 0) {
    if (g_synth.tex) (void)hipFree(g_synth.tex);
    g_synth.tex = nullptr;
    g_synth.valid = false;
    const int tw = p->width + 2 * margin, th = p->height + 2 * margin;
    float *tex = nullptr, *tmp = nullptr;
    unsigned* mx = nullptr;
    if (hipMalloc(&tex, sizeof(float) * tw * th) != hipSuccess) return SSG_E_NO_MEMORY;
    if (hipMalloc(&tmp, sizeof(float) * tw * th) != hipSuccess) { (void)hipFree(tex); return SSG_E_NO_MEMORY; }
    if (hipMalloc(&mx, sizeof(unsigned)) != hipSuccess) { (void)hipFree(tex); (void)hipFree(tmp); return SSG_E_NO_MEMORY; }
    dim3 g((tw + 255) / 256, th);
    hipLaunchKernelGGL(k_tex_noise, g, dim3(256), 0, s, tex, tw, th, (uint32_t)p->seed ^ 0x1234567u, (uint32_t)(p->seed >> 32));
    const float sig = (float)p->texture_sigma;
    hipLaunchKernelGGL(k_gauss_axis, dim3((tw + 255) / 256, th, 1), dim3(256), 0, s, tex, tmp, tw, th, 0LL, nullptr, sig, 0);
    hipLaunchKernelGGL(k_gauss_axis, dim3((tw + 255) / 256, th, 1), dim3(256), 0, s, tmp, tex, tw, th, 0LL, nullptr, sig, 1);
    (void)hipMemsetAsync(mx, 0, sizeof(unsigned), s);
    hipLaunchKernelGGL(k_absmax, dim3(512), dim3(256), 0, s, tex, (long long)tw * th, mx);
    unsigned bits = 0;
    (void)hipMemcpyAsync(&bits, mx, sizeof(unsigned), hipMemcpyDeviceToHost, s);
    if (hipStreamSynchronize(s) != hipSuccess) return SSG_E_HIP;
    float m;
    memcpy(&m, &bits, 4);
    (void)hipFree(tmp);
    (void)hipFree(mx);
    g_synth.p = *p;
    g_synth.tex = tex;
    g_synth.tw = tw;
    g_synth.th = th;
    g_synth.norm = m > 0 ? 1.f / m : 1.f;
    g_synth.valid = true;
  }
  if (n <= 0) return SSG_OK;
  const int chunk = std::max(1, std::min(n, (int)(512LL * 1024 * 1024 / ((long long)p->width * p->height * 8))));
  float *b0 = nullptr, *b1 = nullptr;
  SynthFrame* dfr = nullptr;
  void** dptr = nullptr;
  const long long fs = (long long)p->width * p->height;
  if (hipMalloc(&b0, sizeof(float) * fs * chunk) != hipSuccess) return SSG_E_NO_MEMORY;
  if (hipMalloc(&b1, sizeof(float) * fs * chunk) != hipSuccess) { (void)hipFree(b0); return SSG_E_NO_MEMORY; }
  (void)hipMalloc(&dfr, sizeof(SynthFrame) * chunk);
  (void)hipMalloc(&dptr, sizeof(void*) * chunk);
  std::vector<SynthFrame> fr(chunk);
  std::vector<float> sig(chunk);
  std::vector<void*> ptrs(chunk);
  float* dsig = nullptr;
  (void)hipMalloc(&dsig, sizeof(float) * chunk);
  int rc = SSG_OK;
  for (int c0 = 0; c0 < n && rc == SSG_OK; c0 += chunk) {
    const int m = std::min(chunk, n - c0);
    for (int i = 0; i < m; ++i) {
      const int idx = first < 0 ? -1 : first + c0 + i;
      float blur;
      uint32_t key;
      double A[6];
      frame_transform(p, idx, A, &blur, &key);
      for (int k = 0; k < 6; ++k) fr[i].A[k] = A[k];
      fr[i].blur_sigma = blur;
      fr[i].key = key;
      sig[i] = blur;
      ptrs[i] = frames[c0 + i].data;
      if (truth) {
        // registration truth: the inverse of the frame -> scene map
        const double det = A[0] * A[4] - A[1] * A[3];
        double* t = truth + 6 * (c0 + i);
        t[0] = A[4] / det;
        t[1] = -A[1] / det;
        t[3] = -A[3] / det;
        t[4] = A[0] / det;
        t[2] = -(t[0] * A[2] + t[1] * A[5]);
        t[5] = -(t[3] * A[2] + t[4] * A[5]);
      }
    }
    (void)hipMemcpyAsync(dfr, fr.data(), sizeof(SynthFrame) * m, hipMemcpyHostToDevice, s);
    (void)hipMemcpyAsync(dsig, sig.data(), sizeof(float) * m, hipMemcpyHostToDevice, s);
    (void)hipMemcpyAsync(dptr, ptrs.data(), sizeof(void*) * m, hipMemcpyHostToDevice, s);
    dim3 g((p->width + 255) / 256, p->height, m);
    const double cx = 0.5 * (p->width - 1), cy = 0.5 * (p->height - 1);
    hipLaunchKernelGGL(k_render, g, dim3(256), 0, s, g_synth.tex, g_synth.tw, g_synth.th, (float)margin, g_synth.norm, dfr,
                       p->width, p->height, cx, cy, p->disk_radius, p->limb_darkening, p->texture_amplitude, b0, fs);
    hipLaunchKernelGGL(k_gauss_axis, g, dim3(256), 0, s, b0, b1, p->width, p->height, fs, dsig, 0.f, 0);
    hipLaunchKernelGGL(k_gauss_axis, g, dim3(256), 0, s, b1, b0, p->width, p->height, fs, dsig, 0.f, 1);
    hipLaunchKernelGGL(k_noise_out, g, dim3(256), 0, s, b0, fs, p->width, p->height, dfr, p->peak_electrons,
                       p->read_noise, p->output_scale, (uint32_t)p->seed, dptr, (int)frames[0].step,
                       frames[0].type == SSG_16UC1 ? 1 : 0);
    if (hipStreamSynchronize(s) != hipSuccess) rc = SSG_E_HIP;
  }
  (void)hipFree(b0);
  (void)hipFree(b1);
  (void)hipFree(dfr);
  (void)hipFree(dptr);
  (void)hipFree(dsig);
  return rc;
}

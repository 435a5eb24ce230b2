// ssg_api.hip -- host engine and C ABI (include/ssg.h).
//
// The registration handle mirrors c_frame_registration (ECC path); the
// accumulation handle mirrors c_frame_accumulation (c_weigthed_average).
// Both own one HIP stream (default priority) and their device buffers.
#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <atomic>
#include <mutex>
#include <string>
#include <cstring>
#include <vector>

#include <cmath>
#include <memory>

#include "../../include/ssg.h"
#include "ssg_engine.hpp"

using namespace ssg;

namespace {

thread_local std::string g_err;

int fail(int code, const char* fmt, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_err = buf;
  return code;
}

#define HIPCHK(x)                                                                         \
  do {                                                                                    \
    hipError_t e_ = (x);                                                                  \
    if (e_ != hipSuccess) return fail(SSG_E_HIP, "%s: %s", #x, hipGetErrorString(e_));    \
  } while (0)

#define CHK(x)             \
  do {                     \
    int r_ = (x);          \
    if (r_ != SSG_OK) return r_; \
  } while (0)

struct DevBuf {
  void* p = nullptr;
  size_t n = 0;
  DevBuf() = default;
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  ~DevBuf() { release(); }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
  }
  int ensure(size_t bytes) {
    if (bytes <= n) return SSG_OK;
    release();
    if (!bytes) return SSG_OK;
    hipError_t e = hipMalloc(&p, bytes);
    if (e != hipSuccess) {
      p = nullptr;
      return fail(SSG_E_NO_MEMORY, "hipMalloc(%zu bytes): %s", bytes, hipGetErrorString(e));
    }
    n = bytes;
    return SSG_OK;
  }
  template <class T>
  T* as() const {
    return reinterpret_cast<T*>(p);
  }
};

int elem_size(int type) {
  switch (type) {
    case SSG_8UC1: return 1;
    case SSG_16UC1: return 2;
    case SSG_32SC1: return 4;
    case SSG_32FC1: return 4;
    case SSG_32FC2: return 8;
    case SSG_32FC3: return 12;
    case SSG_32SC3: return 12;
    case SSG_64FC1: return 8;
    case SSG_8UC3: return 3;
    case SSG_16UC3: return 6;
    case SSG_64FC3: return 24;
  }
  return 0;
}

int channels_of(int type) { return type == SSG_32FC3 ? 3 : (type == SSG_32FC2 ? 2 : 1); }

int check_image(const ssg_image* im, const char* what) {
  if (!im || !im->data) return fail(SSG_E_INVALID_ARG, "%s: null image", what);
  const int es = elem_size(im->type);
  if (!es) return fail(SSG_E_INVALID_ARG, "%s: unsupported type %d", what, im->type);
  if (im->width <= 0 || im->height <= 0) return fail(SSG_E_INVALID_ARG, "%s: empty image %dx%d", what, im->width, im->height);
  if (im->step < (int64_t)im->width * es) return fail(SSG_E_INVALID_ARG, "%s: step %lld < row bytes", what, (long long)im->step);
  return SSG_OK;
}

// A device view of an image: device images are used in place, host images are
// copied to tmp.  Returns the device pointer and the row pitch in bytes.
int device_view(const ssg_image* im, DevBuf& tmp, hipStream_t s, const void** dptr, size_t* pitch_bytes) {
  if (im->on_device) {
    *dptr = im->data;
    *pitch_bytes = (size_t)im->step;
    return SSG_OK;
  }
  const size_t row = (size_t)im->width * elem_size(im->type);
  const size_t pitch = (row + 255) / 256 * 256;
  CHK(tmp.ensure(pitch * im->height));
  HIPCHK(hipMemcpy2DAsync(tmp.p, pitch, im->data, (size_t)im->step, row, im->height, hipMemcpyHostToDevice, s));
  *dptr = tmp.p;
  *pitch_bytes = pitch;
  return SSG_OK;
}

// Cubic tables (OpenCV interpolateCubic at t = i/32 and BicubicTab_i), per device
std::mutex g_tab_mu;
std::vector<int> g_tab_done;

void cubic_coeffs(float x, float* c) {
  const float A = -0.75f;
  c[0] = ((A * (x + 1) - 5 * A) * (x + 1) + 8 * A) * (x + 1) - 4 * A;
  c[1] = ((A + 2) * x - (A + 3)) * x * x + 1;
  c[2] = ((A + 2) * (1 - x) - (A + 3)) * (1 - x) * (1 - x) + 1;
  c[3] = 1.f - c[0] - c[1] - c[2];
}

int ensure_tables(hipStream_t s) {
  int dev = 0;
  HIPCHK(hipGetDevice(&dev));
  std::lock_guard<std::mutex> lk(g_tab_mu);
  if ((int)g_tab_done.size() > dev && g_tab_done[dev]) return SSG_OK;
  float tab[32][4];
  for (int i = 0; i < 32; ++i) cubic_coeffs(i * (1.f / 32), tab[i]);
  static short itab[32][32][16];
  for (int fy = 0; fy < 32; ++fy)
    for (int fx = 0; fx < 32; ++fx) {
      int w[16], isum = 0;
      for (int k1 = 0; k1 < 4; ++k1)
        for (int k2 = 0; k2 < 4; ++k2) {
          const float v = tab[fy][k1] * tab[fx][k2];
          long r = lrintf(v * 32768.f);
          const int iv = r < -32768 ? -32768 : (r > 32767 ? 32767 : (int)r);
          w[k1 * 4 + k2] = iv;
          isum += iv;
        }
      if (isum != 32768) {
        int diff = isum - 32768, Mk1 = 2, Mk2 = 2, mk1 = 2, mk2 = 2;
        for (int k1 = 2; k1 < 4; ++k1)
          for (int k2 = 2; k2 < 4; ++k2) {
            if (w[k1 * 4 + k2] < w[mk1 * 4 + mk2]) { mk1 = k1; mk2 = k2; }
            else if (w[k1 * 4 + k2] > w[Mk1 * 4 + Mk2]) { Mk1 = k1; Mk2 = k2; }
          }
        if (diff < 0) w[Mk1 * 4 + Mk2] -= diff;
        else w[mk1 * 4 + mk2] -= diff;
      }
      for (int k = 0; k < 16; ++k) itab[fy][fx][k] = (short)w[k];
    }
  set_cubic_table(tab, s);
  HIPCHK(set_cubic_itab(&itab[0][0][0], s));
  HIPCHK(hipStreamSynchronize(s));
  if ((int)g_tab_done.size() <= dev) g_tab_done.resize(dev + 1, 0);
  g_tab_done[dev] = 1;
  return SSG_OK;
}

int motion_nparams(int m) {
  return m == SSG_MOTION_TRANSLATION ? 2 : (m == SSG_MOTION_EUCLIDEAN ? 3 : (m == SSG_MOTION_SCALED_EUCLIDEAN ? 4 : 6));
}

// host copy of the device map-coefficient rule (ssg_ecc.hip update_map)
MapCoef map_coef(int motion, const float* p, const float* C) {
  MapCoef m{};
  if (motion == SSG_MOTION_TRANSLATION) {
    m.c[0] = p[0];
    m.c[1] = p[1];
  } else if (motion == SSG_MOTION_AFFINE) {
    for (int i = 0; i < 6; ++i) m.c[i] = p[i];
  } else {
    m.c[0] = (float)sin((double)p[2]);
    m.c[1] = (float)cos((double)p[2]);
    m.c[2] = motion == SSG_MOTION_SCALED_EUCLIDEAN ? p[3] : 1.f;
    m.c[3] = p[0];
    m.c[4] = p[1];
    m.c[5] = C ? C[0] : 0.f;
    m.c[6] = C ? C[1] : 0.f;
  }
  return m;
}

void params_matrix(int motion, const float* p, double* m) {
  if (motion == SSG_MOTION_TRANSLATION) {
    m[0] = 1; m[1] = 0; m[2] = p[0];
    m[3] = 0; m[4] = 1; m[5] = p[1];
  } else if (motion == SSG_MOTION_AFFINE) {
    for (int i = 0; i < 6; ++i) m[i] = p[i];
  } else {
    const double s = motion == SSG_MOTION_SCALED_EUCLIDEAN ? p[3] : 1.0;
    const double sa = (float)sin((double)p[2]), ca = (float)cos((double)p[2]);
    m[0] = s * ca; m[1] = -s * sa; m[2] = p[0];
    m[3] = s * sa; m[4] = s * ca; m[5] = p[1];
  }
}

// c_image_transform::scale_transfrom on the host (float parameters, double factor)
void scale_params(int motion, float* p, double f) {
  if (motion == SSG_MOTION_AFFINE) {
    p[2] = (float)((double)p[2] * f);
    p[5] = (float)((double)p[5] * f);
  } else {
    p[0] = (float)((double)p[0] * f);
    p[1] = (float)((double)p[1] * f);
  }
}

SepKernel gaussian_sep(double sigma) {
  // c_ecch smoothing kernel: ksize = max(3, 2*int(3*sigma)+1), cv::getGaussianKernel(ksize, sigma)
  SepKernel k{};
  int n = 2 * (int)(3 * sigma) + 1;
  if (n < 3) n = 3;
  if (n > 31) n = 31;
  double cd[32], sum = 0;
  const double scale2X = -0.5 / (sigma * sigma);
  for (int i = 0; i < n; ++i) {
    const double x = i - (n - 1) * 0.5;
    cd[i] = exp(scale2X * x * x);
    sum += cd[i];
  }
  sum = 1. / sum;
  for (int i = 0; i < n; ++i) k.kx[i] = k.ky[i] = (float)(cd[i] * sum);
  k.nx = k.ny = n;
  // cv::getKernelType on the double kernel: a gaussian is symmetric
  bool sym = true;
  for (int i = 0; i < n; ++i)
    if (cd[i] != cd[n - 1 - i]) sym = false;
  k.tx = k.ty = sym ? 1 : 0;
  return k;
}

// Per-launch HIP-event timing of the hot kernels on a handle's stream.
struct Profiler {
  bool on = false;
  std::vector<hipEvent_t> pool;
  size_t used = 0;
  struct Pending {
    int kind;
    hipEvent_t a, b;
  };
  std::vector<Pending> pend;
  ssg_profile acc{};
  hipEvent_t a_ = nullptr;

  ~Profiler() {
    for (hipEvent_t e : pool) (void)hipEventDestroy(e);
  }
  hipEvent_t ev() {
    if (used == pool.size()) {
      hipEvent_t e = nullptr;
      if (hipEventCreate(&e) != hipSuccess) return nullptr;
      pool.push_back(e);
    }
    return pool[used++];
  }
  void start(hipStream_t s) {
    if (!on) return;
    a_ = ev();
    if (a_) (void)hipEventRecord(a_, s);
  }
  void stop(hipStream_t s, int kind) {
    if (!on || !a_) return;
    hipEvent_t b = ev();
    if (!b) return;
    (void)hipEventRecord(b, s);
    pend.push_back({kind, a_, b});
    a_ = nullptr;
  }
  // after the stream has finished: fold the pending durations in
  void collect() {
    for (const Pending& p : pend) {
      float ms = 0.f;
      if (hipEventElapsedTime(&ms, p.a, p.b) == hipSuccess) {
        acc.ms[p.kind] += ms;
        acc.launches[p.kind] += 1;
      }
    }
    pend.clear();
    used = 0;
  }
  void add_bytes(int kind, double b) {
    if (on) acc.bytes[kind] += b;
  }
};

}  // namespace

// =====================================================================================================
// registration
// =====================================================================================================
struct ssg_registration {
  ssg_image_registration_options o{};
  hipStream_t stream = nullptr;
  int nlevels = 0;
  int ref_w = 0, ref_h = 0;  // full-resolution reference size
  int ew = 0, eh = 0;        // ECC image size (after ecc.scale)
  int lw[MAX_LEVELS] = {}, lh[MAX_LEVELS] = {}, lpitch[MAX_LEVELS] = {};
  long long loff[MAX_LEVELS] = {};
  long long pyr_elems = 0;   // floats per frame pyramid
  long long mask_elems = 0;  // bytes per frame mask pyramid
  DevBuf ref_pyr, ref_mask_pyr;
  bool ref_has_mask[MAX_LEVELS] = {};
  double ref_mean[MAX_LEVELS] = {};
  // inverse-compositional aligners: reference gradients (2 planes per level, pyramid
  // layout), their Hessian moments, reference-mask pixel counts, the J record
  DevBuf ref_grad, ref_hm, ref_part, jrec;
  double rma[MAX_LEVELS] = {};
  // batch workspace
  int cap = 0;
  DevBuf cur_pyr, cur_mask_pyr, half_buf, norm_buf, gray_buf;
  DevBuf xf_t, xf_m, res, partials, active[2], active_count[2], lm;
  DevBuf frame_ptrs, mask_ptrs, zero_count, frame_list, tmp_in, tmp_mask, sum_tmp;
  DevBuf half_ptrs;  // sources of the batched half-size pyrDown
  int* pinned_count = nullptr;
  bool have_ref = false;
  Profiler prof;
  // eccflow refinement (enable_eccflow_registration, c_frame_registration.cc:900-917)
  ssg_eccflow* flow = nullptr;
  DevBuf cur_map;  // the last registered frame's remap at the reference size (refined)
  bool have_status = false;
  ssg_registration_status last{};
};

extern "C" void ssg_default_image_registration_options(ssg_image_registration_options* o) {
  if (!o) return;
  memset(o, 0, sizeof(*o));
  o->motion_type = SSG_MOTION_AFFINE;
  o->ecc_registration_channel = SSG_COLOR_CHANNEL_GRAY;
  o->interpolation = SSG_INTER_LINEAR;
  o->border_mode = SSG_BORDER_REFLECT101;
  o->ecc.scale = 0.5;
  o->ecc.eps = 0.2;
  o->ecc.min_rho = 0.8;
  o->ecc.input_smooth_sigma = 1.0;
  o->ecc.reference_smooth_sigma = 1.0;
  o->ecc.update_step_scale = 1.5;
  o->ecc.se_radius = 5;
  o->ecc.ecc_method = SSG_ECC_ALIGN_LM;
  o->ecc.max_iterations = 50;
  o->ecc.ecch_max_level = 0;
  o->ecc.ecch_minimum_image_size = 16;
  o->ecc.normalization_noise = 0.01;
  o->ecc.normalization_scale = 0;
  o->ecc.ecch_estimate_translation_first = 1;
  o->ecc.replace_planetary_disk_with_mask = 0;
  o->eccflow.update_multiplier = 1.5;
  o->eccflow.noise_level = -1;
  o->eccflow.scale_factor = 0.75;
  o->eccflow.max_iterations = 3;
  o->eccflow.support_scale = 4;
  o->eccflow.min_image_size = -1;
  o->eccflow.max_pyramid_level = -1;
  o->eccflow.downscale_method = 0;
  o->enable_feature_registration = 1;
  o->enable_ecc_registration = 0;
  o->enable_eccflow_registration = 0;
}

// ---- tuning knobs -----------------------------------------------------------------------------
namespace {
std::atomic<int> g_tune[SSG_TUNE_COUNT] = {{4}, {4096}, {8}, {32}, {1}};
const int g_tune_lo[SSG_TUNE_COUNT] = {2, 256, 1, 1, 0}, g_tune_hi[SSG_TUNE_COUNT] = {4, 65536, 16, 32, 1};
}  // namespace

namespace ssg {
int tune_value(int knob) { return knob >= 0 && knob < SSG_TUNE_COUNT ? g_tune[knob].load() : 0; }

int accumulate_chunks(int n, int w, int h) {
  const int th = 4 * tune_value(SSG_TUNE_ACC_ROWS);
  const long long tiles = (long long)((w + 63) / 64) * ((h + th - 1) / th);
  const long long target = tune_value(SSG_TUNE_ACC_BLOCKS);
  int nch = (int)std::max(1LL, std::min<long long>(16, (target + tiles - 1) / tiles));
  nch = std::min(nch, std::max(n, 1));
  const int per = (n + nch - 1) / nch;
  return (n + per - 1) / per;
}
}  // namespace ssg

extern "C" int ssg_tune_set(int32_t knob, int32_t value) {
  if (knob < 0 || knob >= SSG_TUNE_COUNT) return fail(SSG_E_INVALID_ARG, "tuning knob %d", knob);
  if (value < g_tune_lo[knob] || value > g_tune_hi[knob] || (knob == SSG_TUNE_ACC_ROWS && value != 2 && value != 4))
    return fail(SSG_E_INVALID_ARG, "tuning knob %d: value %d out of range", knob, value);
  g_tune[knob].store(value);
  return SSG_OK;
}

extern "C" int ssg_tune_get(int32_t knob, int32_t* value) {
  if (knob < 0 || knob >= SSG_TUNE_COUNT || !value) return fail(SSG_E_INVALID_ARG, "tuning knob %d", knob);
  *value = g_tune[knob].load();
  return SSG_OK;
}

extern "C" int ssg_abi_version(void) { return SSG_ABI_VERSION; }
extern "C" const char* ssg_last_error(void) { return g_err.c_str(); }

extern "C" int ssg_device_info(int* num_devices, char* name, int name_len) {
  int n = 0;
  HIPCHK(hipGetDeviceCount(&n));
  if (num_devices) *num_devices = n;
  if (n <= 0) return fail(SSG_E_HIP, "no HIP device");
  if (name && name_len > 0) {
    int dev = 0;
    HIPCHK(hipGetDevice(&dev));
    hipDeviceProp_t prop;
    HIPCHK(hipGetDeviceProperties(&prop, dev));
    snprintf(name, name_len, "%s (%s, %d CUs)", prop.name, prop.gcnArchName, prop.multiProcessorCount);
  }
  return SSG_OK;
}

extern "C" int ssg_registration_create(const ssg_image_registration_options* opts, ssg_registration** out) {
  if (!out) return fail(SSG_E_INVALID_ARG, "out is null");
  *out = nullptr;
  ssg_image_registration_options o;
  if (opts) o = *opts;
  else ssg_default_image_registration_options(&o);
  if (!o.enable_ecc_registration) return fail(SSG_E_UNSUPPORTED, "only ECC registration is on this path (enable_ecc_registration=0)");
  if (o.enable_feature_registration) return fail(SSG_E_UNSUPPORTED, "sparse feature registration is not on this path");
  if (o.ecc.ecc_method < SSG_ECC_ALIGN_FORWARD_ADDITIVE || o.ecc.ecc_method > SSG_ECC_ALIGN_INVERSE_COMPOSITIONAL_LM)
    return fail(SSG_E_INVALID_ARG, "ecc_method %d", o.ecc.ecc_method);
  if (o.ecc.max_iterations < 1) return fail(SSG_E_INVALID_ARG, "ecc.max_iterations %d", o.ecc.max_iterations);
  if (o.motion_type < SSG_MOTION_TRANSLATION || o.motion_type > SSG_MOTION_AFFINE) return fail(SSG_E_UNSUPPORTED, "motion_type %d", o.motion_type);
  if (!(o.ecc.scale <= 0 || o.ecc.scale == 1 || fabs(o.ecc.scale - 0.5) < 1e-2)) return fail(SSG_E_UNSUPPORTED, "ecc.scale %g: only 1 and 0.5", o.ecc.scale);
  if (o.ecc.replace_planetary_disk_with_mask) return fail(SSG_E_UNSUPPORTED, "replace_planetary_disk_with_mask");
  if (o.enable_eccflow_registration && o.eccflow.downscale_method != SSG_ECCFLOW_DOWNSCALE_PYRAMID &&
      !(o.eccflow.scale_factor > 0 && o.eccflow.scale_factor < 1))
    return fail(SSG_E_INVALID_ARG, "eccflow.scale_factor must be in (0, 1)");
  ssg_registration* r = new ssg_registration();
  r->o = o;
  hipError_t e = hipStreamCreateWithFlags(&r->stream, hipStreamNonBlocking);
  if (e != hipSuccess) {
    delete r;
    return fail(SSG_E_HIP, "hipStreamCreate: %s", hipGetErrorString(e));
  }
  e = hipHostMalloc((void**)&r->pinned_count, sizeof(int) * 4, hipHostMallocDefault);
  if (e != hipSuccess) {
    (void)hipStreamDestroy(r->stream);
    delete r;
    return fail(SSG_E_HIP, "hipHostMalloc: %s", hipGetErrorString(e));
  }
  int rc = ensure_tables(r->stream);
  if (rc == SSG_OK && o.enable_eccflow_registration) {
    // the reference sets every c_eccflow option from c_eccflow_registration_options
    // (c_frame_registration.cc:640-650)
    ssg_eccflow_options fo;
    ssg_default_eccflow_options(&fo);
    fo.update_multiplier = o.eccflow.update_multiplier;
    fo.input_smooth_sigma = o.eccflow.input_smooth_sigma;
    fo.reference_smooth_sigma = o.eccflow.reference_smooth_sigma;
    fo.noise_level = o.eccflow.noise_level;
    fo.scale_factor = o.eccflow.scale_factor;
    fo.max_iterations = o.eccflow.max_iterations;
    fo.support_scale = o.eccflow.support_scale;
    fo.min_image_size = o.eccflow.min_image_size;
    fo.max_pyramid_level = o.eccflow.max_pyramid_level;
    fo.downscale = o.eccflow.downscale_method;
    rc = ssg_eccflow_create(&fo, &r->flow);
  }
  if (rc != SSG_OK) {
    if (r->flow) ssg_eccflow_destroy(r->flow);
    (void)hipHostFree(r->pinned_count);
    (void)hipStreamDestroy(r->stream);
    delete r;
    return rc;
  }
  *out = r;
  return SSG_OK;
}

namespace {
void drop_stack_state(const ssg_registration* r);  // ssg_stack_api.inc
}

extern "C" void ssg_registration_destroy(ssg_registration* r) {
  if (r) drop_stack_state(r);
  if (!r) return;
  if (r->stream) (void)hipStreamSynchronize(r->stream);
  if (r->flow) ssg_eccflow_destroy(r->flow);
  if (r->pinned_count) (void)hipHostFree(r->pinned_count);
  if (r->stream) (void)hipStreamDestroy(r->stream);
  delete r;
}

namespace {

bool use_half(const ssg_image_registration_options& o) { return o.ecc.scale > 0 && o.ecc.scale != 1; }

// scaleImage's pyrDown of a batch in one strip launch (SSG_TUNE_PYR_STRIP 1) when the geometry
// allows it; else one tiled launch per frame (the same values)
bool half_batch(const ssg_registration* r, int src_pitch, const float* dst, long long dst_fstride, int dst_pitch,
                int n) {
  return tune_value(SSG_TUNE_PYR_STRIP) != 0 &&
         strip_pyrdown_supported(nullptr, 0, src_pitch, r->ref_w, r->ref_h, dst, dst_fstride, dst_pitch, r->ew, r->eh,
                                 n);
}

// Builds one level-0 image (smoothed or copied) and the pyramid for n frames.
// src_ptrs: device array of n frame pointers (f32, pitch src_pitch floats).
int build_pyramids(ssg_registration* r, const float* const* src_ptrs, int src_pitch, int n, double sigma,
                   float* pyr) {
  hipStream_t s = r->stream;
  // strip kernels (SSG_TUNE_PYR_STRIP 1): level 0 smoothed together with level 1, the other levels
  // by pyrDown strips; else the tiled smoothing and pyrDown (the same values)
  const bool strip = tune_value(SSG_TUNE_PYR_STRIP) != 0;
  int l1 = 1;  // first level formed by pyrDown alone
  if (sigma > 0) {
    const SepKernel k = gaussian_sep(sigma);
    if (strip && r->nlevels > 1 &&
        smooth_pyrdown_supported(k, src_pitch, pyr + r->loff[0], r->pyr_elems, r->lpitch[0], r->lw[0], r->lh[0],
                                 r->loff[1] - r->loff[0], r->lpitch[1], r->lw[1], r->lh[1], n)) {
      HIPCHK(launch_smooth_pyrdown(src_ptrs, src_pitch, pyr + r->loff[0], r->pyr_elems, r->lpitch[0], r->lw[0],
                                   r->lh[0], r->loff[1] - r->loff[0], r->lpitch[1], r->lw[1], r->lh[1], n, k, s));
      l1 = 2;
    } else {
      HIPCHK(launch_sepfilter(src_ptrs, src_pitch, pyr + r->loff[0], r->pyr_elems, r->lpitch[0], r->lw[0], r->lh[0],
                              n, k, BORDER_REPLICATE, s));
    }
  } else {
    HIPCHK(launch_copy_frames(src_ptrs, src_pitch, pyr + r->loff[0], r->pyr_elems, r->lpitch[0], r->lw[0], r->lh[0], n, s));
  }
  for (int l = l1; l < r->nlevels; ++l) {
    const float* sp = pyr + r->loff[l - 1];
    float* dp = pyr + r->loff[l];
    if (strip && strip_pyrdown_supported(sp, r->pyr_elems, r->lpitch[l - 1], r->lw[l - 1], r->lh[l - 1], dp,
                                         r->pyr_elems, r->lpitch[l], r->lw[l], r->lh[l], n))
      HIPCHK(launch_strip_pyrdown(nullptr, sp, r->pyr_elems, r->lpitch[l - 1], r->lw[l - 1], r->lh[l - 1], dp,
                                  r->pyr_elems, r->lpitch[l], r->lw[l], r->lh[l], n, s));
    else
      HIPCHK(launch_pyrdown(sp, r->pyr_elems, r->lpitch[l - 1], r->lw[l - 1], r->lh[l - 1], dp, r->pyr_elems,
                            r->lpitch[l], r->lw[l], r->lh[l], n, BORDER_REFLECT101, s));
  }
  return SSG_OK;
}

// Mask pyramids: level l = INTER_NEAREST resize of the un-eroded level l-1
// mask; then each level is eroded 5x5 (BORDER_REPLICATE) when it has a zero
// (c_ecc_forward_additive::set_current_image / set_reference_image).
// has_mask[l*n + f] receives whether frame f keeps a mask at level l
// (reference masks without zeros are dropped).
int build_mask_pyramids(ssg_registration* r, const uint8_t* const* src_ptrs, int src_pitch, int n, uint8_t* mpyr,
                        bool is_ref, std::vector<char>& keep, unsigned* zeros_out = nullptr) {
  // partial masks are eroded by the forward-additive aligner (reference and current,
  // ecc2.cc:1187-1234) and by c_ecclm (reference only, ecc2.cc:1387-1400); the
  // inverse-compositional aligners keep them as given (ecc2.cc:585-630)
  const int method = r->o.ecc.ecc_method;
  const bool erode = method == SSG_ECC_ALIGN_FORWARD_ADDITIVE || (method == SSG_ECC_ALIGN_LM && is_ref);
  hipStream_t s = r->stream;
  // raw (un-eroded) chain into tmp_mask, one level at a time, then erode into mpyr
  CHK(r->tmp_mask.ensure((size_t)r->mask_elems * n));
  uint8_t* raw = r->tmp_mask.as<uint8_t>();
  CHK(r->frame_ptrs.ensure(sizeof(void*) * n * 2));
  std::vector<const uint8_t*> ptrs(n);
  CHK(r->zero_count.ensure(sizeof(unsigned) * n));
  keep.assign((size_t)r->nlevels * n, 0);
  std::vector<unsigned> zc(n);
  for (int l = 0; l < r->nlevels; ++l) {
    const int lp = r->lpitch[l];
    if (l == 0) {
      HIPCHK(launch_mask_prepare(src_ptrs, src_pitch, raw + r->loff[0], r->mask_elems, lp, r->lw[0], r->lh[0], r->lw[0],
                                 r->lh[0], n, s));
    } else {
      for (int f = 0; f < n; ++f) ptrs[f] = raw + (size_t)f * r->mask_elems + r->loff[l - 1];
      const uint8_t** dptrs = r->frame_ptrs.as<const uint8_t*>() + n;
      HIPCHK(hipMemcpyAsync((void*)dptrs, ptrs.data(), sizeof(void*) * n, hipMemcpyHostToDevice, s));
      HIPCHK(launch_mask_prepare(dptrs, r->lpitch[l - 1], raw + r->loff[l], r->mask_elems, lp, r->lw[l - 1],
                                 r->lh[l - 1], r->lw[l], r->lh[l], n, s));
    }
    HIPCHK(hipMemsetAsync(r->zero_count.p, 0, sizeof(unsigned) * n, s));
    HIPCHK(launch_count_zero_u8(raw + r->loff[l], r->mask_elems, lp, r->lw[l], r->lh[l], n, r->zero_count.as<unsigned>(), s));
    HIPCHK(hipMemcpyAsync(zc.data(), r->zero_count.p, sizeof(unsigned) * n, hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    for (int f = 0; f < n; ++f) {
      uint8_t* src = raw + (size_t)f * r->mask_elems + r->loff[l];
      uint8_t* dst = mpyr + (size_t)f * r->mask_elems + r->loff[l];
      if (zeros_out && f == 0) zeros_out[l] = zc[0];
      if (zc[f] == 0) {
        keep[(size_t)l * n + f] = is_ref ? 0 : 1;
        if (!is_ref) HIPCHK(hipMemcpy2DAsync(dst, lp, src, lp, r->lw[l], r->lh[l], hipMemcpyDeviceToDevice, s));
      } else {
        keep[(size_t)l * n + f] = 1;
        if (erode) HIPCHK(launch_erode5(src, 0, lp, dst, 0, lp, r->lw[l], r->lh[l], 1, BORDER_REPLICATE, 255, s));
        else HIPCHK(hipMemcpy2DAsync(dst, lp, src, lp, r->lw[l], r->lh[l], hipMemcpyDeviceToDevice, s));
      }
    }
  }
  return SSG_OK;
}

int ensure_workspace(ssg_registration* r, int n) {
  if (n <= r->cap) return SSG_OK;
  const int tx0 = (r->lw[0] + ECC_TW - 1) / ECC_TW, ty0 = (r->lh[0] + ECC_TH - 1) / ECC_TH;
  CHK(r->cur_pyr.ensure(sizeof(float) * (size_t)r->pyr_elems * n));
  CHK(r->xf_t.ensure(sizeof(FrameXf) * n));
  CHK(r->xf_m.ensure(sizeof(FrameXf) * n));
  CHK(r->res.ensure(sizeof(FrameRes) * n));
  CHK(r->partials.ensure(sizeof(double) * (size_t)n * tx0 * ty0 * NMOM_FULL));
  if (r->o.ecc.ecc_method != SSG_ECC_ALIGN_FORWARD_ADDITIVE) CHK(r->lm.ensure(sizeof(FrameLm) * n));
  for (int i = 0; i < 2; ++i) {
    CHK(r->active[i].ensure(sizeof(int) * n));
    CHK(r->active_count[i].ensure(sizeof(int) * 4));
  }
  CHK(r->frame_ptrs.ensure(sizeof(void*) * n * 2));
  CHK(r->mask_ptrs.ensure(sizeof(void*) * n));
  CHK(r->frame_list.ensure(sizeof(int) * n));
  std::vector<int> list(n);
  for (int i = 0; i < n; ++i) list[i] = i;
  HIPCHK(hipMemcpyAsync(r->frame_list.p, list.data(), sizeof(int) * n, hipMemcpyHostToDevice, r->stream));
  HIPCHK(hipStreamSynchronize(r->stream));
  r->cap = n;
  return SSG_OK;
}

// One ECC pass (translation-first or main) over all levels for n frames.
// plain: the host waits for each iteration and stops when no frame is active,
// the per-frame form of the reference loop.
int run_ecc(ssg_registration* r, int n, int pass, bool has_cmask, bool plain, bool poll = true) {
  hipStream_t s = r->stream;
  const int motion = pass == 0 ? SSG_MOTION_TRANSLATION : r->o.motion_type;
  const int method = r->o.ecc.ecc_method;
  const bool tfirst = r->o.motion_type != SSG_MOTION_TRANSLATION && r->o.ecc.ecch_estimate_translation_first &&
                      r->o.ecc.ecch_max_level != 0;
  const bool ic = method == ECC_IC || method == ECC_ICLM;
  const int L = r->nlevels;
  const int grid = ecc_persistent_grid();
  const bool translation_first_ran = pass == 1 && r->o.motion_type != SSG_MOTION_TRANSLATION &&
                                     r->o.ecc.ecch_estimate_translation_first && r->o.ecc.ecch_max_level != 0;
  for (int l = L - 1; l >= 0; --l) {
    // pass 0 works on xf_t; pass 1 on xf_m and seeds its translation from xf_t
    // when the translation-first pass ran (xf_t == nullptr otherwise)
    LevelBeginArgs lb{};
    lb.xf_t = (pass == 0 || translation_first_ran) ? r->xf_t.as<FrameXf>() : nullptr;
    lb.xf_m = r->xf_m.as<FrameXf>();
    lb.res = r->res.as<FrameRes>();
    lb.nframes = n;
    lb.pass = pass;
    lb.level = l;
    lb.nlevels = L;
    lb.ratio = l == L - 1 ? (double)r->lw[L - 1] / (double)r->lw[0] : (double)r->lw[l] / (double)r->lw[l + 1];
    lb.main_motion = r->o.motion_type;
    lb.cf = (float)r->ref_mean[l];
    lb.active_out = r->active[0].as<int>();
    lb.active_count_out = r->active_count[0].as<int>();
    lb.method = method;
    lb.lm = method != ECC_FA ? r->lm.as<FrameLm>() : nullptr;
    // the inverse-compositional J is dropped whenever the translation-first pass swaps
    // the aligner's transform (c_frame_registration.cc:823,844); otherwise it persists
    lb.jrec = (method == ECC_IC || method == ECC_ICLM) && !tfirst ? r->jrec.as<JRec>() : nullptr;
    HIPCHK(launch_level_begin(lb, s));

    EccPassArgs a{};
    a.cur = r->cur_pyr.as<float>() + r->loff[l];
    a.cur_fstride = r->pyr_elems;
    a.cur_pitch = r->lpitch[l];
    a.cw = r->lw[l];
    a.ch = r->lh[l];
    a.cmask = has_cmask ? r->cur_mask_pyr.as<uint8_t>() + r->loff[l] : nullptr;
    a.cmask_fstride = r->mask_elems;
    a.cmask_pitch = r->lpitch[l];
    a.ref = r->ref_pyr.as<float>() + r->loff[l];
    a.ref_pitch = r->lpitch[l];
    a.w = r->lw[l];
    a.h = r->lh[l];
    a.rmask = r->ref_has_mask[l] ? r->ref_mask_pyr.as<uint8_t>() + r->loff[l] : nullptr;
    a.rmask_pitch = r->lpitch[l];
    a.cf = (float)r->ref_mean[l];
    a.tiles_x = (a.w + ECC_TW - 1) / ECC_TW;
    a.tiles_y = (a.h + ECC_TH - 1) / ECC_TH;
    a.ntiles = a.tiles_x * a.tiles_y;
    a.xf = pass == 0 ? r->xf_t.as<FrameXf>() : r->xf_m.as<FrameXf>();
    a.res = r->res.as<FrameRes>();
    a.partials = r->partials.as<float>();
    a.nframes = n;
    a.pass = pass;
    a.level = l;
    a.max_iterations = r->o.ecc.max_iterations;
    double max_eps = r->o.ecc.eps * (double)(1 << l);
    // only the forward-additive aligner replaces a non-positive max_eps (ecc2.cc:1274-1276)
    if (max_eps <= 0 && method == ECC_FA) max_eps = 1e-3;
    a.max_eps = (float)max_eps;
    a.update_step_scale = r->o.ecc.update_step_scale;
    a.method = method;
    a.lm = method != ECC_FA ? r->lm.as<FrameLm>() : nullptr;
    if (ic) {
      a.rgx = r->ref_grad.as<float>() + r->loff[l];
      a.rgy = r->ref_grad.as<float>() + r->pyr_elems + r->loff[l];
      a.ref_h = r->ref_hm.as<double>() + (size_t)NMOM_REF * l;
      a.rma = r->rma[l];
      a.tiles_x = (a.w + CORR_TW - 1) / CORR_TW;
      a.tiles_y = (a.h + CORR_TH - 1) / CORR_TH;
      a.ntiles = a.tiles_x * a.tiles_y;
    }
    const int g = std::min(grid, std::max(8, n * a.ntiles));
    // the Levenberg-Marquardt aligners evaluate the starting parameters before their
    // first counted iteration: one pass more
    const int max_passes = r->o.ecc.max_iterations + (method == ECC_LM || method == ECC_ICLM ? 1 : 0);
    for (int it = 0; it < max_passes; ++it) {
      a.active_in = r->active[it & 1].as<int>();
      a.active_count_in = r->active_count[it & 1].as<int>();
      a.active_out = r->active[(it + 1) & 1].as<int>();
      a.active_count_out = r->active_count[(it + 1) & 1].as<int>();
      r->prof.start(s);
      if (method == ECC_FA) HIPCHK(launch_ecc_pass(motion, a, g, s));
      else if (method == ECC_LM) HIPCHK(launch_lm_pass(motion, a, g, s));
      else HIPCHK(launch_ic_pass(motion, a, g, s));
      r->prof.stop(s, pass == 0 ? SSG_PROF_ECC_TRANSLATION : SSG_PROF_ECC_MAIN);
      r->prof.start(s);
      if (method == ECC_FA) HIPCHK(launch_ecc_solve(motion, a, s));
      else HIPCHK(launch_lm_solve(motion, a, s));
      r->prof.stop(s, SSG_PROF_ECC_SOLVE);
      // device-driven batches (poll == false) never wait on the host: passes over an empty active
      // list return at once, so the level's full pass budget is queued without a round trip
      const bool check = poll && (plain || ((it & 3) == 3));
      if (check && it + 1 < max_passes) {
        HIPCHK(hipMemcpyAsync(r->pinned_count, a.active_count_out, sizeof(int), hipMemcpyDeviceToHost, s));
        HIPCHK(hipStreamSynchronize(s));
        if (*r->pinned_count == 0) break;
      }
    }
  }
  return SSG_OK;
}

int run_corr(ssg_registration* r, int n, int pass, bool has_cmask) {
  CorrPassArgs c{};
  c.cur = r->cur_pyr.as<float>() + r->loff[0];
  c.cur_fstride = r->pyr_elems;
  c.cur_pitch = r->lpitch[0];
  c.cw = r->lw[0];
  c.ch = r->lh[0];
  c.cmask = has_cmask ? r->cur_mask_pyr.as<uint8_t>() + r->loff[0] : nullptr;
  c.cmask_fstride = r->mask_elems;
  c.cmask_pitch = r->lpitch[0];
  c.ref = r->ref_pyr.as<float>() + r->loff[0];
  c.ref_pitch = r->lpitch[0];
  c.w = r->lw[0];
  c.h = r->lh[0];
  c.rmask = r->ref_has_mask[0] ? r->ref_mask_pyr.as<uint8_t>() + r->loff[0] : nullptr;
  c.rmask_pitch = r->lpitch[0];
  c.cf = (float)r->ref_mean[0];
  c.tiles_x = (c.w + CORR_TW - 1) / CORR_TW;
  c.tiles_y = (c.h + CORR_TH - 1) / CORR_TH;
  c.ntiles = c.tiles_x * c.tiles_y;
  c.xf = pass == 0 ? r->xf_t.as<FrameXf>() : r->xf_m.as<FrameXf>();
  c.res = r->res.as<FrameRes>();
  c.partials = r->partials.as<float>();
  c.frames = r->frame_list.as<int>();
  c.nlist = n;
  c.pass = pass;
  c.min_rho = r->o.ecc.min_rho;
  const int g = std::min(ecc_persistent_grid(), std::max(8, n * c.ntiles));
  r->prof.start(r->stream);
  HIPCHK(launch_corr_pass(pass == 0 ? SSG_MOTION_TRANSLATION : r->o.motion_type, c, g, r->stream));
  r->prof.stop(r->stream, SSG_PROF_CORRELATION);
  return SSG_OK;
}

// prepares level-0 sources (optionally the 0.5 pyrDown of scaleImage) and the pyramids of n frames
// ecc_normalize(src, mask, dst, lvl) (ecc2.cc:385-397) of n frames given by device pointers:
// dst = src - pyrUp chain(pyrDown^lvl(src, BORDER_REPLICATE)) back through the same sizes
// (ecc_downscale / ecc_upscale, ecc2.cc:224-275), then 0 where the mask is not 255.  masks:
// device pointers (entries may be null) or an empty vector.  Chunks of 16 frames.
int ecc_normalize_frames(const std::vector<const float*>& src, int spitch, const std::vector<const uint8_t*>& masks,
                         int mpitch, int w, int h, int lvl, float* out, int opitch, long long ofs, hipStream_t s) {
  const int n = (int)src.size();
  if (lvl < 1) return fail(SSG_E_INVALID_ARG, "normalization level %d", lvl);
  std::vector<int> lw(1, w), lh(1, h);
  std::vector<size_t> loff(1, 0);
  size_t tot = 0;
  for (int l = 1; l <= lvl; ++l) {
    lw.push_back((lw.back() + 1) / 2);
    lh.push_back((lh.back() + 1) / 2);
    if (lw[l] == lw[l - 1] && lh[l] == lh[l - 1])
      return fail(SSG_E_UNSUPPORTED, "normalization level %d below a 1-pixel image (%dx%d)", lvl, w, h);
    loff.push_back(tot);
    tot += (size_t)lw[l] * lh[l];
  }
  constexpr int CH = 16;
  const size_t n0 = (size_t)w * h;
  DevBuf down, upa, upb, ptrs;
  CHK(down.ensure(sizeof(float) * tot * CH));
  CHK(upa.ensure(sizeof(float) * n0 * CH));
  CHK(upb.ensure(sizeof(float) * n0 * CH));
  CHK(ptrs.ensure(sizeof(void*) * 2 * CH));
  auto L = [&](int l) { return down.as<float>() + loff[l] * CH; };
  auto N = [&](int l) { return (long long)lw[l] * lh[l]; };
  std::vector<const void*> hp(2 * CH);
  for (int b = 0; b < n; b += CH) {
    const int m = std::min(CH, n - b);
    for (int i = 0; i < m; ++i)
      HIPCHK(launch_pyrdown(src[b + i], 0, spitch, w, h, L(1) + i * N(1), 0, lw[1], lw[1], lh[1], 1, BORDER_REPLICATE,
                            s));
    for (int l = 2; l <= lvl; ++l)
      HIPCHK(launch_pyrdown(L(l - 1), N(l - 1), lw[l - 1], lw[l - 1], lh[l - 1], L(l), N(l), lw[l], lw[l], lh[l], m,
                            BORDER_REPLICATE, s));
    const float* cur = L(lvl);
    float* bufs[2] = {upa.as<float>(), upb.as<float>()};
    for (int l = lvl - 1; l >= 0; --l) {
      float* dst = bufs[l & 1];
      HIPCHK(launch_pyrup(cur, N(l + 1), lw[l + 1], lh[l + 1], dst, N(l), lw[l], lh[l], m, s));
      cur = dst;
    }
    for (int i = 0; i < m; ++i) {
      hp[i] = src[b + i];
      hp[CH + i] = masks.empty() ? nullptr : masks[b + i];
    }
    HIPCHK(hipMemcpyAsync(ptrs.p, hp.data(), sizeof(void*) * 2 * CH, hipMemcpyHostToDevice, s));
    HIPCHK(launch_norm_sub(ptrs.as<const float*>(), spitch, cur, N(0),
                           masks.empty() ? nullptr : ptrs.as<const uint8_t*>() + CH, mpitch, out + b * ofs, opitch, ofs,
                           w, h, m, s));
    HIPCHK(hipStreamSynchronize(s));  // the pointer table is reused
  }
  return SSG_OK;
}

bool use_normalize(const ssg_image_registration_options& o) {
  return o.ecc.normalization_scale > 0 && o.ecc.normalization_noise > 0;
}

int prepare_batch(ssg_registration* r, int n, const ssg_image* frames, const ssg_image* masks, bool* has_cmask) {
  hipStream_t s = r->stream;
  CHK(ensure_workspace(r, n));
  std::vector<const float*> ptrs(n), src_list;
  int src_pitch = 0;
  const bool half = use_half(r->o);
  size_t in_pitch = 0;
  for (int f = 0; f < n; ++f) {
    CHK(check_image(&frames[f], "frame"));
    if (frames[f].type != SSG_32FC1) return fail(SSG_E_UNSUPPORTED, "frame %d: registration takes SSG_32FC1 (gray) images", f);
    if (frames[f].width != r->ref_w || frames[f].height != r->ref_h)
      return fail(SSG_E_INVALID_ARG, "frame %d: size %dx%d differs from the reference %dx%d", f, frames[f].width,
                  frames[f].height, r->ref_w, r->ref_h);
  }
  // frames not on the device, or with different steps, go through one staging buffer
  bool direct = true;
  for (int f = 0; f < n; ++f)
    if (!frames[f].on_device || frames[f].step != frames[0].step || (frames[f].step % 4)) direct = false;
  if (direct) {
    for (int f = 0; f < n; ++f) ptrs[f] = (const float*)frames[f].data;
    src_pitch = (int)(frames[0].step / 4);
  } else {
    in_pitch = ((size_t)r->ref_w * 4 + 255) / 256 * 256;
    CHK(r->tmp_in.ensure(in_pitch * r->ref_h * n));
    for (int f = 0; f < n; ++f) {
      uint8_t* dst = r->tmp_in.as<uint8_t>() + in_pitch * r->ref_h * f;
      HIPCHK(hipMemcpy2DAsync(dst, in_pitch, frames[f].data, (size_t)frames[f].step, (size_t)r->ref_w * 4, r->ref_h,
                              frames[f].on_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, s));
      ptrs[f] = (const float*)dst;
    }
    src_pitch = (int)(in_pitch / 4);
  }
  if (half) {
    // scaleImage: pyrDown to the default half size (c_frame_registration.cc:230-250)
    const size_t hp = ((size_t)r->ew * 4 + 255) / 256 * 256;
    CHK(r->half_buf.ensure(hp * r->eh * n));
    float* hb = r->half_buf.as<float>();
    const long long hfs = (long long)(hp / 4) * r->eh;
    if (half_batch(r, src_pitch, hb, hfs, (int)(hp / 4), n)) {
      // the whole batch in one launch, through a table of the sources (kept unchanged until the
      // stream has read it: the function synchronises before it returns)
      src_list.assign(ptrs.begin(), ptrs.end());
      CHK(r->half_ptrs.ensure(sizeof(void*) * n));
      HIPCHK(hipMemcpyAsync(r->half_ptrs.p, src_list.data(), sizeof(void*) * n, hipMemcpyHostToDevice, s));
      HIPCHK(launch_strip_pyrdown(r->half_ptrs.as<const float*>(), nullptr, 0, src_pitch, r->ref_w, r->ref_h, hb, hfs,
                                  (int)(hp / 4), r->ew, r->eh, n, s));
    } else {
      // pyrdown takes base + stride: run frame by frame on the pointers
      for (int f = 0; f < n; ++f)
        HIPCHK(launch_pyrdown(ptrs[f], 0, src_pitch, r->ref_w, r->ref_h, hb + hfs * f, 0, (int)(hp / 4), r->ew, r->eh,
                              1, BORDER_REFLECT101, s));
    }
    for (int f = 0; f < n; ++f) ptrs[f] = hb + hfs * f;
    src_pitch = (int)(hp / 4);
  }
  DevBuf tmpnm;
  if (use_normalize(r->o)) {
    // ecc_normalize of the (scaled) frames with their masks (c_frame_registration.cc:807-812)
    std::vector<const uint8_t*> nm;
    int nmp = 0;
    if (masks) {
      if (half) return fail(SSG_E_UNSUPPORTED, "current masks with ecc.scale 0.5");
      const size_t mp = ((size_t)r->ref_w + 255) / 256 * 256;
      CHK(tmpnm.ensure(mp * r->ref_h * n));
      for (int f = 0; f < n; ++f) {
        CHK(check_image(&masks[f], "mask"));
        if (masks[f].type != SSG_8UC1 || masks[f].width != r->ref_w || masks[f].height != r->ref_h)
          return fail(SSG_E_INVALID_ARG, "mask %d must be SSG_8UC1 of the frame size", f);
        uint8_t* dst = tmpnm.as<uint8_t>() + mp * r->ref_h * f;
        HIPCHK(hipMemcpy2DAsync(dst, mp, masks[f].data, (size_t)masks[f].step, r->ref_w, r->ref_h,
                                masks[f].on_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, s));
        nm.push_back(dst);
      }
      nmp = (int)mp;
    }
    const size_t np = ((size_t)r->ew * 4 + 255) / 256 * 256;
    CHK(r->norm_buf.ensure(np * r->eh * n));
    const long long nfs = (long long)(np / 4) * r->eh;
    CHK(ecc_normalize_frames(ptrs, src_pitch, nm, nmp, r->ew, r->eh, r->o.ecc.normalization_scale,
                             r->norm_buf.as<float>(), (int)(np / 4), nfs, s));
    for (int f = 0; f < n; ++f) ptrs[f] = r->norm_buf.as<float>() + nfs * f;
    src_pitch = (int)(np / 4);
  }
  HIPCHK(hipMemcpyAsync(r->frame_ptrs.p, ptrs.data(), sizeof(void*) * n, hipMemcpyHostToDevice, s));
  r->prof.start(s);
  CHK(build_pyramids(r, r->frame_ptrs.as<const float*>(), src_pitch, n, r->o.ecc.input_smooth_sigma, r->cur_pyr.as<float>()));
  r->prof.stop(s, SSG_PROF_PYRAMID);
  {
    // level 0 reads the frame and writes the smoothed level; each further level reads its parent
    double b = 8.0 * r->lw[0] * r->lh[0];
    for (int l = 1; l < r->nlevels; ++l) b += 4.0 * r->lw[l - 1] * r->lh[l - 1] + 4.0 * r->lw[l] * r->lh[l];
    r->prof.add_bytes(SSG_PROF_PYRAMID, b * n);
  }
  *has_cmask = false;
  if (masks) {
    if (half) return fail(SSG_E_UNSUPPORTED, "current masks with ecc.scale 0.5");
    std::vector<const uint8_t*> mp(n);
    DevBuf tmpm;
    const size_t mpitch = ((size_t)r->ref_w + 255) / 256 * 256;
    bool mdirect = true;
    for (int f = 0; f < n; ++f) {
      CHK(check_image(&masks[f], "mask"));
      if (masks[f].type != SSG_8UC1 || masks[f].width != r->ref_w || masks[f].height != r->ref_h)
        return fail(SSG_E_INVALID_ARG, "mask %d must be SSG_8UC1 of the frame size", f);
      if (!masks[f].on_device || masks[f].step != masks[0].step) mdirect = false;
    }
    int mpitch_el;
    if (mdirect) {
      for (int f = 0; f < n; ++f) mp[f] = (const uint8_t*)masks[f].data;
      mpitch_el = (int)masks[0].step;
    } else {
      CHK(tmpm.ensure(mpitch * r->ref_h * n));
      for (int f = 0; f < n; ++f) {
        uint8_t* dst = tmpm.as<uint8_t>() + mpitch * r->ref_h * f;
        HIPCHK(hipMemcpy2DAsync(dst, mpitch, masks[f].data, (size_t)masks[f].step, r->ref_w, r->ref_h,
                                masks[f].on_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, s));
        mp[f] = dst;
      }
      mpitch_el = (int)mpitch;
    }
    CHK(r->mask_ptrs.ensure(sizeof(void*) * n));
    HIPCHK(hipMemcpyAsync(r->mask_ptrs.p, mp.data(), sizeof(void*) * n, hipMemcpyHostToDevice, s));
    CHK(r->cur_mask_pyr.ensure((size_t)r->mask_elems * n));
    std::vector<char> keep;
    CHK(build_mask_pyramids(r, r->mask_ptrs.as<const uint8_t*>(), mpitch_el, n, r->cur_mask_pyr.as<uint8_t>(), false, keep));
    HIPCHK(hipStreamSynchronize(s));
    *has_cmask = true;
  }
  return SSG_OK;
}

// create_ecc_image's channel extraction (c_frame_registration.cc:1009-1022, extract_channel with
// ecc_registration_channel): colour (BGR f32) frames become device gray images in `buf`, `views`
// holds their views and *out points at them; gray frames pass through unchanged.
int ecc_channel_views(ssg_registration* r, int n, const ssg_image* frames, DevBuf& buf, std::vector<ssg_image>& views,
                      const ssg_image** out) {
  *out = frames;
  if (n <= 0 || frames[0].type != SSG_32FC3) return SSG_OK;
  int ch;
  switch (r->o.ecc_registration_channel) {
    case SSG_COLOR_CHANNEL_GRAY: ch = -1; break;
    case SSG_COLOR_CHANNEL_BLUE: case SSG_COLOR_CHANNEL_0: ch = 0; break;
    case SSG_COLOR_CHANNEL_GREEN: case SSG_COLOR_CHANNEL_1: ch = 1; break;
    case SSG_COLOR_CHANNEL_RED: case SSG_COLOR_CHANNEL_2: ch = 2; break;
    default:
      return fail(SSG_E_UNSUPPORTED, "ecc_registration_channel %d on colour frames (gray, blue, green, red, 0..2)",
                  r->o.ecc_registration_channel);
  }
  const int w = frames[0].width, h = frames[0].height;
  for (int f = 0; f < n; ++f) {
    CHK(check_image(&frames[f], "frame"));
    if (frames[f].type != SSG_32FC3 || frames[f].width != w || frames[f].height != h)
      return fail(SSG_E_INVALID_ARG, "frame %d: colour frames of one size", f);
  }
  hipStream_t s = r->stream;
  const size_t gp = ((size_t)w * 4 + 255) / 256 * 256, cp = ((size_t)w * 12 + 255) / 256 * 256;
  bool direct = true;
  for (int f = 0; f < n; ++f)
    if (!frames[f].on_device || frames[f].step != frames[0].step || frames[f].step % 4) direct = false;
  CHK(buf.ensure(gp * h * n + (direct ? 0 : cp * h * n) + sizeof(void*) * n));
  char* gbase = buf.as<char>();
  char* cbase = gbase + gp * h * n;
  std::vector<const float*> src(n);
  int spitch;
  if (direct) {
    for (int f = 0; f < n; ++f) src[f] = (const float*)frames[f].data;
    spitch = (int)(frames[0].step / 4);
  } else {
    for (int f = 0; f < n; ++f) {
      char* d = cbase + cp * h * f;
      HIPCHK(hipMemcpy2DAsync(d, cp, frames[f].data, (size_t)frames[f].step, (size_t)w * 12, h,
                              frames[f].on_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, s));
      src[f] = (const float*)d;
    }
    spitch = (int)(cp / 4);
  }
  const float** dsrc = (const float**)(gbase + gp * h * n + (direct ? 0 : cp * h * n));
  HIPCHK(hipMemcpyAsync((void*)dsrc, src.data(), sizeof(void*) * n, hipMemcpyHostToDevice, s));
  HIPCHK(launch_ecc_channel(dsrc, spitch, w, h, n, ch, (float*)gbase, (long long)(gp / 4) * h, (int)(gp / 4), s));
  HIPCHK(hipStreamSynchronize(s));
  views.assign(n, ssg_image{});
  for (int f = 0; f < n; ++f) {
    views[f].data = gbase + gp * h * f;
    views[f].width = w;
    views[f].height = h;
    views[f].step = (int64_t)gp;
    views[f].type = SSG_32FC1;
    views[f].on_device = 1;
  }
  *out = views.data();
  return SSG_OK;
}

void fill_statuses(ssg_registration* r, int n, const FrameXf* xf, const FrameRes* res, ssg_registration_status* st);

void account_ecc_bytes(ssg_registration* r, int n, const FrameRes* res, bool has_cmask);

int register_n(ssg_registration* r, int n, const ssg_image* frames, const ssg_image* masks, ssg_registration_status* st,
               bool plain) {
  if (!r->have_ref) return fail(SSG_E_NOT_READY, "setup_reference_frame() was not called");
  if (n <= 0) return SSG_OK;
  if (!frames || !st) return fail(SSG_E_INVALID_ARG, "null frames or status");
  hipStream_t s = r->stream;
  std::vector<ssg_image> gv;
  CHK(ecc_channel_views(r, n, frames, r->gray_buf, gv, &frames));
  bool has_cmask = false;
  CHK(prepare_batch(r, n, frames, masks, &has_cmask));
  HIPCHK(launch_zero(r->res.p, sizeof(FrameRes) * n, s));
  const bool tfirst = r->o.motion_type != SSG_MOTION_TRANSLATION && r->o.ecc.ecch_estimate_translation_first &&
                      r->o.ecc.ecch_max_level != 0;
  if (tfirst) {
    CHK(run_ecc(r, n, 0, has_cmask, plain));
    CHK(run_corr(r, n, 0, has_cmask));
  }
  CHK(run_ecc(r, n, 1, has_cmask, plain));
  CHK(run_corr(r, n, 1, has_cmask));
  std::vector<FrameXf> xf(n);
  std::vector<FrameRes> res(n);
  HIPCHK(hipMemcpyAsync(xf.data(), r->xf_m.p, sizeof(FrameXf) * n, hipMemcpyDeviceToHost, s));
  HIPCHK(hipMemcpyAsync(res.data(), r->res.p, sizeof(FrameRes) * n, hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  if (r->prof.on) {
    account_ecc_bytes(r, n, res.data(), has_cmask);
    r->prof.collect();
  }
  fill_statuses(r, n, xf.data(), res.data(), st);
  return SSG_OK;
}

// algorithmic bytes of the ECC passes of a batch, from its per-frame records
void account_ecc_bytes(ssg_registration* r, int n, const FrameRes* res, bool has_cmask) {
  const bool tfirst = r->o.motion_type != SSG_MOTION_TRANSLATION && r->o.ecc.ecch_estimate_translation_first &&
                      r->o.ecc.ecch_max_level != 0;
  {
    // every iteration of a frame at level l reads the warped current level and the reference
    // level once (4 + 4 B/px, +1 B/px per mask); correlation passes read both level-0 images
    const int maxit = r->o.ecc.max_iterations;
    for (int f = 0; f < n; ++f) {
      for (int p = 0; p < 2; ++p) {
        if (p == 0 && !tfirst) continue;
        if (p == 1 && res[f].rejected == SSG_REG_RHO_TRANSLATION) continue;
        for (int l = 0; l < r->nlevels; ++l) {
          const double bpp = 8.0 + (has_cmask ? 1.0 : 0.0) + (r->ref_has_mask[l] ? 1.0 : 0.0);
          const int passes = std::min(res[f].level_iters[p][l], maxit);
          r->prof.add_bytes(p == 0 ? SSG_PROF_ECC_TRANSLATION : SSG_PROF_ECC_MAIN, passes * bpp * r->lw[l] * r->lh[l]);
        }
      }
      const int checks = (tfirst ? 1 : 0) + (res[f].rejected == SSG_REG_RHO_TRANSLATION ? 0 : 1);
      r->prof.add_bytes(SSG_PROF_CORRELATION, checks * 8.0 * r->lw[0] * r->lh[0]);
    }
  }
}

// ssg_registration_status of n frames from the device records of a batch
void fill_statuses(ssg_registration* r, int n, const FrameXf* xf, const FrameRes* res, ssg_registration_status* st) {
  const bool tfirst = r->o.motion_type != SSG_MOTION_TRANSLATION && r->o.ecc.ecch_estimate_translation_first &&
                      r->o.ecc.ecch_max_level != 0;
  const int np = motion_nparams(r->o.motion_type);
  for (int f = 0; f < n; ++f) {
    ssg_registration_status& o = st[f];
    memset(&o, 0, sizeof(o));
    const FrameRes& R = res[f];
    o.code = R.rejected ? R.rejected : SSG_REG_OK;
    o.ok = R.rejected == 0;
    o.rho_translation = tfirst ? R.rho[0] : 0.0;
    o.rho = R.rejected == SSG_REG_RHO_TRANSLATION ? 0.0 : R.rho[1];
    o.eps = R.eps;
    for (int l = 0; l < r->nlevels; ++l) {
      o.iterations_translation += tfirst ? R.level_iters[0][l] : 0;
      o.iterations += R.level_iters[1][l];
      o.levels_failed += (tfirst ? R.level_failed[0][l] : 0) + R.level_failed[1][l];
      o.level_iterations[l] = R.level_iters[1][l];
    }
    o.nparams = np;
    float p[6];
    if (R.rejected == SSG_REG_RHO_TRANSLATION) {
      // the translation gate returns before the main transform takes the translation estimate
      // (c_frame_registration.cc:842-849): the transform keeps its default (identity) parameters
      const float ident[4][6] = {{0, 0}, {0, 0, 0}, {0, 0, 0, 1}, {1, 0, 0, 0, 1, 0}};
      for (int i = 0; i < 6; ++i) p[i] = ident[r->o.motion_type][i];
    } else {
      for (int i = 0; i < 6; ++i) p[i] = xf[f].p[i];
      if (use_half(r->o)) scale_params(r->o.motion_type, p, 1.0 / r->o.ecc.scale);
    }
    for (int i = 0; i < 6; ++i) o.params[i] = i < np ? p[i] : 0.f;
    params_matrix(r->o.motion_type, p, o.matrix);
  }
}

}  // namespace

extern "C" int ssg_registration_setup_reference_frame(ssg_registration* r, const ssg_image* image, const ssg_image* mask) {
  if (!r) return fail(SSG_E_INVALID_ARG, "null handle");
  CHK(check_image(image, "reference image"));
  DevBuf gref;
  std::vector<ssg_image> gview;
  CHK(ecc_channel_views(r, 1, image, gref, gview, &image));
  if (image->type != SSG_32FC1) return fail(SSG_E_UNSUPPORTED, "reference image: SSG_32FC1 (gray) or SSG_32FC3 (BGR)");
  if (mask) {
    CHK(check_image(mask, "reference mask"));
    if (mask->type != SSG_8UC1 || mask->width != image->width || mask->height != image->height)
      return fail(SSG_E_INVALID_ARG, "reference mask must be SSG_8UC1 of the image size");
  }
  hipStream_t s = r->stream;
  r->have_ref = false;
  r->cap = 0;
  drop_stack_state(r);  // a captured batch graph holds this reference's levels and constants
  r->ref_w = image->width;
  r->ref_h = image->height;
  const bool half = use_half(r->o);
  r->ew = half ? (r->ref_w + 1) / 2 : r->ref_w;
  r->eh = half ? (r->ref_h + 1) / 2 : r->ref_h;
  if (half && mask) return fail(SSG_E_UNSUPPORTED, "reference mask with ecc.scale 0.5");
  // the ECC kernels address a level's samples with 24-bit products of 18-bit tap indices
  if (r->ref_w >= (1 << 17) || r->ref_h >= (1 << 17))
    return fail(SSG_E_UNSUPPORTED, "reference frame %dx%d: sides of 131072 pixels or more", r->ref_w, r->ref_h);
  // level count (ecc2.cc:1009-1026)
  const int min_size = std::max(4, r->o.ecc.ecch_minimum_image_size);
  int tw = r->ew, th = r->eh, lv = 1;
  for (;;) {
    const int maxlevel = r->o.ecc.ecch_max_level;
    if (maxlevel >= 0 && lv >= std::max(0, maxlevel)) break;
    const int nw = ((tw + 1) >> 1) & ~1, nh = ((th + 1) >> 1) & ~1;
    if (nw < min_size || nh < min_size) break;
    tw = nw;
    th = nh;
    if (++lv >= MAX_LEVELS) break;
  }
  r->nlevels = lv;
  long long off = 0, moff = 0;
  int w = r->ew, h = r->eh;
  for (int l = 0; l < lv; ++l) {
    r->lw[l] = w;
    r->lh[l] = h;
    r->lpitch[l] = (w + 63) / 64 * 64;
    r->loff[l] = off;
    off += (long long)r->lpitch[l] * h;
    moff += (long long)r->lpitch[l] * h;
    w = ((w + 1) >> 1) & ~1;
    h = ((h + 1) >> 1) & ~1;
  }
  r->pyr_elems = off;
  r->mask_elems = moff;
  CHK(r->ref_pyr.ensure(sizeof(float) * off));
  CHK(r->frame_ptrs.ensure(sizeof(void*) * 2));
  // level-0 source
  DevBuf tmp;
  const void* dptr;
  size_t pitch_b;
  CHK(device_view(image, tmp, s, &dptr, &pitch_b));
  const float* src = (const float*)dptr;
  int src_pitch = (int)(pitch_b / 4);
  DevBuf halfb;
  if (half) {
    const size_t hp = ((size_t)r->ew * 4 + 255) / 256 * 256;
    CHK(halfb.ensure(hp * r->eh));
    HIPCHK(launch_pyrdown(src, 0, src_pitch, r->ref_w, r->ref_h, halfb.as<float>(), 0, (int)(hp / 4), r->ew, r->eh, 1,
                          BORDER_REFLECT101, s));
    src = halfb.as<const float>();
    src_pitch = (int)(hp / 4);
  }
  DevBuf normb, tmpnm;
  if (use_normalize(r->o)) {
    // ecc_normalize of the reference with its mask (c_frame_registration.cc:624-628)
    std::vector<const uint8_t*> nm;
    size_t nmp = 0;
    if (mask) {
      const void* mptr;
      CHK(device_view(mask, tmpnm, s, &mptr, &nmp));
      nm.push_back((const uint8_t*)mptr);
    }
    const size_t np = ((size_t)r->ew * 4 + 255) / 256 * 256;
    CHK(normb.ensure(np * r->eh));
    CHK(ecc_normalize_frames({src}, src_pitch, nm, (int)nmp, r->ew, r->eh, r->o.ecc.normalization_scale,
                             normb.as<float>(), (int)(np / 4), 0, s));
    src = normb.as<const float>();
    src_pitch = (int)(np / 4);
  }
  HIPCHK(hipMemcpyAsync(r->frame_ptrs.p, &src, sizeof(void*), hipMemcpyHostToDevice, s));
  CHK(build_pyramids(r, r->frame_ptrs.as<const float*>(), src_pitch, 1, r->o.ecc.reference_smooth_sigma,
                     r->ref_pyr.as<float>()));
  for (int l = 0; l < lv; ++l) r->ref_has_mask[l] = false;
  if (mask) {
    DevBuf tmpm;
    const void* mptr;
    size_t mpitch;
    CHK(device_view(mask, tmpm, s, &mptr, &mpitch));
    CHK(r->ref_mask_pyr.ensure((size_t)moff));
    CHK(r->mask_ptrs.ensure(sizeof(void*)));
    HIPCHK(hipMemcpyAsync(r->mask_ptrs.p, &mptr, sizeof(void*), hipMemcpyHostToDevice, s));
    std::vector<char> keep;
    unsigned zeros[MAX_LEVELS] = {};
    CHK(build_mask_pyramids(r, r->mask_ptrs.as<const uint8_t*>(), (int)mpitch, 1, r->ref_mask_pyr.as<uint8_t>(), true, keep,
                            zeros));
    for (int l = 0; l < lv; ++l) {
      r->ref_has_mask[l] = keep[l] != 0;
      r->rma[l] = (double)r->lw[l] * r->lh[l] - zeros[l];
    }
  } else {
    for (int l = 0; l < lv; ++l) r->rma[l] = (double)r->lw[l] * r->lh[l];
  }
  const int method = r->o.ecc.ecc_method;
  if (method == ECC_IC || method == ECC_ICLM) {
    // J precompute on the reference (ecc2.cc:1733-1738, 1985-1991): gradients zeroed off
    // the reference mask and their Hessian moments, per level
    CHK(r->ref_grad.ensure(sizeof(float) * 2 * (size_t)off));
    CHK(r->ref_hm.ensure(sizeof(double) * NMOM_REF * MAX_LEVELS));
    const size_t tiles0 = (size_t)((r->lw[0] + 63) / 64) * ((r->lh[0] + 63) / 64);
    CHK(r->ref_part.ensure(sizeof(double) * NMOM_REF * tiles0));
    for (int l = 0; l < lv; ++l) {
      const uint8_t* m = r->ref_has_mask[l] ? r->ref_mask_pyr.as<uint8_t>() + r->loff[l] : nullptr;
      HIPCHK(launch_ic_reference(r->ref_pyr.as<float>() + r->loff[l], r->lpitch[l], r->lw[l], r->lh[l], m, r->lpitch[l],
                                 r->ref_grad.as<float>() + r->loff[l], r->ref_grad.as<float>() + off + r->loff[l],
                                 r->ref_part.as<double>(), r->ref_hm.as<double>() + (size_t)NMOM_REF * l, s));
    }
    CHK(r->jrec.ensure(sizeof(JRec) * MAX_LEVELS));
    HIPCHK(hipMemsetAsync(r->jrec.p, 0, sizeof(JRec) * MAX_LEVELS, s));
  }
  // level means (centring constants)
  CHK(r->sum_tmp.ensure(sizeof(double) * (r->lh[0] + 1)));
  for (int l = 0; l < lv; ++l) {
    HIPCHK(launch_sum_f32(r->ref_pyr.as<float>() + r->loff[l], r->lpitch[l], r->lw[l], r->lh[l], r->sum_tmp.as<double>(), s));
    double sum = 0;
    HIPCHK(hipMemcpyAsync(&sum, r->sum_tmp.as<double>() + r->lh[l], sizeof(double), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    r->ref_mean[l] = sum / ((double)r->lw[l] * r->lh[l]);
  }
  HIPCHK(hipStreamSynchronize(s));
  // eccflow works on the full-resolution ECC image and mask (c_frame_registration.cc:637-660)
  if (r->flow) CHK(ssg_eccflow_set_reference_image(r->flow, image, mask));
  r->have_status = false;
  r->have_ref = true;
  return SSG_OK;
}

namespace {
int refine_current_remap(ssg_registration* r, const ssg_image* image, const ssg_image* mask);
}

extern "C" int ssg_registration_register_frame(ssg_registration* r, const ssg_image* image, const ssg_image* mask,
                                               ssg_registration_status* status) {
  if (!r) return fail(SSG_E_INVALID_ARG, "null handle");
  r->have_status = false;
  CHK(register_n(r, 1, image, mask, status, true));
  r->last = *status;
  // create_remap(reference size) then the eccflow refine (c_frame_registration.cc:883-917) on the
  // registration channel (register_n left it in gray_buf for colour frames)
  ssg_image gv{};
  if (image && image->type == SSG_32FC3) {
    gv = *image;
    gv.data = r->gray_buf.p;
    gv.step = (int64_t)(((size_t)image->width * 4 + 255) / 256 * 256);
    gv.type = SSG_32FC1;
    gv.on_device = 1;
    image = &gv;
  }
  if (r->flow && status->ok) CHK(refine_current_remap(r, image, mask));
  r->have_status = true;
  return SSG_OK;
}

extern "C" int ssg_registration_register_batch(ssg_registration* r, int n, const ssg_image* frames,
                                               const ssg_image* masks, ssg_registration_status* status) {
  if (!r) return fail(SSG_E_INVALID_ARG, "null handle");
  if (n < 0) return fail(SSG_E_INVALID_ARG, "n < 0");
  return register_n(r, n, frames, masks, status, false);
}

extern "C" int ssg_registration_create_remap(ssg_registration* r, const ssg_registration_status* st, ssg_image* map) {
  if (!r || !st) return fail(SSG_E_INVALID_ARG, "null argument");
  if (!r->have_ref) return fail(SSG_E_NOT_READY, "no reference frame");
  CHK(check_image(map, "map"));
  if (map->type != SSG_32FC2 || map->width != r->ref_w || map->height != r->ref_h)
    return fail(SSG_E_INVALID_ARG, "map must be SSG_32FC2 of the reference size %dx%d", r->ref_w, r->ref_h);
  const MapCoef mc = map_coef(r->o.motion_type, st->params, nullptr);
  hipStream_t s = r->stream;
  DevBuf tmp;
  float* dst;
  size_t pitch;
  if (map->on_device) {
    dst = (float*)map->data;
    pitch = (size_t)map->step;
  } else {
    pitch = ((size_t)map->width * 8 + 255) / 256 * 256;
    CHK(tmp.ensure(pitch * map->height));
    dst = tmp.as<float>();
  }
  HIPCHK(launch_create_map(r->o.motion_type, mc, dst, (int)(pitch / 4), map->width, map->height, s));
  if (!map->on_device)
    HIPCHK(hipMemcpy2DAsync(map->data, (size_t)map->step, dst, pitch, (size_t)map->width * 8, map->height,
                            hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  return SSG_OK;
}

extern "C" int ssg_remap(const ssg_image* src, const ssg_image* src_mask, const ssg_image* map, const double* matrix,
                         int32_t motion_type, const float* params, ssg_image* dst, ssg_image* dst_mask,
                         int32_t interpolation, int32_t border_mode, double border_value) {
  (void)matrix;
  if (!dst && !dst_mask) return SSG_OK;
  if (!map && !params) return fail(SSG_E_INVALID_ARG, "need a map field or transform parameters");
  if (interpolation != SSG_INTER_LINEAR && interpolation != SSG_INTER_CUBIC && interpolation != SSG_INTER_NEAREST)
    return fail(SSG_E_UNSUPPORTED, "interpolation %d", interpolation);
  hipStream_t s = 0;
  HIPCHK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
  struct StreamGuard {
    hipStream_t s;
    ~StreamGuard() { (void)hipStreamSynchronize(s); (void)hipStreamDestroy(s); }
  } guard{s};
  CHK(ensure_tables(s));
  int dw = dst ? dst->width : dst_mask->width, dh = dst ? dst->height : dst_mask->height;
  int motion = map ? -1 : motion_type;
  MapCoef mc = map ? MapCoef{} : map_coef(motion_type, params, nullptr);
  DevBuf tmap;
  const float* dmap = nullptr;
  int map_pitch = 0;
  if (map) {
    CHK(check_image(map, "map"));
    if (map->type != SSG_32FC2 || map->width != dw || map->height != dh)
      return fail(SSG_E_INVALID_ARG, "map must be SSG_32FC2 of the destination size");
    const void* p;
    size_t pb;
    CHK(device_view(map, tmap, s, &p, &pb));
    dmap = (const float*)p;
    map_pitch = (int)(pb / 4);
  }
  if (dst) {
    CHK(check_image(src, "src"));
    CHK(check_image(dst, "dst"));
    if ((src->type != SSG_32FC1 && src->type != SSG_32FC3) || dst->type != src->type)
      return fail(SSG_E_UNSUPPORTED, "remap: SSG_32FC1/SSG_32FC3 images of the same type");
    const int cn = channels_of(src->type);
    DevBuf tsrc, tdst;
    const void* sp;
    size_t spb;
    CHK(device_view(src, tsrc, s, &sp, &spb));
    float* dp;
    size_t dpb;
    if (dst->on_device) {
      dp = (float*)dst->data;
      dpb = (size_t)dst->step;
    } else {
      dpb = ((size_t)dw * 4 * cn + 255) / 256 * 256;
      CHK(tdst.ensure(dpb * dh));
      dp = tdst.as<float>();
      // TRANSPARENT keeps the destination: start from its current contents
      HIPCHK(hipMemcpy2DAsync(dp, dpb, dst->data, (size_t)dst->step, (size_t)dw * 4 * cn, dh, hipMemcpyHostToDevice, s));
    }
    HIPCHK(launch_remap(motion, (const float*)sp, (int)(spb / 4), src->width, src->height, cn, mc, dmap, map_pitch, dp,
                        (int)(dpb / 4), dw, dh, interpolation, border_mode, (float)border_value, s));
    if (!dst->on_device)
      HIPCHK(hipMemcpy2DAsync(dst->data, (size_t)dst->step, dp, dpb, (size_t)dw * 4 * cn, dh, hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
  }
  if (dst_mask) {
    CHK(check_image(dst_mask, "dst_mask"));
    if (dst_mask->type != SSG_8UC1) return fail(SSG_E_INVALID_ARG, "dst_mask must be SSG_8UC1");
    int sw = src ? src->width : 0, sh = src ? src->height : 0;
    const uint8_t* smp = nullptr;
    int smpitch = 0;
    DevBuf tsm;
    if (src_mask) {
      CHK(check_image(src_mask, "src_mask"));
      sw = src_mask->width;
      sh = src_mask->height;
      const void* p;
      size_t pb;
      CHK(device_view(src_mask, tsm, s, &p, &pb));
      smp = (const uint8_t*)p;
      smpitch = (int)pb;
    }
    if (sw <= 0) return fail(SSG_E_INVALID_ARG, "mask source size unknown");
    DevBuf m1, m2;
    const size_t mp = ((size_t)dw + 255) / 256 * 256;
    CHK(m1.ensure(mp * dh));
    CHK(m2.ensure(mp * dh));
    HIPCHK(launch_remap_mask(motion, smp, smpitch, sw, sh, mc, dmap, map_pitch, m1.as<uint8_t>(), (int)mp, dw, dh, s,
                             interpolation == SSG_INTER_CUBIC ? INTER_CUBIC : INTER_LINEAR));
    HIPCHK(launch_erode5(m1.as<uint8_t>(), 0, (int)mp, m2.as<uint8_t>(), 0, (int)mp, dw, dh, 1, BORDER_CONSTANT, 255, s));
    HIPCHK(hipMemcpy2DAsync(dst_mask->data, (size_t)dst_mask->step, m2.p, mp, dw, dh,
                            dst_mask->on_device ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
  }
  return SSG_OK;
}

// =====================================================================================================
// accumulation
// =====================================================================================================
struct ssg_accumulation {
  int method = SSG_ACC_WEIGHTED_AVERAGE;
  int w = 0, h = 0, cn = 1;
  long long plane = 0;
  hipStream_t stream = nullptr;
  DevBuf acc, cnt, part, part_cnt, maps, weights, valid, frame_ptrs, tmp;
  int accumulated = 0;
  Profiler prof;
  // c_bayer_average
  int colorid = 0;
  unsigned pattern = 0;
  int bpp = 16;
  DevBuf rmap;
  int rmap_pitch = 0;  // floats per row; 0 = no remap set
  // c_laplacian_pyramid_focus_stacking: level-major pyramids, the channel planes of a
  // level contiguous: element offset of (level i, channel c) = cn * loff[i] + c * n_i
  ssg_focus_stacking_options fo{};
  std::vector<int> lw, lh;
  std::vector<long long> loff;
  long long P = 0;  // elements of one single-plane pyramid
  int L = 0, wc = 1;
  DevBuf facc, fw, fp, fup, ft0, ft1, ft2, ft3, fgray, fm, fW, fptrs, fstd;
  DevBuf fbpyr, fbup, fbW;  // batch planes of add_frames, kept for the next batch
  DevBuf fms, fmtab;        // ssg_focus_stack_batch: chunk scratch and job tables (first handle)
  std::vector<const float*> hptrs;
  int slot = 0;
  SepKernel G{}, Wk{};
  bool hasG = false;
  bool bayer() const { return method == SSG_ACC_BAYER_AVERAGE; }
  bool focus() const { return method == SSG_ACC_FOCUS_STACK; }
  // accumulator planes: weighted average (2*cn + 2) f32 planes; bayer acc[3] + cntr[3];
  // focus stacking keeps its own pyramids
  size_t acc_bytes() const { return sizeof(float) * (focus() ? 1 : plane * (bayer() ? 6 : 2 * cn + 2)); }
  size_t cnt_bytes() const { return sizeof(unsigned) * (focus() ? 1 : plane * (bayer() ? 3 : 1)); }
  BayerState bayer_state() {
    BayerState st;
    st.acc = acc.as<float>();
    st.cntr = acc.as<float>() + 3 * plane;
    st.cnt = cnt.as<unsigned>();
    st.plane = plane;
    st.w = w;
    st.h = h;
    st.pattern = pattern;
    return st;
  }
};

extern "C" int ssg_accumulation_create(int32_t method, int32_t width, int32_t height, int32_t channels,
                                       ssg_accumulation** out) {
  if (!out) return fail(SSG_E_INVALID_ARG, "out is null");
  *out = nullptr;
  if (method != SSG_ACC_WEIGHTED_AVERAGE && method != SSG_ACC_AVERAGE && method != SSG_ACC_BAYER_AVERAGE &&
      method != SSG_ACC_FOCUS_STACK)
    return fail(SSG_E_UNSUPPORTED, "accumulation method %d", method);
  if (method == SSG_ACC_FOCUS_STACK && (width < 5 || height < 5))
    return fail(SSG_E_INVALID_ARG, "focus stacking needs frames of at least 5x5");
  if (width <= 0 || height <= 0 || (channels != 1 && channels != 3))
    return fail(SSG_E_INVALID_ARG, "accumulator %dx%dx%d", width, height, channels);
  if (method == SSG_ACC_BAYER_AVERAGE && channels != 1)
    return fail(SSG_E_INVALID_ARG, "a bayer accumulator takes 1-channel raw frames");
  ssg_accumulation* a = new ssg_accumulation();
  a->method = method;
  a->w = width;
  a->h = height;
  a->cn = channels;
  a->plane = (long long)width * height;
  ssg_default_focus_stacking_options(&a->fo);
  hipError_t e = hipStreamCreateWithFlags(&a->stream, hipStreamNonBlocking);
  if (e != hipSuccess) {
    delete a;
    return fail(SSG_E_HIP, "hipStreamCreate: %s", hipGetErrorString(e));
  }
  int rc = a->acc.ensure(a->acc_bytes());
  if (rc == SSG_OK) rc = a->cnt.ensure(a->cnt_bytes());
  if (rc == SSG_OK) rc = ensure_tables(a->stream);
  if (rc != SSG_OK) {
    (void)hipStreamDestroy(a->stream);
    delete a;
    return rc;
  }
  *out = a;
  return ssg_accumulation_clear(a);
}

extern "C" void ssg_accumulation_destroy(ssg_accumulation* a) {
  if (!a) return;
  if (a->stream) {
    (void)hipStreamSynchronize(a->stream);
    (void)hipStreamDestroy(a->stream);
  }
  delete a;
}

extern "C" int ssg_accumulation_clear(ssg_accumulation* a) {
  if (!a) return fail(SSG_E_INVALID_ARG, "null handle");
  HIPCHK(hipMemsetAsync(a->acc.p, 0, a->acc_bytes(), a->stream));
  HIPCHK(hipMemsetAsync(a->cnt.p, 0, a->cnt_bytes(), a->stream));
  HIPCHK(hipStreamSynchronize(a->stream));
  a->accumulated = 0;
  a->rmap_pitch = 0;  // c_bayer_average::clear releases the remap (:1154-1161)
  return SSG_OK;
}

extern "C" int ssg_accumulation_set_bayer_pattern(ssg_accumulation* a, int32_t colorid) {
  if (!a) return fail(SSG_E_INVALID_ARG, "null handle");
  if (!a->bayer()) return fail(SSG_E_INVALID_ARG, "not a bayer accumulator");
  const unsigned bits = bayer_pattern_bits(colorid);
  if (!bits) return fail(SSG_E_INVALID_ARG, "colorid %d is not a bayer pattern", colorid);
  a->colorid = colorid;
  a->pattern = bits;
  return SSG_OK;
}

extern "C" int ssg_accumulation_set_remap(ssg_accumulation* a, const ssg_image* map) {
  if (!a) return fail(SSG_E_INVALID_ARG, "null handle");
  if (!a->bayer()) return fail(SSG_E_INVALID_ARG, "not a bayer accumulator");
  if (!map) {
    a->rmap_pitch = 0;
    return SSG_OK;
  }
  CHK(check_image(map, "map"));
  if (map->type != SSG_32FC2 || map->width != a->w || map->height != a->h)
    return fail(SSG_E_INVALID_ARG, "remap must be SSG_32FC2 of the accumulator size (%dx%d)", a->w, a->h);
  const size_t row = (size_t)a->w * 8;
  CHK(a->rmap.ensure(row * a->h));
  HIPCHK(hipMemcpy2DAsync(a->rmap.p, row, map->data, (size_t)map->step, row, a->h,
                          map->on_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, a->stream));
  HIPCHK(hipStreamSynchronize(a->stream));
  a->rmap_pitch = 2 * a->w;
  return SSG_OK;
}

extern "C" int ssg_accumulation_set_input_bpp(ssg_accumulation* a, int32_t bpp) {
  if (!a) return fail(SSG_E_INVALID_ARG, "null handle");
  if (bpp < 1 || bpp > 16) return fail(SSG_E_INVALID_ARG, "bpp %d", bpp);
  a->bpp = bpp;
  return SSG_OK;
}

namespace {
int bayer_add(ssg_accumulation* a, const ssg_image* src, const ssg_image* weights) {
  if (!a->pattern) return fail(SSG_E_NOT_READY, "no bayer pattern set");
  if (src->type != SSG_32FC1 && src->type != SSG_16UC1)
    return fail(SSG_E_INVALID_ARG, "bayer frames are SSG_32FC1 or SSG_16UC1 raw images");
  hipStream_t s = a->stream;
  DevBuf ts, tw;
  const void* sp;
  size_t spb;
  CHK(device_view(src, ts, s, &sp, &spb));
  const uint8_t* mp = nullptr;
  const float* wp = nullptr;
  size_t wpb = 0;
  if (weights) {
    CHK(check_image(weights, "weights"));
    if (weights->width != a->w || weights->height != a->h) return fail(SSG_E_INVALID_ARG, "weights size");
    const void* p;
    CHK(device_view(weights, tw, s, &p, &wpb));
    if (weights->type == SSG_8UC1) mp = (const uint8_t*)p;
    else if (weights->type == SSG_32FC1) wp = (const float*)p;
    else return fail(SSG_E_INVALID_ARG, "weights must be SSG_8UC1 or SSG_32FC1");
  }
  const bool u16 = src->type == SSG_16UC1;
  HIPCHK(launch_bayer_add(a->bayer_state(), sp, (int)(spb / (u16 ? 2 : 4)), u16, 1.f / (float)(1 << a->bpp),
                          a->rmap_pitch ? a->rmap.as<float>() : nullptr, a->rmap_pitch, mp, (int)wpb, wp,
                          (int)(wpb / 4), s));
  HIPCHK(hipStreamSynchronize(s));
  ++a->accumulated;
  return SSG_OK;
}
}  // namespace

extern "C" void ssg_default_focus_stacking_options(ssg_focus_stacking_options* o) {
  if (!o) return;
  *o = ssg_focus_stacking_options{};
  o->fusing_policy = SSG_FOCUS_SELECT_MAX_ENERGY;  // c_frame_accumulation.h:156-161
  o->inpaint_mask_holes = 1;
  o->avgchannel = 1;
  o->kradius = 0;
  o->ksigma = 0;
  o->lpg_k = 6;
  o->lpg_p = 2;
  o->lpg_sigma = 5;
  o->max_levels = 0;
}

extern "C" int ssg_accumulation_set_focus_options(ssg_accumulation* a, const ssg_focus_stacking_options* o) {
  if (!a || !o) return fail(SSG_E_INVALID_ARG, "null argument");
  if (!a->focus()) return fail(SSG_E_INVALID_ARG, "not a focus stacking accumulator");
  if (a->accumulated) return fail(SSG_E_INVALID_ARG, "options are set before the first add()");
  if (o->fusing_policy < SSG_FOCUS_SELECT_MAX_ENERGY || o->fusing_policy > SSG_FOCUS_LPG_WEIGHTS)
    return fail(SSG_E_INVALID_ARG, "fusing_policy %d", o->fusing_policy);
  if (o->max_levels < 0 || o->kradius < 0) return fail(SSG_E_INVALID_ARG, "max_levels / kradius");
  if (o->fusing_policy == SSG_FOCUS_LPG_WEIGHTS && !(o->lpg_sigma > 0))
    return fail(SSG_E_INVALID_ARG, "lpg_sigma must be > 0");
  a->fo = *o;
  return SSG_OK;
}

namespace {
// cv::getGaussianKernel (fixed tables for sigma <= 0 and n <= 7), float taps
void gaussian_taps(int n, double sigma, float* out) {
  static const float small_tab[4][7] = {{1.f},
                                        {0.25f, 0.5f, 0.25f},
                                        {0.0625f, 0.25f, 0.375f, 0.25f, 0.0625f},
                                        {0.03125f, 0.109375f, 0.21875f, 0.28125f, 0.21875f, 0.109375f, 0.03125f}};
  const float* fixed = (n % 2 == 1 && n <= 7 && sigma <= 0) ? small_tab[n >> 1] : nullptr;
  const double sx = sigma > 0 ? sigma : ((n - 1) * 0.5 - 1) * 0.3 + 0.8;
  const double scale2X = -0.5 / (sx * sx);
  std::vector<double> cd(n);
  double sum = 0;
  for (int i = 0; i < n; ++i) {
    const double x = i - (n - 1) * 0.5;
    cd[i] = fixed ? (double)fixed[i] : exp(scale2X * x * x);
    sum += cd[i];
  }
  sum = 1. / sum;
  for (int i = 0; i < n; ++i) out[i] = (float)(cd[i] * sum);
}

SepKernel square_sep(const float* k, int n, float delta) {
  SepKernel K{};
  for (int i = 0; i < n; ++i) K.kx[i] = K.ky[i] = k[i];
  K.nx = K.ny = n;
  bool sym = true, asym = true;
  for (int i = 0; i < n; ++i) {
    if ((double)k[i] != (double)k[n - 1 - i]) sym = false;
    if ((double)k[i] != -(double)k[n - 1 - i]) asym = false;
  }
  K.tx = K.ty = sym ? 1 : (asym ? 2 : 0);
  K.delta = delta;
  return K;
}

// separable filter of nplanes dense planes (w x h) from src into dst (REFLECT101)
int blur_planes(ssg_accumulation* a, const float* src, float* dst, int w, int h, int nplanes, const SepKernel& K) {
  const size_t n = (size_t)w * h;
  if (a->slot + nplanes > (int)a->hptrs.size()) return fail(SSG_E_INVALID_ARG, "too many filter passes");
  for (int c = 0; c < nplanes; ++c) a->hptrs[a->slot + c] = src + c * n;
  const float** dp = a->fptrs.as<const float*>() + a->slot;
  HIPCHK(hipMemcpyAsync((void*)dp, &a->hptrs[a->slot], sizeof(float*) * nplanes, hipMemcpyHostToDevice, a->stream));
  a->slot += nplanes;
  HIPCHK(launch_sepfilter(dp, w, dst, (long long)n, w, w, h, nplanes, K, SSG_BORDER_REFLECT101, a->stream));
  return SSG_OK;
}

int focus_prepare(ssg_accumulation* a) {
  const ssg_focus_stacking_options& o = a->fo;
  a->lw.assign(1, a->w);
  a->lh.assign(1, a->h);
  for (;;) {
    const int nw = (a->lw.back() + 1) / 2, nh = (a->lh.back() + 1) / 2;
    if (std::min(nw, nh) <= 8 || (o.max_levels && (int)a->lw.size() >= o.max_levels)) break;
    a->lw.push_back(nw);
    a->lh.push_back(nh);
  }
  a->L = (int)a->lw.size();
  a->loff.assign(a->L + 1, 0);
  for (int i = 0; i < a->L; ++i) a->loff[i + 1] = a->loff[i] + (long long)a->lw[i] * a->lh[i];
  a->P = a->loff[a->L];
  const int cn = a->cn;
  a->wc = (o.fusing_policy == SSG_FOCUS_WEIGHTED_AVERAGE && o.avgchannel && cn > 1) ? 1 : cn;
  const size_t n0 = (size_t)a->w * a->h;
  CHK(a->facc.ensure(sizeof(float) * cn * a->P));
  CHK(a->fp.ensure(sizeof(float) * cn * a->P));
  CHK(a->fup.ensure(sizeof(float) * cn * n0));
  if (o.fusing_policy != SSG_FOCUS_LPG_WEIGHTS) {  // energy scratch of the reference's own policies
    CHK(a->ft0.ensure(sizeof(float) * cn * n0));
    CHK(a->ft1.ensure(sizeof(float) * cn * n0));
    CHK(a->ft2.ensure(sizeof(float) * cn * n0));
    CHK(a->ft3.ensure(sizeof(float) * cn * n0));
  }
  CHK(a->fstd.ensure(sizeof(double) * 2 * cn));
  CHK(a->fptrs.ensure(sizeof(float*) * 256));
  a->hptrs.assign(256, nullptr);
  if (o.fusing_policy == SSG_FOCUS_LPG_WEIGHTS) {
    CHK(a->fw.ensure(sizeof(float) * a->P));
    CHK(a->fW.ensure(sizeof(float) * a->P));
    CHK(a->fgray.ensure(sizeof(float) * n0));
    CHK(a->fm.ensure(sizeof(float) * n0));
    HIPCHK(hipMemsetAsync(a->facc.p, 0, sizeof(float) * cn * a->P, a->stream));
    HIPCHK(hipMemsetAsync(a->fw.p, 0, sizeof(float) * a->P, a->stream));
    const int nk = (int)std::floor(o.lpg_sigma * 8 + 1 + 0.5) | 1;  // cv::GaussianBlur on float images
    if (nk > SEP_MAX_TAPS) return fail(SSG_E_UNSUPPORTED, "lpg_sigma %g needs %d taps (max %d)", o.lpg_sigma, nk,
                                       SEP_MAX_TAPS);
    std::vector<float> t(nk);
    gaussian_taps(nk, o.lpg_sigma, t.data());
    a->Wk = square_sep(t.data(), nk, 0.f);
  } else if (o.fusing_policy == SSG_FOCUS_WEIGHTED_AVERAGE) {
    CHK(a->fw.ensure(sizeof(float) * a->wc * a->P));
  }
  a->hasG = o.ksigma > 0 || o.kradius > 0;
  if (a->hasG) {
    const int nk = std::max(0, 2 * o.kradius + 1);
    if (nk > SEP_MAX_TAPS) return fail(SSG_E_UNSUPPORTED, "kradius %d", o.kradius);
    std::vector<float> t(nk);
    gaussian_taps(nk, std::max(0.0, o.ksigma), t.data());
    a->G = square_sep(t.data(), nk, 0.f);
  }
  return SSG_OK;
}

// c_laplacian_pyramid_focus_stacking::add (c_frame_accumulation.cc:491-627)
int focus_add(ssg_accumulation* a, const ssg_image* src) {
  const int cn = a->cn;
  if (src->type != (cn == 3 ? SSG_32FC3 : SSG_32FC1))
    return fail(SSG_E_INVALID_ARG, "focus frames are f32 with the accumulator's %d channels", cn);
  const bool first = a->accumulated == 0;
  if (first) CHK(focus_prepare(a));
  hipStream_t s = a->stream;
  a->slot = 0;
  const ssg_focus_stacking_options& o = a->fo;
  DevBuf ts;
  const void* sp;
  size_t spb;
  CHK(device_view(src, ts, s, &sp, &spb));
  float* fp = a->fp.as<float>();
  float* acc = a->facc.as<float>();
  const size_t n0 = (size_t)a->w * a->h;
  auto lvl = [&](float* base, int i, int planes) { return base + (size_t)planes * a->loff[i]; };
  auto nl = [&](int i) { return (size_t)a->lw[i] * a->lh[i]; };
  HIPCHK(launch_split((const float*)sp, (int)(spb / 4), a->w, a->h, cn, fp, s));
  if (o.fusing_policy == SSG_FOCUS_LPG_WEIGHTS) {
    // W = gaussian(lpg(gray)), reduced to every level
    const double k = o.lpg_k;
    const float l2 = (100.f / 4.f) * (100.f / 4.f), g2 = (100.f / 36.f) * (100.f / 36.f);
    const float alpha = (float)(k / (k + 1)) * l2, beta = (float)(1. / (k + 1)) * g2;
    HIPCHK(launch_gray(fp, n0, cn, a->fgray.as<float>(), s));
    HIPCHK(launch_lpg(a->fgray.as<float>(), a->w, a->h, alpha, beta, 1e-9f, (float)o.lpg_p, a->fm.as<float>(), s));
    CHK(blur_planes(a, a->fm.as<float>(), a->fW.as<float>(), a->w, a->h, 1, a->Wk));
    for (int i = 0; i + 1 < a->L; ++i)
      HIPCHK(launch_pyrdown(lvl(a->fW.as<float>(), i, 1), (long long)nl(i), a->lw[i], a->lw[i], a->lh[i],
                            lvl(a->fW.as<float>(), i + 1, 1), (long long)nl(i + 1), a->lw[i + 1], a->lw[i + 1],
                            a->lh[i + 1], 1, SSG_BORDER_REFLECT101, s));
  }
  // build_laplacian_pyramid (laplacian_pyramid.cc:44-88)
  for (int i = 0; i + 1 < a->L; ++i) {
    HIPCHK(launch_pyrdown(lvl(fp, i, cn), (long long)nl(i), a->lw[i], a->lw[i], a->lh[i], lvl(fp, i + 1, cn),
                          (long long)nl(i + 1), a->lw[i + 1], a->lw[i + 1], a->lh[i + 1], cn, SSG_BORDER_REFLECT101, s));
    HIPCHK(launch_pyrup(lvl(fp, i + 1, cn), (long long)nl(i + 1), a->lw[i + 1], a->lh[i + 1], a->fup.as<float>(),
                        (long long)nl(i), a->lw[i], a->lh[i], cn, s));
    HIPCHK(launch_ew(0, lvl(fp, i, cn), a->fup.as<float>(), lvl(fp, i, cn), cn * nl(i), s));
  }
  const int top = a->L - 1;
  if (o.fusing_policy == SSG_FOCUS_LPG_WEIGHTS) {
    for (int i = 0; i < a->L; ++i)
      HIPCHK(launch_wacc(lvl(acc, i, cn), lvl(a->fw.as<float>(), i, 1), lvl(a->fW.as<float>(), i, 1), lvl(fp, i, cn),
                         nl(i), cn, s));
  } else {
    if (!first) {
      // top level: blend weighted by the channel-summed standard deviations (:579-588)
      HIPCHK(launch_plane_stdev(lvl(acc, top, cn), nl(top), cn, a->fstd.as<double>(), s));
      HIPCHK(launch_plane_stdev(lvl(fp, top, cn), nl(top), cn, a->fstd.as<double>() + cn, s));
      double sd[6];
      HIPCHK(hipMemcpyAsync(sd, a->fstd.p, sizeof(double) * 2 * cn, hipMemcpyDeviceToHost, s));
      HIPCHK(hipStreamSynchronize(s));
      double s0 = 0, s1 = 0;
      for (int c = 0; c < cn; ++c) {
        s0 += sd[c];
        s1 += sd[cn + c];
      }
      HIPCHK(launch_blend(lvl(acc, top, cn), lvl(fp, top, cn), (float)(s0 / (s0 + s1)), (float)(s1 / (s0 + s1)),
                          cn * nl(top), s));
    } else {
      HIPCHK(hipMemcpyAsync(lvl(acc, top, cn), lvl(fp, top, cn), sizeof(float) * cn * nl(top),
                            hipMemcpyDeviceToDevice, s));
    }
    for (int i = 0; i < top; ++i) {
      float* L = lvl(fp, i, cn);
      float* A = lvl(acc, i, cn);
      const size_t n = nl(i);
      if (o.fusing_policy == SSG_FOCUS_SELECT_MAX_ENERGY) {
        if (first) {
          HIPCHK(hipMemcpyAsync(A, L, sizeof(float) * cn * n, hipMemcpyDeviceToDevice, s));
          continue;
        }
        float* w0 = a->ft0.as<float>();
        float* w1 = a->ft1.as<float>();
        HIPCHK(launch_ew(2, A, nullptr, w0, cn * n, s));
        HIPCHK(launch_ew(2, L, nullptr, w1, cn * n, s));
        if (a->hasG) {
          CHK(blur_planes(a, w0, a->ft2.as<float>(), a->lw[i], a->lh[i], cn, a->G));
          CHK(blur_planes(a, w1, a->ft3.as<float>(), a->lw[i], a->lh[i], cn, a->G));
          w0 = a->ft2.as<float>();
          w1 = a->ft3.as<float>();
        }
        HIPCHK(launch_select(A, L, w0, w1, cn * n, s));
      } else {
        // compute_energy (:511-531), then acc += lap*w, wwp += w
        float* ww = a->ft0.as<float>();
        HIPCHK(launch_energy(L, ww, n, cn, o.avgchannel, s));
        if (a->hasG) {
          SepKernel K = a->G;
          K.delta = 1e-12f;
          CHK(blur_planes(a, ww, a->ft1.as<float>(), a->lw[i], a->lh[i], a->wc, K));
          ww = a->ft1.as<float>();
        }
        HIPCHK(launch_mulw(L, ww, a->wc, n, cn, s));
        float* Wp = lvl(a->fw.as<float>(), i, a->wc);
        if (first) {
          HIPCHK(hipMemcpyAsync(A, L, sizeof(float) * cn * n, hipMemcpyDeviceToDevice, s));
          HIPCHK(hipMemcpyAsync(Wp, ww, sizeof(float) * a->wc * n, hipMemcpyDeviceToDevice, s));
        } else {
          HIPCHK(launch_ew(1, L, A, A, cn * n, s));
          HIPCHK(launch_ew(1, ww, Wp, Wp, a->wc * n, s));
        }
      }
    }
  }
  HIPCHK(hipStreamSynchronize(s));
  ++a->accumulated;
  return SSG_OK;
}

// n successive add() calls of a focus stack in batches of m frames: the laplacian pyramids of a
// batch are built together (level planes [frame][channel][pixel], one launch per level and
// step), and each level is fused frame after frame per pixel by one kernel; the results are
// those of n single adds.  Smoothed energies (kradius/ksigma) couple neighbouring pixels between
// frames, so that configuration adds frame by frame.
int focus_add_frames(ssg_accumulation* a, int n, const ssg_image* frames) {
  const int cn = a->cn;
  for (int f = 0; f < n; ++f) {
    CHK(check_image(&frames[f], "frame"));
    if (frames[f].type != (cn == 3 ? SSG_32FC3 : SSG_32FC1) || frames[f].width != a->w || frames[f].height != a->h)
      return fail(SSG_E_INVALID_ARG, "frame %d: f32 with the accumulator's size and %d channels", f, cn);
  }
  if (a->accumulated == 0) CHK(focus_prepare(a));
  const ssg_focus_stacking_options& o = a->fo;
  if (a->hasG && o.fusing_policy != SSG_FOCUS_LPG_WEIGHTS) {
    for (int f = 0; f < n; ++f) CHK(focus_add(a, &frames[f]));
    return SSG_OK;
  }
  hipStream_t s = a->stream;
  const size_t n0 = (size_t)a->w * a->h;
  const bool lpgw = o.fusing_policy == SSG_FOCUS_LPG_WEIGHTS;
  const size_t per = sizeof(float) * ((size_t)cn * a->P + (size_t)cn * n0 + (lpgw ? a->P : 0));
  const int chunk = (int)std::max<size_t>(1, std::min<size_t>({(size_t)n, (size_t)64, ((size_t)2 << 30) / per}));
  DevBuf &pyr = a->fbpyr, &up = a->fbup, &W = a->fbW;
  CHK(pyr.ensure(sizeof(float) * (size_t)cn * a->P * chunk));
  CHK(up.ensure(sizeof(float) * (size_t)cn * n0 * chunk));
  if (lpgw) CHK(W.ensure(sizeof(float) * a->P * chunk));
  auto nl = [&](int i) { return (size_t)a->lw[i] * a->lh[i]; };
  float* acc = a->facc.as<float>();
  const int top = a->L - 1;
  for (int f0 = 0; f0 < n; f0 += chunk) {
    const int m = std::min(chunk, n - f0);
    const bool first = a->accumulated == 0;
    // level i of the batch: m * cn planes of nl(i) at pyr + m * cn * loff[i]
    auto lv = [&](int i) { return pyr.as<float>() + (size_t)m * cn * a->loff[i]; };
    auto lw1 = [&](int i) { return W.as<float>() + (size_t)m * a->loff[i]; };
    a->slot = 0;
    for (int k = 0; k < m; ++k) {
      DevBuf ts;
      const void* sp;
      size_t spb;
      CHK(device_view(&frames[f0 + k], ts, s, &sp, &spb));
      float* planes = lv(0) + (size_t)k * cn * n0;
      HIPCHK(launch_split((const float*)sp, (int)(spb / 4), a->w, a->h, cn, planes, s));
      if (lpgw) {
        // lpg(gray) of frame k into plane k of the pyrup scratch, blurred below for the whole batch
        const double kk = o.lpg_k;
        const float l2 = (100.f / 4.f) * (100.f / 4.f), g2 = (100.f / 36.f) * (100.f / 36.f);
        const float alpha = (float)(kk / (kk + 1)) * l2, beta = (float)(1. / (kk + 1)) * g2;
        HIPCHK(launch_gray(planes, n0, cn, a->fgray.as<float>(), s));
        HIPCHK(launch_lpg(a->fgray.as<float>(), a->w, a->h, alpha, beta, 1e-9f, (float)o.lpg_p,
                          up.as<float>() + (size_t)k * n0, s));
      }
      if (!frames[f0 + k].on_device) HIPCHK(hipStreamSynchronize(s));  // the staging copy goes out of scope
    }
    if (lpgw) CHK(blur_planes(a, up.as<float>(), lw1(0), a->w, a->h, m, a->Wk));
    for (int i = 0; i + 1 < a->L; ++i) {
      if (lpgw)
        HIPCHK(launch_pyrdown(lw1(i), (long long)nl(i), a->lw[i], a->lw[i], a->lh[i], lw1(i + 1), (long long)nl(i + 1),
                              a->lw[i + 1], a->lw[i + 1], a->lh[i + 1], m, SSG_BORDER_REFLECT101, s));
      HIPCHK(launch_pyrdown(lv(i), (long long)nl(i), a->lw[i], a->lw[i], a->lh[i], lv(i + 1), (long long)nl(i + 1),
                            a->lw[i + 1], a->lw[i + 1], a->lh[i + 1], m * cn, SSG_BORDER_REFLECT101, s));
      HIPCHK(launch_pyrup(lv(i + 1), (long long)nl(i + 1), a->lw[i + 1], a->lh[i + 1], up.as<float>(), (long long)nl(i),
                          a->lw[i], a->lh[i], m * cn, s));
      HIPCHK(launch_ew(0, lv(i), up.as<float>(), lv(i), (size_t)m * cn * nl(i), s));
    }
    for (int i = 0; i < a->L; ++i) {
      float* A = acc + (size_t)cn * a->loff[i];
      if (lpgw) {
        HIPCHK(launch_wacc_frames(A, a->fw.as<float>() + a->loff[i], lw1(i), lv(i), nl(i), cn, m, s));
      } else if (i == top) {
        HIPCHK(launch_top_frames(A, lv(i), nl(i), cn, m, first, s));
      } else if (o.fusing_policy == SSG_FOCUS_SELECT_MAX_ENERGY) {
        HIPCHK(launch_select_frames(A, lv(i), (size_t)cn * nl(i), (size_t)cn * nl(i), m, first, s));
      } else {
        HIPCHK(launch_wavg_frames(A, a->fw.as<float>() + (size_t)a->wc * a->loff[i], lv(i), nl(i), cn, o.avgchannel,
                                  a->wc, m, first, s));
      }
    }
    HIPCHK(hipStreamSynchronize(s));
    a->accumulated += m;
  }
  return SSG_OK;
}

// c_laplacian_pyramid_focus_stacking::compute (:631-662) + reconstruct_laplacian_pyramid (:90-127):
// the reconstructed image is left in level 0 of fp
int focus_compute(ssg_accumulation* a) {
  const int cn = a->cn;
  hipStream_t s = a->stream;
  float* fp = a->fp.as<float>();
  float* acc = a->facc.as<float>();
  auto lvl = [&](float* base, int i, int planes) { return base + (size_t)planes * a->loff[i]; };
  auto nl = [&](int i) { return (size_t)a->lw[i] * a->lh[i]; };
  const int top = a->L - 1;
  const int pol = a->fo.fusing_policy;
  // levels to reconstruct: lv_i in fp (select: the accumulator itself)
  std::vector<const float*> lv(a->L);
  for (int i = 0; i < a->L; ++i) {
    if (pol == SSG_FOCUS_SELECT_MAX_ENERGY || (pol == SSG_FOCUS_WEIGHTED_AVERAGE && i == top)) {
      lv[i] = lvl(acc, i, cn);
    } else {
      const int wc = pol == SSG_FOCUS_LPG_WEIGHTS ? 1 : a->wc;
      HIPCHK(launch_div(lvl(acc, i, cn), lvl(a->fw.as<float>(), i, wc), wc, lvl(fp, i, cn), nl(i), cn, s));
      lv[i] = lvl(fp, i, cn);
    }
  }
  // cur = top; cur = pyrUp(cur) + lv_i, into fp level i
  if (lv[top] != lvl(fp, top, cn))
    HIPCHK(hipMemcpyAsync(lvl(fp, top, cn), lv[top], sizeof(float) * cn * nl(top), hipMemcpyDeviceToDevice, s));
  for (int i = top - 1; i >= 0; --i) {
    HIPCHK(launch_pyrup(lvl(fp, i + 1, cn), (long long)nl(i + 1), a->lw[i + 1], a->lh[i + 1], a->fup.as<float>(),
                        (long long)nl(i), a->lw[i], a->lh[i], cn, s));
    HIPCHK(launch_ew(1, a->fup.as<float>(), lv[i], lvl(fp, i, cn), cn * nl(i), s));
  }
  return SSG_OK;
}
}  // namespace

extern "C" int ssg_ecc_normalize(const ssg_image* src, const ssg_image* mask, ssg_image* dst, int32_t lvl) {
  CHK(check_image(src, "src"));
  CHK(check_image(dst, "dst"));
  if (src->type != SSG_32FC1 || dst->type != SSG_32FC1 || dst->width != src->width || dst->height != src->height)
    return fail(SSG_E_INVALID_ARG, "src and dst must be SSG_32FC1 of one size");
  const int w = src->width, h = src->height;
  if (mask) {
    CHK(check_image(mask, "mask"));
    if (mask->type != SSG_8UC1 || mask->width != w || mask->height != h)
      return fail(SSG_E_INVALID_ARG, "mask must be SSG_8UC1 of the image size");
  }
  hipStream_t s = nullptr;
  HIPCHK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
  struct StreamGuard {
    hipStream_t s;
    ~StreamGuard() {
      (void)hipStreamSynchronize(s);
      (void)hipStreamDestroy(s);
    }
  } guard{s};
  DevBuf ts, tm, out;
  const void* sp;
  size_t spb;
  CHK(device_view(src, ts, s, &sp, &spb));
  std::vector<const uint8_t*> nm;
  size_t mpb = 0;
  if (mask) {
    const void* mp;
    CHK(device_view(mask, tm, s, &mp, &mpb));
    nm.push_back((const uint8_t*)mp);
  }
  float* op = (float*)dst->data;
  size_t opb = (size_t)dst->step;
  if (!dst->on_device || (opb & 3)) {
    opb = (size_t)w * 4;
    CHK(out.ensure(opb * h));
    op = out.as<float>();
  }
  CHK(ecc_normalize_frames({(const float*)sp}, (int)(spb / 4), nm, (int)mpb, w, h, lvl, op, (int)(opb / 4), 0, s));
  if (op != dst->data)
    HIPCHK(hipMemcpy2DAsync(dst->data, (size_t)dst->step, op, opb, (size_t)w * 4, h,
                            dst->on_device ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  return SSG_OK;
}

namespace {
// local_variance_map's measure of one image, enqueued on s (core/proc/focus.cc: the gray image,
// pdownscale :30-52, the morphological gradient, the block sums): a and b hold w * h floats,
// the block sums go to part[0 .. 2 * nb).  The last level comes back in *cur (cw x ch) and the
// gradient image in *grad.
int lvm_enqueue(const void* sp, size_t spb, int type, int w, int h, int kradius, int dscale, float* a, float* b,
                double* part, hipStream_t s, float** cur_out, float** grad, int* cw_out, int* ch_out, int* nb_out) {
  const bool integral = type == SSG_8UC1 || type == SSG_16UC1;
  // gray, dense
  if (type == SSG_32FC3) {
    HIPCHK(launch_bgr2gray((const float*)sp, (int)(spb / 4), w, h, a, s));
  } else if (integral) {
    HIPCHK(launch_int_to_f32(sp, (int)spb, w, h, type == SSG_16UC1, a, s));
  } else {
    HIPCHK(hipMemcpy2DAsync(a, (size_t)w * 4, sp, spb, (size_t)w * 4, h, hipMemcpyDeviceToDevice, s));
  }
  // pdownscale (:30-52): pyrDown while the image keeps a side >= 4
  int cw = w, ch = h;
  float* cur = a;
  float* oth = b;
  if (dscale > 0 && std::min(cw, ch) >= 4) {
    for (int l = 0; l < dscale; ++l) {
      const int nw = (cw + 1) / 2, nh = (ch + 1) / 2;
      HIPCHK(launch_pyrdown(cur, 0, cw, cw, ch, oth, 0, nw, nw, nh, 1, SSG_BORDER_REFLECT101, s));
      if (integral) HIPCHK(launch_round_half_up(oth, (size_t)nw * nh, s));
      std::swap(cur, oth);
      cw = nw;
      ch = nh;
      if (std::min(cw, ch) < 4) break;
    }
  }
  const size_t nl = (size_t)cw * ch;
  HIPCHK(launch_morph_grad(cur, cw, ch, std::max(1, kradius), oth, s));
  HIPCHK(launch_lvm_sums(oth, nl, part, s));
  *cur_out = cur;
  *grad = oth;
  *cw_out = cw;
  *ch_out = ch;
  *nb_out = lvm_sum_blocks(nl);
  return SSG_OK;
}

// The measure from the block sums, folded in block order; *W receives the weight sum
double lvm_fold(const double* hp, int nb, int type, double* W_out) {
  double W = 0, S4 = 0;
  for (int i = 0; i < nb; ++i) {
    W += hp[2 * i];
    S4 += hp[2 * i + 1];
  }
  // depthScale = 20 * maxval(CV_32F) / maxval(depth) (getMaxValForPixelDepth: 1, 255, 65535)
  const double ds = 20.0 / (type == SSG_8UC1 ? 255.0 : (type == SSG_16UC1 ? 65535.0 : 1.0));
  if (W_out) *W_out = W;
  return W > 0 ? ds * ds * ds / W * S4 : 0.0;
}
}  // namespace

extern "C" int ssg_local_variance_map(const ssg_image* image, int32_t kradius, int32_t dscale, ssg_image* map,
                                      double* q) {
  CHK(check_image(image, "image"));
  if (image->type != SSG_32FC1 && image->type != SSG_32FC3 && image->type != SSG_8UC1 && image->type != SSG_16UC1)
    return fail(SSG_E_INVALID_ARG, "image must be SSG_32FC1/3, SSG_8UC1 or SSG_16UC1");
  if (map) {
    CHK(check_image(map, "map"));
    if (map->type != SSG_32FC1 || map->width != image->width || map->height != image->height)
      return fail(SSG_E_INVALID_ARG, "map must be SSG_32FC1 of the image size");
  }
  if (dscale < 0 || dscale > 16) return fail(SSG_E_INVALID_ARG, "dscale %d", dscale);
  const int w = image->width, h = image->height;
  hipStream_t s = nullptr;
  HIPCHK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
  struct StreamGuard {
    hipStream_t s;
    ~StreamGuard() {
      (void)hipStreamSynchronize(s);
      (void)hipStreamDestroy(s);
    }
  } guard{s};
  DevBuf ts, a, b, part, out;
  const void* sp;
  size_t spb;
  CHK(device_view(image, ts, s, &sp, &spb));
  const size_t n = (size_t)w * h;
  CHK(a.ensure(sizeof(float) * n));
  CHK(b.ensure(sizeof(float) * n));
  CHK(part.ensure(sizeof(double) * 2 * lvm_sum_blocks(n)));
  float *cur, *G;
  int cw, ch, nb;
  CHK(lvm_enqueue(sp, spb, image->type, w, h, kradius, dscale, a.as<float>(), b.as<float>(), part.as<double>(), s,
                  &cur, &G, &cw, &ch, &nb));
  const size_t nl = (size_t)cw * ch;
  std::vector<double> hp(2 * nb);
  HIPCHK(hipMemcpyAsync(hp.data(), part.p, sizeof(double) * 2 * nb, hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  double W = 0;
  const double Q = lvm_fold(hp.data(), nb, image->type, &W);
  const double ds = 20.0 / (image->type == SSG_8UC1 ? 255.0 : (image->type == SSG_16UC1 ? 65535.0 : 1.0));
  if (q) *q = Q;
  if (!map) return SSG_OK;
  if (!(W > 0)) {
    if (map->on_device) {
      HIPCHK(hipMemset2DAsync(map->data, (size_t)map->step, 0, (size_t)w * 4, h, s));
      HIPCHK(hipStreamSynchronize(s));
    } else {
      for (int y = 0; y < h; ++y) std::memset((char*)map->data + (size_t)y * map->step, 0, (size_t)w * 4);
    }
    return SSG_OK;
  }
  HIPCHK(launch_lvm_map(G, nl, (float)(ds * ds * ds), (float)(0.05 * Q), cur, s));
  float* dptr = (float*)map->data;
  size_t dpb = (size_t)map->step;
  if (!map->on_device || (cw == w && ch == h)) {
    CHK(out.ensure(sizeof(float) * n));
    dptr = out.as<float>();
    dpb = (size_t)w * 4;
  }
  if (cw != w || ch != h) {
    HIPCHK(launch_resize_linear(cur, cw, ch, dptr, (int)(dpb / 4), w, h, s));
  } else {
    HIPCHK(hipMemcpyAsync(dptr, cur, sizeof(float) * n, hipMemcpyDeviceToDevice, s));
  }
  if (dptr != map->data)
    HIPCHK(hipMemcpy2DAsync(map->data, (size_t)map->step, dptr, dpb, (size_t)w * 4, h,
                            map->on_device ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  return SSG_OK;
}

extern "C" int ssg_lpg(const ssg_image* src, ssg_image* dst, double k, double p, double* sharpness) {
  CHK(check_image(src, "src"));
  CHK(check_image(dst, "dst"));
  if (src->type != SSG_32FC1 && src->type != SSG_32FC3) return fail(SSG_E_INVALID_ARG, "src must be SSG_32FC1/3");
  if (dst->type != SSG_32FC1 || dst->width != src->width || dst->height != src->height)
    return fail(SSG_E_INVALID_ARG, "dst must be SSG_32FC1 of the source size");
  if (src->width < 5 || src->height < 5) return fail(SSG_E_INVALID_ARG, "lpg needs at least 5x5 pixels");
  const int w = src->width, h = src->height, cn = channels_of(src->type);
  hipStream_t s = nullptr;
  HIPCHK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
  struct StreamGuard {
    hipStream_t s;
    ~StreamGuard() {
      (void)hipStreamSynchronize(s);
      (void)hipStreamDestroy(s);
    }
  } guard{s};
  DevBuf ts, planes, g, m, sums;
  const void* sp;
  size_t spb;
  CHK(device_view(src, ts, s, &sp, &spb));
  const size_t n = (size_t)w * h;
  CHK(planes.ensure(sizeof(float) * cn * n));
  CHK(g.ensure(sizeof(float) * n));
  CHK(m.ensure(sizeof(float) * n));
  HIPCHK(launch_split((const float*)sp, (int)(spb / 4), w, h, cn, planes.as<float>(), s));
  HIPCHK(launch_gray(planes.as<float>(), n, cn, g.as<float>(), s));
  const float l2 = (100.f / 4.f) * (100.f / 4.f), g2 = (100.f / 36.f) * (100.f / 36.f);
  HIPCHK(launch_lpg(g.as<float>(), w, h, (float)(k / (k + 1)) * l2, (float)(1. / (k + 1)) * g2, 1e-9f, (float)p,
                    m.as<float>(), s));
  if (sharpness) {
    CHK(sums.ensure(sizeof(double) * (h + 1)));
    HIPCHK(launch_sum_f32(m.as<float>(), w, w, h, sums.as<double>(), s));
    double total = 0;
    HIPCHK(hipMemcpyAsync(&total, sums.as<double>() + h, sizeof(double), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    *sharpness = total / (double)n;
  }
  HIPCHK(hipMemcpy2DAsync(dst->data, (size_t)dst->step, m.p, (size_t)w * 4, (size_t)w * 4, h,
                          dst->on_device ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  return SSG_OK;
}

extern "C" int ssg_accumulation_add_frames(ssg_accumulation* a, int n, const ssg_image* frames) {
  if (!a) return fail(SSG_E_INVALID_ARG, "null handle");
  if (n <= 0) return SSG_OK;
  if (!frames) return fail(SSG_E_INVALID_ARG, "null frames");
  if (a->focus()) return focus_add_frames(a, n, frames);
  for (int f = 0; f < n; ++f) CHK(ssg_accumulation_add(a, &frames[f], nullptr));
  return SSG_OK;
}

// Several focus sequences at once (see ssg_focus_multi.hip).  The frames of every sequence go
// through in chunks of m frames; each step of a chunk is one launch over all sequences.
extern "C" int ssg_focus_stack_batch(int32_t nseq, ssg_accumulation* const* accs, const int32_t* nframes,
                                     const ssg_image* const* frames, const float* const* shifts) {
  if (nseq <= 0) return SSG_OK;
  if (!accs || !nframes || !frames) return fail(SSG_E_INVALID_ARG, "null argument");
  if (nseq > FM_MAXJOBS / 2) return fail(SSG_E_UNSUPPORTED, "at most %d sequences per call", FM_MAXJOBS / 2);
  int maxn = 0;
  for (int q = 0; q < nseq; ++q) {
    ssg_accumulation* a = accs[q];
    if (!a || !a->focus()) return fail(SSG_E_INVALID_ARG, "sequence %d: not a focus stacking accumulator", q);
    for (int r = 0; r < q; ++r)
      if (accs[r] == a) return fail(SSG_E_INVALID_ARG, "sequence %d repeats the accumulator of sequence %d", q, r);
    if (a->fo.fusing_policy != SSG_FOCUS_LPG_WEIGHTS)
      return fail(SSG_E_UNSUPPORTED, "sequence %d: the batched form takes the lpg-weight fusing policy", q);
    if (nframes[q] < 0) return fail(SSG_E_INVALID_ARG, "sequence %d: %d frames", q, nframes[q]);
    if (nframes[q] > 0 && !frames[q]) return fail(SSG_E_INVALID_ARG, "sequence %d: null frames", q);
    for (int k = 0; k < nframes[q]; ++k) {
      const ssg_image& im = frames[q][k];
      CHK(check_image(&im, "frame"));
      if (!im.on_device || im.type != (a->cn == 3 ? SSG_32FC3 : SSG_32FC1) || im.width != a->w || im.height != a->h ||
          im.step != frames[q][0].step || im.step % 4)
        return fail(SSG_E_INVALID_ARG,
                    "sequence %d frame %d: f32 device images with the accumulator's size and channels, one step", q, k);
    }
    maxn = std::max(maxn, (int)nframes[q]);
  }
  if (maxn == 0) return SSG_OK;
  for (int q = 0; q < nseq; ++q) {
    ssg_accumulation* a = accs[q];
    if (a->accumulated == 0) CHK(focus_prepare(a));
    if (a->L > FM_MAXL) return fail(SSG_E_UNSUPPORTED, "sequence %d: %d pyramid levels (max %d)", q, a->L, FM_MAXL);
    if (a->Wk.nx != a->Wk.ny || !(a->Wk.nx & 1) || a->Wk.nx > SEP_MAX_TAPS)
      return fail(SSG_E_UNSUPPORTED, "sequence %d: smoothing kernel of %d taps", q, a->Wk.nx);
    HIPCHK(hipStreamSynchronize(a->stream));
  }
  ssg_accumulation* a0 = accs[0];
  hipStream_t s = a0->stream;
  // chunk size: all sequences' chunk planes within a fixed scratch budget
  std::vector<size_t> per(nseq);
  size_t per_all = 0;
  for (int q = 0; q < nseq; ++q) {
    const ssg_accumulation* a = accs[q];
    const size_t n0 = (size_t)a->w * a->h;
    per[q] = sizeof(float) * ((size_t)a->cn * a->P + (size_t)a->P + 2 * n0);
    if (nframes[q] > 0) per_all += per[q];
  }
  // at most 24 GiB and a third of the free device memory: m frames per sequence and chunk (<= 32)
  size_t free_b = 0, total_b = 0;
  HIPCHK(hipMemGetInfo(&free_b, &total_b));
  const size_t budget = std::min<size_t>((size_t)24 << 30, free_b / 3);
  const int m = (int)std::max<size_t>(1, std::min<size_t>({(size_t)tune_value(SSG_TUNE_FOCUS_CHUNK), (size_t)maxn,
                                                             budget / std::max<size_t>(per_all, 1)}));
  std::vector<size_t> soff(nseq + 1, 0);
  for (int q = 0; q < nseq; ++q) soff[q + 1] = soff[q] + (nframes[q] > 0 ? ((per[q] * m + 255) / 256) * 256 : 0);
  CHK(a0->fms.ensure(soff[nseq]));
  // device tables for every chunk: FmSeq[nchunks][nseq], the frame pointers, the shifts
  const int nchunks = (maxn + m - 1) / m;
  int nptr = 0;
  for (int q = 0; q < nseq; ++q) nptr += nframes[q];
  const size_t tab_bytes = sizeof(FmSeq) * (size_t)nchunks * nseq;
  const size_t ptr_off = (tab_bytes + 255) / 256 * 256;
  const size_t sh_off = ptr_off + ((sizeof(void*) * (size_t)nptr + 255) / 256) * 256;
  const size_t total = sh_off + sizeof(float) * 2 * (size_t)nptr;
  CHK(a0->fmtab.ensure(total));
  std::vector<char> host(total, 0);
  FmSeq* tab = (FmSeq*)host.data();
  const float** hptr = (const float**)(host.data() + ptr_off);
  float* hsh = (float*)(host.data() + sh_off);
  char* dbase = a0->fmtab.as<char>();
  std::vector<int> pbase(nseq, 0);
  for (int q = 0, o = 0; q < nseq; ++q) {
    pbase[q] = o;
    for (int k = 0; k < nframes[q]; ++k) {
      hptr[o + k] = (const float*)frames[q][k].data;
      hsh[2 * (o + k)] = shifts && shifts[q] ? shifts[q][2 * k] : 0.f;
      hsh[2 * (o + k) + 1] = shifts && shifts[q] ? shifts[q][2 * k + 1] : 0.f;
    }
    o += nframes[q];
  }
  const double l2 = (100.f / 4.f) * (100.f / 4.f), g2 = (100.f / 36.f) * (100.f / 36.f);
  for (int c = 0; c < nchunks; ++c)
    for (int q = 0; q < nseq; ++q) {
      const ssg_accumulation* a = accs[q];
      FmSeq& t = tab[(size_t)c * nseq + q];
      t.w = a->w;
      t.h = a->h;
      t.cn = a->cn;
      t.L = a->L;
      t.m = std::max(0, std::min(m, nframes[q] - c * m));
      for (int i = 0; i < a->L; ++i) {
        t.lw[i] = a->lw[i];
        t.lh[i] = a->lh[i];
        t.loff[i] = a->loff[i];
      }
      const size_t n0 = (size_t)a->w * a->h;
      float* base = (float*)(a0->fms.as<char>() + soff[q]);
      t.G = base;
      t.W = base + (size_t)a->cn * a->P * m;
      t.gray = t.W + (size_t)a->P * m;
      t.M = t.gray + n0 * m;
      t.acc = a->facc.as<float>();
      t.wsum = a->fw.as<float>();
      t.frames = (const float* const*)(dbase + ptr_off) + pbase[q] + (size_t)c * m;
      t.shift = shifts && shifts[q] ? (const float*)(dbase + sh_off) + 2 * ((size_t)pbase[q] + (size_t)c * m) : nullptr;
      t.frame_pitch = nframes[q] > 0 ? (int)(frames[q][0].step / 4) : 0;
      const double kk = a->fo.lpg_k;
      t.alpha = (float)(kk / (kk + 1)) * (float)l2;
      t.beta = (float)(1. / (kk + 1)) * (float)g2;
      t.eps = 1e-9f;
      t.p = (float)a->fo.lpg_p;
      t.K = a->Wk;
    }
  HIPCHK(hipMemcpyAsync(dbase, host.data(), total, hipMemcpyHostToDevice, s));
  HIPCHK(hipStreamSynchronize(s));
  int maxL = 0;
  for (int q = 0; q < nseq; ++q) maxL = std::max(maxL, accs[q]->L);
  // one smoothing kernel for every sequence: the blur runs its register-blocked form
  int blur_r = accs[0]->Wk.nx / 2;
  for (int q = 1; q < nseq; ++q)
    if (memcmp(&accs[q]->Wk, &accs[0]->Wk, sizeof(SepKernel)) != 0) blur_r = 0;
  if (accs[0]->Wk.tx != 1 || accs[0]->Wk.ty != 1) blur_r = 0;
  auto& prof = a0->prof;
  for (int c = 0; c < nchunks; ++c) {
    const FmSeq* hs = tab + (size_t)c * nseq;
    const FmSeq* ds = (const FmSeq*)dbase + (size_t)c * nseq;
    auto run = [&](int kind, auto&& add_jobs, int prof_kind) -> int {
      FmJobs J{};
      long long end = 0;
      add_jobs([&](int q, int k, int level) {
        const long long b = fm_job_blocks(hs[q], k, level);
        if (b <= 0) return;
        end += b;
        J.j[J.n++] = FmJob{q, level, k, (int)end};
      });
      if (J.n == 0) return SSG_OK;
      if (end > INT32_MAX) return fail(SSG_E_UNSUPPORTED, "step of %lld blocks", end);
      prof.start(s);
      HIPCHK(launch_fm_step(kind, J, ds, blur_r, s));
      prof.stop(s, prof_kind);
      return SSG_OK;
    };
    auto each = [&](int kind) {
      return [&, kind](auto&& add) {
        for (int q = 0; q < nseq; ++q)
          if (hs[q].m > 0) add(q, kind, 0);
      };
    };
    CHK(run(FM_PREP, each(FM_PREP), SSG_PROF_FOCUS_PREP));
    CHK(run(FM_LPG, each(FM_LPG), SSG_PROF_FOCUS_PREP));
    CHK(run(FM_BLUR, each(FM_BLUR), SSG_PROF_FOCUS_BLUR));
    for (int i = 0; i + 1 < maxL; ++i) {
      CHK(run(FM_DOWN_G, [&](auto&& add) {
        for (int q = 0; q < nseq; ++q)
          if (hs[q].m > 0 && i + 1 < hs[q].L) {
            add(q, FM_DOWN_G, i);
            add(q, FM_DOWN_W, i);
          }
      }, SSG_PROF_FOCUS_DOWN));
      CHK(run(FM_FUSE, [&](auto&& add) {
        for (int q = 0; q < nseq; ++q)
          if (hs[q].m > 0 && i + 1 < hs[q].L) add(q, FM_FUSE, i);
      }, SSG_PROF_FOCUS_FUSE));
    }
    CHK(run(FM_TOP, [&](auto&& add) {
      for (int q = 0; q < nseq; ++q)
        if (hs[q].m > 0) add(q, FM_TOP, hs[q].L - 1);
    }, SSG_PROF_FOCUS_FUSE));
  }
  HIPCHK(hipStreamSynchronize(s));
  if (prof.on) {
    // algorithmic bytes per frame (cn channels, N0 level-0 pixels, P pyramid pixels): prep reads
    // the frame and writes the planes and the gray image (8 cn + 4) N0; lpg 8 N0; the smoothing
    // 8 N0; each pyrDown reads its level and writes the next, (cn + 1)(n_i + n_i+1); the fuse reads
    // each level's planes and weights once, (cn + 1) P, plus the next level through L2
    for (int q = 0; q < nseq; ++q) {
      const ssg_accumulation* a = accs[q];
      const double n0 = (double)a->w * a->h, nf = nframes[q];
      double down = 0;
      for (int i = 0; i + 1 < a->L; ++i) down += (double)(a->cn + 1) * 4.0 * ((double)a->lw[i] * a->lh[i] + (double)a->lw[i + 1] * a->lh[i + 1]);
      prof.add_bytes(SSG_PROF_FOCUS_PREP, nf * ((8.0 * a->cn + 4.0) * n0 + 8.0 * n0));
      prof.add_bytes(SSG_PROF_FOCUS_BLUR, nf * 8.0 * n0);
      prof.add_bytes(SSG_PROF_FOCUS_DOWN, nf * down);
      prof.add_bytes(SSG_PROF_FOCUS_FUSE, nf * (a->cn + 1.0) * 4.0 * (double)a->P);
    }
    prof.collect();
  }
  for (int q = 0; q < nseq; ++q) accs[q]->accumulated += nframes[q];
  return SSG_OK;
}

extern "C" int ssg_accumulation_add(ssg_accumulation* a, const ssg_image* src, const ssg_image* weights) {
  if (!a) return fail(SSG_E_INVALID_ARG, "null handle");
  CHK(check_image(src, "src"));
  if (src->width != a->w || src->height != a->h)
    return fail(SSG_E_INVALID_ARG, "current frame (%dx%d) and accumulator (%dx%d) sizes not match", src->width, src->height,
                a->w, a->h);
  if (a->bayer()) return bayer_add(a, src, weights);
  if (a->focus()) {
    if (weights) return fail(SSG_E_UNSUPPORTED, "focus stacking takes frames without masks on this path");
    return focus_add(a, src);
  }
  if (channels_of(src->type) != a->cn || (src->type != SSG_32FC1 && src->type != SSG_32FC3))
    return fail(SSG_E_INVALID_ARG, "frame type %d does not match the accumulator (%d channels, f32)", src->type, a->cn);
  hipStream_t s = a->stream;
  DevBuf ts, tw;
  const void* sp;
  size_t spb;
  CHK(device_view(src, ts, s, &sp, &spb));
  const uint8_t* mp = nullptr;
  const float* wp = nullptr;
  size_t wpb = 0;
  if (weights) {
    CHK(check_image(weights, "weights"));
    if (weights->width != a->w || weights->height != a->h) return fail(SSG_E_INVALID_ARG, "weights size");
    const void* p;
    CHK(device_view(weights, tw, s, &p, &wpb));
    if (weights->type == SSG_8UC1) mp = (const uint8_t*)p;
    else if (weights->type == SSG_32FC1) wp = (const float*)p;
    else return fail(SSG_E_INVALID_ARG, "weights must be SSG_8UC1 or SSG_32FC1");
  }
  HIPCHK(launch_acc_add((const float*)sp, (int)(spb / 4), a->w, a->h, a->cn, mp, (int)wpb, wp, (int)(wpb / 4),
                        a->acc.as<float>(), nullptr, a->cnt.as<unsigned>(), a->plane, s));
  HIPCHK(hipStreamSynchronize(s));
  ++a->accumulated;
  return SSG_OK;
}

namespace {
int bayer_batch(ssg_accumulation* a, int n, const ssg_image* frames, const ssg_registration_status* statuses,
                int motion_type, const float* frame_weights, int interpolation) {
  if (!a->pattern) return fail(SSG_E_NOT_READY, "no bayer pattern set");
  if (frame_weights) return fail(SSG_E_INVALID_ARG, "bayer batches take the base_remap mask only (frame_weights NULL)");
  if (interpolation != SSG_INTER_LINEAR) return fail(SSG_E_UNSUPPORTED, "bayer batch masks are remapped LINEAR");
  const int type = frames[0].type;
  if (type != SSG_32FC1 && type != SSG_16UC1)
    return fail(SSG_E_INVALID_ARG, "bayer frames are SSG_32FC1 or SSG_16UC1 raw images");
  const int es = type == SSG_16UC1 ? 2 : 4;
  for (int f = 0; f < n; ++f) {
    CHK(check_image(&frames[f], "frame"));
    if (!frames[f].on_device) return fail(SSG_E_INVALID_ARG, "frame %d: batched frames must be device images", f);
    if (frames[f].width != a->w || frames[f].height != a->h || frames[f].type != type ||
        frames[f].step != frames[0].step)
      return fail(SSG_E_INVALID_ARG, "frame %d: bayer batch frames share the accumulator size, type and step", f);
  }
  if (frames[0].step % es) return fail(SSG_E_INVALID_ARG, "frame step must be a multiple of the element size");
  hipStream_t s = a->stream;
  std::vector<MapCoef> maps(n);
  std::vector<int> valid(n);
  std::vector<const void*> ptrs(n);
  int nvalid = 0;
  for (int f = 0; f < n; ++f) {
    maps[f] = map_coef(motion_type, statuses[f].params, nullptr);
    valid[f] = statuses[f].ok != 0;
    nvalid += valid[f];
    ptrs[f] = frames[f].data;
  }
  CHK(a->maps.ensure(sizeof(MapCoef) * n));
  CHK(a->valid.ensure(sizeof(int) * n));
  CHK(a->frame_ptrs.ensure(sizeof(void*) * n));
  HIPCHK(hipMemcpyAsync(a->maps.p, maps.data(), sizeof(MapCoef) * n, hipMemcpyHostToDevice, s));
  HIPCHK(hipMemcpyAsync(a->valid.p, valid.data(), sizeof(int) * n, hipMemcpyHostToDevice, s));
  HIPCHK(hipMemcpyAsync(a->frame_ptrs.p, ptrs.data(), sizeof(void*) * n, hipMemcpyHostToDevice, s));
  BayerBatchArgs g{};
  g.frames = a->frame_ptrs.as<const void*>();
  g.src_pitch = (int)(frames[0].step / es);
  g.scale = 1.f / (float)(1 << a->bpp);
  g.maps = a->maps.as<MapCoef>();
  g.valid = a->valid.as<int>();
  g.nframes = n;
  g.st = a->bayer_state();
  a->prof.start(s);
  HIPCHK(launch_bayer_batch(motion_type, type == SSG_16UC1, g, s));
  a->prof.stop(s, SSG_PROF_ACCUMULATE);
  HIPCHK(hipStreamSynchronize(s));
  // algorithmic bytes: every valid raw frame read once; accumulator planes read and written once
  a->prof.add_bytes(SSG_PROF_ACCUMULATE, (double)nvalid * es * a->plane + 2.0 * (6 * 4 + 3 * 4) * a->plane);
  a->prof.collect();
  a->accumulated += nvalid;
  return SSG_OK;
}
}  // namespace

namespace {
// Shared validation of a fused batch: frames (device, one geometry) and optional weight maps.
int check_batch(ssg_accumulation* a, int n, const ssg_image* frames, const ssg_image* weight_maps) {
  const int sw = frames[0].width, sh = frames[0].height;
  const int64_t step = frames[0].step;
  for (int f = 0; f < n; ++f) {
    CHK(check_image(&frames[f], "frame"));
    if (!frames[f].on_device) return fail(SSG_E_INVALID_ARG, "frame %d: batched frames must be device images", f);
    if (frames[f].width != sw || frames[f].height != sh || frames[f].step != step || frames[f].type != frames[0].type)
      return fail(SSG_E_INVALID_ARG, "frame %d: all frames of a batch share size, step and type", f);
    if (channels_of(frames[f].type) != a->cn) return fail(SSG_E_INVALID_ARG, "frame %d: channel count", f);
    if (step % 4) return fail(SSG_E_INVALID_ARG, "frame step must be a multiple of 4");
  }
  if (weight_maps) {
    for (int f = 0; f < n; ++f) {
      CHK(check_image(&weight_maps[f], "weight map"));
      if (!weight_maps[f].on_device || weight_maps[f].type != SSG_32FC1 || weight_maps[f].width != sw ||
          weight_maps[f].height != sh || weight_maps[f].step != weight_maps[0].step || weight_maps[0].step % 4)
        return fail(SSG_E_INVALID_ARG, "weight map %d: SSG_32FC1 device images of the frame size and one step", f);
    }
  }
  return SSG_OK;
}

// The fused remap -> mask -> weight -> accumulate of n frames.  motion is a parametric
// motion type (maps[f] per frame) or MOTION_MAP (rmaps[f]: dense fields at the accumulator size).
int accumulate_fused(ssg_accumulation* a, int n, const ssg_image* frames, int motion, const MapCoef* maps,
                     const ssg_image* rmaps, const int* valid, const float* w, const ssg_image* weight_maps,
                     int32_t interpolation, int32_t border_mode, double border_value) {
  hipStream_t s = a->stream;
  const int sw = frames[0].width, sh = frames[0].height;
  const int64_t step = frames[0].step;
  std::vector<const float*> ptrs(n);
  int nvalid = 0;
  for (int f = 0; f < n; ++f) {
    ptrs[f] = (const float*)frames[f].data;
    nvalid += valid[f];
  }
  const int nptr = 1 + (weight_maps ? 1 : 0) + (rmaps ? 1 : 0);
  CHK(a->weights.ensure(sizeof(float) * n));
  CHK(a->valid.ensure(sizeof(int) * n));
  CHK(a->frame_ptrs.ensure(sizeof(void*) * n * nptr));
  const float** dptr = a->frame_ptrs.as<const float*>();
  int slot = 1;
  const float* const* wmaps = nullptr;
  const float2* const* dmaps = nullptr;
  if (weight_maps) {
    std::vector<const float*> wp(n);
    for (int f = 0; f < n; ++f) wp[f] = (const float*)weight_maps[f].data;
    HIPCHK(hipMemcpyAsync(dptr + slot * n, wp.data(), sizeof(void*) * n, hipMemcpyHostToDevice, s));
    wmaps = dptr + slot * n;
    ++slot;
  }
  if (rmaps) {
    std::vector<const void*> mp(n);
    for (int f = 0; f < n; ++f) mp[f] = rmaps[f].data;
    HIPCHK(hipMemcpyAsync(dptr + slot * n, mp.data(), sizeof(void*) * n, hipMemcpyHostToDevice, s));
    dmaps = (const float2* const*)(dptr + slot * n);
  } else {
    CHK(a->maps.ensure(sizeof(MapCoef) * n));
    HIPCHK(hipMemcpyAsync(a->maps.p, maps, sizeof(MapCoef) * n, hipMemcpyHostToDevice, s));
  }
  HIPCHK(hipMemcpyAsync(a->weights.p, w, sizeof(float) * n, hipMemcpyHostToDevice, s));
  HIPCHK(hipMemcpyAsync(a->valid.p, valid, sizeof(int) * n, hipMemcpyHostToDevice, s));
  HIPCHK(hipMemcpyAsync(a->frame_ptrs.p, ptrs.data(), sizeof(void*) * n, hipMemcpyHostToDevice, s));
  AccumArgs g{};
  g.cn = a->cn;
  g.interp = interpolation == SSG_INTER_CUBIC ? INTER_CUBIC : INTER_LINEAR;
  g.wmaps = wmaps;
  // chunks: about SSG_TUNE_ACC_BLOCKS blocks over the tiles, each chunk a contiguous frame range
  const int nch = accumulate_chunks(n, a->w, a->h);
  const int per = (n + nch - 1) / nch;
  CHK(a->part.ensure(sizeof(float) * a->plane * (2 * a->cn + 2) * nch));
  CHK(a->part_cnt.ensure(sizeof(unsigned) * a->plane * nch));
  g.frames = a->frame_ptrs.as<const float*>();
  g.src_pitch = (int)(step / 4);
  g.sw = sw;
  g.sh = sh;
  g.maps = rmaps ? nullptr : a->maps.as<MapCoef>();
  g.rmaps = dmaps;
  g.rmap_pitch = rmaps ? (int)(rmaps[0].step / 8) : 0;
  g.weights = a->weights.as<float>();
  g.wmap_pitch = weight_maps ? (int)(weight_maps[0].step / 4) : 0;
  g.valid = a->valid.as<int>();
  g.nframes = n;
  g.per_chunk = per;
  g.dw = a->w;
  g.dh = a->h;
  g.border = border_mode;
  g.cval = (float)border_value;
  g.part = a->part.as<float>();
  g.part_cnt = a->part_cnt.as<unsigned>();
  g.plane = a->plane;
  g.nchunks = nch;
  g.rows_per_thread = tune_value(SSG_TUNE_ACC_ROWS);
  a->prof.start(s);
  HIPCHK(launch_accumulate(motion, g, s));
  a->prof.stop(s, SSG_PROF_ACCUMULATE);
  HIPCHK(launch_accum_merge(a->part.as<float>(), a->part_cnt.as<unsigned>(), nch, a->plane, a->cn, a->acc.as<float>(),
                            nullptr, a->cnt.as<unsigned>(), s));
  HIPCHK(hipStreamSynchronize(s));
  // algorithmic bytes of the fused kernel: each frame read once (4 B per channel and destination
  // pixel; +4 B for a weight map, +8 B for a dense field); each chunk writes its partial planes
  const double per_px = 4.0 * a->cn + (weight_maps ? 4.0 : 0.0) + (rmaps ? 8.0 : 0.0);
  a->prof.add_bytes(SSG_PROF_ACCUMULATE, (double)nvalid * per_px * a->plane +
                                             (double)nch * ((2.0 * a->cn + 2.0) * 4.0 + 4.0) * a->plane);
  a->prof.collect();
  a->accumulated += nvalid;
  return SSG_OK;
}
}  // namespace

extern "C" int ssg_accumulation_add_registered_batch(ssg_accumulation* a, int n, const ssg_image* frames,
                                                     const ssg_registration_status* statuses, int32_t motion_type,
                                                     const float* frame_weights, const ssg_image* weight_maps,
                                                     int32_t interpolation, int32_t border_mode,
                                                     double border_value) {
  if (!a) return fail(SSG_E_INVALID_ARG, "null handle");
  if (n <= 0) return SSG_OK;
  if (!frames || !statuses) return fail(SSG_E_INVALID_ARG, "null frames or statuses");
  if (interpolation != SSG_INTER_LINEAR && interpolation != SSG_INTER_CUBIC)
    return fail(SSG_E_UNSUPPORTED, "interpolation %d", interpolation);
  if (motion_type < SSG_MOTION_TRANSLATION || motion_type > SSG_MOTION_AFFINE)
    return fail(SSG_E_INVALID_ARG, "motion_type %d", motion_type);
  if (a->bayer()) {
    if (weight_maps) return fail(SSG_E_INVALID_ARG, "bayer batches take the base_remap mask only (weight_maps NULL)");
    return bayer_batch(a, n, frames, statuses, motion_type, frame_weights, interpolation);
  }
  if (frame_weights && weight_maps) return fail(SSG_E_INVALID_ARG, "pass frame_weights or weight_maps, not both");
  if (a->focus()) return fail(SSG_E_UNSUPPORTED, "focus stacking adds frames one by one (ssg_accumulation_add)");
  CHK(check_batch(a, n, frames, weight_maps));
  std::vector<MapCoef> maps(n);
  std::vector<float> w(n);
  std::vector<int> valid(n);
  for (int f = 0; f < n; ++f) {
    maps[f] = map_coef(motion_type, statuses[f].params, nullptr);
    w[f] = frame_weights ? frame_weights[f] : 1.f;
    // a weight <= 0 contributes nothing (c_frame_accumulation.cc:112: w > 0)
    valid[f] = statuses[f].ok && w[f] > 0.f;
  }
  return accumulate_fused(a, n, frames, motion_type, maps.data(), nullptr, valid.data(), w.data(), weight_maps,
                          interpolation, border_mode, border_value);
}

extern "C" int ssg_accumulation_add_remapped_batch(ssg_accumulation* a, int n, const ssg_image* frames,
                                                   const ssg_image* rmaps, const float* frame_weights,
                                                   const ssg_image* weight_maps, int32_t interpolation,
                                                   int32_t border_mode, double border_value) {
  if (!a) return fail(SSG_E_INVALID_ARG, "null handle");
  if (n <= 0) return SSG_OK;
  if (!frames || !rmaps) return fail(SSG_E_INVALID_ARG, "null frames or rmaps");
  if (interpolation != SSG_INTER_LINEAR && interpolation != SSG_INTER_CUBIC)
    return fail(SSG_E_UNSUPPORTED, "interpolation %d", interpolation);
  if (a->bayer() || a->focus())
    return fail(SSG_E_UNSUPPORTED, "dense-field batches go to the weighted average accumulator");
  if (frame_weights && weight_maps) return fail(SSG_E_INVALID_ARG, "pass frame_weights or weight_maps, not both");
  CHK(check_batch(a, n, frames, weight_maps));
  for (int f = 0; f < n; ++f) {
    CHK(check_image(&rmaps[f], "rmap"));
    if (!rmaps[f].on_device || rmaps[f].type != SSG_32FC2 || rmaps[f].width != a->w || rmaps[f].height != a->h ||
        rmaps[f].step != rmaps[0].step || rmaps[0].step % 8)
      return fail(SSG_E_INVALID_ARG, "rmap %d: SSG_32FC2 device maps of the accumulator size and one step", f);
  }
  std::vector<float> w(n);
  std::vector<int> valid(n);
  for (int f = 0; f < n; ++f) {
    w[f] = frame_weights ? frame_weights[f] : 1.f;
    valid[f] = w[f] > 0.f;
  }
  return accumulate_fused(a, n, frames, MOTION_MAP, nullptr, rmaps, valid.data(), w.data(), weight_maps,
                          interpolation, border_mode, border_value);
}

extern "C" int ssg_accumulation_compute(ssg_accumulation* a, ssg_image* avg, ssg_image* mask, double dscale) {
  if (!a) return fail(SSG_E_INVALID_ARG, "null handle");
  if (a->accumulated < 1) return fail(SSG_E_NOT_READY, "no frames accumulated");
  hipStream_t s = a->stream;
  DevBuf ta, tm;
  float* ap = nullptr;
  size_t apb = 0;
  uint8_t* mp = nullptr;
  size_t mpb = 0;
  const int ocn = a->bayer() ? 3 : a->cn;
  if (avg) {
    CHK(check_image(avg, "avg"));
    if (avg->width != a->w || avg->height != a->h || channels_of(avg->type) != ocn ||
        (avg->type != SSG_32FC1 && avg->type != SSG_32FC3))
      return fail(SSG_E_INVALID_ARG, "avg must be f32 with the accumulator's size and channels");
    if (avg->on_device) {
      ap = (float*)avg->data;
      apb = (size_t)avg->step;
    } else {
      apb = ((size_t)a->w * 4 * ocn + 255) / 256 * 256;
      CHK(ta.ensure(apb * a->h));
      ap = ta.as<float>();
    }
  }
  if (mask) {
    CHK(check_image(mask, "mask"));
    if (mask->type != SSG_8UC1 || mask->width != a->w || mask->height != a->h) return fail(SSG_E_INVALID_ARG, "mask");
    if (mask->on_device) {
      mp = (uint8_t*)mask->data;
      mpb = (size_t)mask->step;
    } else {
      mpb = ((size_t)a->w + 255) / 256 * 256;
      CHK(tm.ensure(mpb * a->h));
      mp = tm.as<uint8_t>();
    }
  }
  if (a->bayer()) {
    HIPCHK(launch_bayer_compute(a->bayer_state(), ap, (int)(apb / 4), mp, (int)mpb, s));
  } else if (a->focus()) {
    CHK(focus_compute(a));
    if (ap) HIPCHK(launch_merge(a->fp.as<float>(), a->w, a->h, a->cn, ap, (int)(apb / 4), s));
    if (mp) HIPCHK(hipMemset2DAsync(mp, mpb, 255, a->w, a->h, s));
  } else {
    HIPCHK(launch_acc_compute(a->acc.as<float>(), nullptr, a->w, a->h, a->cn, a->plane, ap, (int)(apb / 4), mp,
                              (int)mpb, dscale, s));
  }
  if (avg && !avg->on_device)
    HIPCHK(hipMemcpy2DAsync(avg->data, (size_t)avg->step, ap, apb, (size_t)a->w * 4 * ocn, a->h, hipMemcpyDeviceToHost, s));
  if (mask && !mask->on_device)
    HIPCHK(hipMemcpy2DAsync(mask->data, (size_t)mask->step, mp, mpb, a->w, a->h, hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  return SSG_OK;
}

extern "C" int ssg_accumulation_get_acc_counters(ssg_accumulation* a, ssg_image* accw) {
  if (!a) return fail(SSG_E_INVALID_ARG, "null handle");
  if (a->focus()) return SSG_OK;  // the reference releases accw (c_frame_accumulation.cc:683-687)
  CHK(check_image(accw, "accw"));
  if (a->bayer()) {
    if (accw->type != SSG_32FC3 || accw->width != a->w || accw->height != a->h)
      return fail(SSG_E_INVALID_ARG, "bayer counters are SSG_32FC3 of the accumulator size");
    DevBuf t;
    float* dp = (float*)accw->data;
    size_t pb = (size_t)accw->step;
    if (!accw->on_device) {
      pb = (size_t)a->w * 12;
      CHK(t.ensure(pb * a->h));
      dp = t.as<float>();
    }
    HIPCHK(launch_bayer_counters(a->bayer_state(), dp, (int)(pb / 4), a->stream));
    if (!accw->on_device)
      HIPCHK(hipMemcpy2DAsync(accw->data, (size_t)accw->step, dp, pb, pb, a->h, hipMemcpyDeviceToHost, a->stream));
    HIPCHK(hipStreamSynchronize(a->stream));
    return SSG_OK;
  }
  if (accw->type != SSG_32FC1 || accw->width != a->w || accw->height != a->h) return fail(SSG_E_INVALID_ARG, "accw");
  // weight plane hi (the lo plane holds the rounding residue)
  const float* hi = a->acc.as<float>() + (size_t)(2 * a->cn) * a->plane;
  HIPCHK(hipMemcpy2DAsync(accw->data, (size_t)accw->step, hi, (size_t)a->w * 4, (size_t)a->w * 4, a->h,
                          accw->on_device ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost, a->stream));
  HIPCHK(hipStreamSynchronize(a->stream));
  return SSG_OK;
}

extern "C" int ssg_accumulation_get_counts(ssg_accumulation* a, ssg_image* counts) {
  if (!a) return fail(SSG_E_INVALID_ARG, "null handle");
  CHK(check_image(counts, "counts"));
  if (a->bayer()) {
    // planes [c][h*w] -> interleaved SSG_32SC3 on the host side of the copy
    if (counts->type != SSG_32SC3 || counts->width != a->w || counts->height != a->h)
      return fail(SSG_E_INVALID_ARG, "bayer counts are SSG_32SC3 of the accumulator size");
    std::vector<unsigned> planes((size_t)3 * a->plane);
    HIPCHK(hipMemcpyAsync(planes.data(), a->cnt.p, sizeof(unsigned) * 3 * a->plane, hipMemcpyDeviceToHost, a->stream));
    HIPCHK(hipStreamSynchronize(a->stream));
    std::vector<unsigned> inter((size_t)3 * a->plane);
    for (long long o = 0; o < a->plane; ++o)
      for (int c = 0; c < 3; ++c) inter[(size_t)o * 3 + c] = planes[(size_t)c * a->plane + o];
    HIPCHK(hipMemcpy2DAsync(counts->data, (size_t)counts->step, inter.data(), (size_t)a->w * 12, (size_t)a->w * 12, a->h,
                            counts->on_device ? hipMemcpyHostToDevice : hipMemcpyHostToHost, a->stream));
    HIPCHK(hipStreamSynchronize(a->stream));
    return SSG_OK;
  }
  if (counts->type != SSG_32SC1 || counts->width != a->w || counts->height != a->h) return fail(SSG_E_INVALID_ARG, "counts");
  HIPCHK(hipMemcpy2DAsync(counts->data, (size_t)counts->step, a->cnt.p, (size_t)a->w * 4, (size_t)a->w * 4, a->h,
                          counts->on_device ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost, a->stream));
  HIPCHK(hipStreamSynchronize(a->stream));
  return SSG_OK;
}

extern "C" int ssg_accumulation_accumulated_frames(ssg_accumulation* a, int32_t* n) {
  if (!a || !n) return fail(SSG_E_INVALID_ARG, "null argument");
  *n = a->accumulated;
  return SSG_OK;
}

extern "C" int ssg_accumulation_reinitialize(ssg_accumulation* a, const ssg_image* src, const ssg_image* accw) {
  if (!a) return fail(SSG_E_INVALID_ARG, "null handle");
  // c_bayer_average and c_laplacian_pyramid_focus_stacking return false (c_frame_accumulation.cc:486-490,1170-1173)
  if (a->bayer() || a->focus()) return fail(SSG_E_UNSUPPORTED, "reinitialize: weighted average accumulators only");
  CHK(check_image(src, "src"));
  CHK(check_image(accw, "accw"));
  if (src->width != a->w || src->height != a->h || src->type != (a->cn == 3 ? SSG_32FC3 : SSG_32FC1))
    return fail(SSG_E_INVALID_ARG, "reinitialize: src must be f32 with the accumulator's size and channels");
  if (accw->width != a->w || accw->height != a->h || accw->type != SSG_32FC1)
    return fail(SSG_E_INVALID_ARG, "reinitialize: accw must be SSG_32FC1 of the accumulator size");
  hipStream_t s = a->stream;
  DevBuf ts, tw;
  const void *sp, *wp;
  size_t spb, wpb;
  CHK(device_view(src, ts, s, &sp, &spb));
  CHK(device_view(accw, tw, s, &wp, &wpb));
  if (spb % 4 || wpb % 4) return fail(SSG_E_INVALID_ARG, "reinitialize: row steps must be multiples of 4 bytes");
  HIPCHK(launch_acc_reinit((const float*)sp, (int)(spb / 4), (const float*)wp, (int)(wpb / 4), a->w, a->h, a->cn,
                           a->plane, a->acc.as<float>(), a->cnt.as<unsigned>(), s));
  HIPCHK(hipStreamSynchronize(s));
  a->accumulated = 1;
  return SSG_OK;
}

namespace {
// Accumulator state blob: a fixed header, then the device planes in a fixed order.
struct AccStateHeader {
  char magic[8];
  int32_t version, method, width, height, channels, accumulated;
  int32_t colorid, levels, policy, wchannels, reserved;
  uint64_t payload;
};
constexpr char kAccMagic[8] = {'S', 'S', 'G', 'A', 'C', 'C', 'S', 'T'};
constexpr int kAccStateVersion = 1;

struct Seg {
  void* p;
  size_t bytes;
};

// the device buffers that hold the state: weighted average (2 cn + 2) f32 planes + u32 counts;
// Bayer 6 f32 planes + 3 u32 count planes; focus stacking the fused pyramid and, for the
// weighted policies, its weight pyramid (the layout fixed by the options at the first add)
std::vector<Seg> acc_segments(ssg_accumulation* a) {
  std::vector<Seg> v;
  if (a->focus()) {
    if (a->L == 0) return v;
    v.push_back({a->facc.p, sizeof(float) * a->cn * a->P});
    if (a->fo.fusing_policy == SSG_FOCUS_LPG_WEIGHTS) v.push_back({a->fw.p, sizeof(float) * a->P});
    else if (a->fo.fusing_policy == SSG_FOCUS_WEIGHTED_AVERAGE) v.push_back({a->fw.p, sizeof(float) * a->wc * a->P});
    return v;
  }
  v.push_back({a->acc.p, a->acc_bytes()});
  v.push_back({a->cnt.p, a->cnt_bytes()});
  return v;
}

AccStateHeader acc_header(ssg_accumulation* a) {
  AccStateHeader h{};
  memcpy(h.magic, kAccMagic, 8);
  h.version = kAccStateVersion;
  h.method = a->method;
  h.width = a->w;
  h.height = a->h;
  h.channels = a->cn;
  h.accumulated = a->accumulated;
  h.colorid = a->bayer() ? a->colorid : 0;
  h.levels = a->focus() ? a->L : 0;
  h.policy = a->focus() ? a->fo.fusing_policy : 0;
  h.wchannels = a->focus() ? a->wc : 0;
  for (const Seg& g : acc_segments(a)) h.payload += g.bytes;
  return h;
}
}  // namespace

extern "C" int ssg_accumulation_state_size(ssg_accumulation* a, uint64_t* bytes) {
  if (!a || !bytes) return fail(SSG_E_INVALID_ARG, "null argument");
  *bytes = sizeof(AccStateHeader) + acc_header(a).payload;
  return SSG_OK;
}

extern "C" int ssg_accumulation_state_save(ssg_accumulation* a, void* buf, uint64_t bytes) {
  if (!a || !buf) return fail(SSG_E_INVALID_ARG, "null argument");
  const AccStateHeader h = acc_header(a);
  if (bytes < sizeof(h) + h.payload)
    return fail(SSG_E_INVALID_ARG, "state_save: %llu bytes given, %llu needed", (unsigned long long)bytes,
                (unsigned long long)(sizeof(h) + h.payload));
  memcpy(buf, &h, sizeof(h));
  char* dst = (char*)buf + sizeof(h);
  for (const Seg& g : acc_segments(a)) {
    HIPCHK(hipMemcpyAsync(dst, g.p, g.bytes, hipMemcpyDeviceToHost, a->stream));
    dst += g.bytes;
  }
  HIPCHK(hipStreamSynchronize(a->stream));
  return SSG_OK;
}

extern "C" int ssg_accumulation_state_load(ssg_accumulation* a, const void* buf, uint64_t bytes) {
  if (!a || !buf) return fail(SSG_E_INVALID_ARG, "null argument");
  AccStateHeader h;
  if (bytes < sizeof(h)) return fail(SSG_E_INVALID_ARG, "state_load: %llu bytes is no accumulator state",
                                     (unsigned long long)bytes);
  memcpy(&h, buf, sizeof(h));
  if (memcmp(h.magic, kAccMagic, 8) || h.version != kAccStateVersion)
    return fail(SSG_E_INVALID_ARG, "state_load: not an accumulator state of version %d", kAccStateVersion);
  if (h.method != a->method || h.width != a->w || h.height != a->h || h.channels != a->cn)
    return fail(SSG_E_INVALID_ARG, "state_load: state of method %d %dx%dx%d, accumulator of method %d %dx%dx%d",
                h.method, h.width, h.height, h.channels, a->method, a->w, a->h, a->cn);
  if (a->bayer() && h.colorid != a->colorid)
    return fail(SSG_E_INVALID_ARG, "state_load: bayer pattern %d, accumulator %d", h.colorid, a->colorid);
  if (a->focus()) {
    if (h.policy != a->fo.fusing_policy) return fail(SSG_E_INVALID_ARG, "state_load: another fusing policy");
    if (h.accumulated > 0 && a->L == 0) CHK(focus_prepare(a));
    if (h.accumulated > 0 && (h.levels != a->L || h.wchannels != a->wc))
      return fail(SSG_E_INVALID_ARG, "state_load: %d pyramid levels, the options give %d", h.levels, a->L);
  }
  if (bytes < sizeof(h) + h.payload) return fail(SSG_E_INVALID_ARG, "state_load: truncated state");
  const std::vector<Seg> segs = h.accumulated > 0 || !a->focus() ? acc_segments(a) : std::vector<Seg>();
  uint64_t tot = 0;
  for (const Seg& g : segs) tot += g.bytes;
  if (tot != h.payload) return fail(SSG_E_INVALID_ARG, "state_load: payload of %llu bytes, %llu expected",
                                    (unsigned long long)h.payload, (unsigned long long)tot);
  const char* src = (const char*)buf + sizeof(h);
  for (const Seg& g : segs) {
    HIPCHK(hipMemcpyAsync(g.p, src, g.bytes, hipMemcpyHostToDevice, a->stream));
    src += g.bytes;
  }
  HIPCHK(hipStreamSynchronize(a->stream));
  a->accumulated = h.accumulated;
  return SSG_OK;
}

extern "C" int ssg_registration_set_profiling(ssg_registration* r, int enable) {
  if (!r) return fail(SSG_E_INVALID_ARG, "null handle");
  r->prof.on = enable != 0;
  return SSG_OK;
}

extern "C" int ssg_registration_get_profile(ssg_registration* r, ssg_profile* out, int reset) {
  if (!r || !out) return fail(SSG_E_INVALID_ARG, "null argument");
  *out = r->prof.acc;
  if (reset) r->prof.acc = ssg_profile{};
  return SSG_OK;
}

extern "C" int ssg_accumulation_set_profiling(ssg_accumulation* a, int enable) {
  if (!a) return fail(SSG_E_INVALID_ARG, "null handle");
  a->prof.on = enable != 0;
  return SSG_OK;
}

extern "C" int ssg_accumulation_get_profile(ssg_accumulation* a, ssg_profile* out, int reset) {
  if (!a || !out) return fail(SSG_E_INVALID_ARG, "null argument");
  *out = a->prof.acc;
  if (reset) a->prof.acc = ssg_profile{};
  return SSG_OK;
}

// ---- FFT phase correlation ------------------------------------------------------------------------
struct ssg_phasecorr {
  int w = 0, h = 0, binning = 1, bw = 0, bh = 0, M = 0, N = 0;
  hipStream_t stream = nullptr;
  DevBuf wx, wy, tw_m, tw_n, ref_spec, buf, tmp, frame_ptrs, out, mean, part, pv, pi;
  int chunk = 1;
  bool have_ref = false;
  Profiler prof;
  ~ssg_phasecorr() {
    if (stream) {
      (void)hipStreamSynchronize(stream);
      (void)hipStreamDestroy(stream);
    }
  }
};

namespace {
int pow2_at_least(int n) {
  int p = 1;
  while (p < n) p <<= 1;
  return p;
}

// frames -> spectra in pc->tmp ([M][N] per frame)
int pc_spectra(ssg_phasecorr* pc, int n, const void* const* dev_ptrs, bool u16, int pitch) {
  hipStream_t s = pc->stream;
  CHK(pc->mean.ensure(sizeof(float) * n));
  CHK(pc->part.ensure(sizeof(double) * PC_PARTS_MAX * n));
  HIPCHK(launch_pc_load(dev_ptrs, u16, pitch, pc->bw, pc->bh, pc->binning, pc->mean.as<float>(), pc->part.as<double>(),
                        pc->wx.as<float>(), pc->wy.as<float>(), pc->buf.as<float2>(), pc->M, pc->N, n, s));
  HIPCHK(launch_fft2(pc->buf.as<float2>(), pc->tmp.as<float2>(), pc->M, pc->N, n, pc->tw_m.as<float2>(),
                     pc->tw_n.as<float2>(), false, s));
  return SSG_OK;
}
}  // namespace

extern "C" int ssg_phasecorr_create(int32_t width, int32_t height, int32_t binning, ssg_phasecorr** out) {
  if (!out) return fail(SSG_E_INVALID_ARG, "out is null");
  *out = nullptr;
  if (binning != 1 && binning != 2) return fail(SSG_E_INVALID_ARG, "binning must be 1 or 2");
  if (width < 4 * binning || height < 4 * binning) return fail(SSG_E_INVALID_ARG, "frame %dx%d too small", width, height);
  auto pc = std::make_unique<ssg_phasecorr>();
  pc->w = width;
  pc->h = height;
  pc->binning = binning;
  pc->bw = width / binning;
  pc->bh = height / binning;
  pc->M = pow2_at_least(pc->bw);
  pc->N = pow2_at_least(pc->bh);
  if (pc->M > 4096 || pc->N > 4096)
    return fail(SSG_E_UNSUPPORTED, "padded FFT size %dx%d above 4096", pc->M, pc->N);
  HIPCHK(hipStreamCreateWithFlags(&pc->stream, hipStreamNonBlocking));
  std::vector<float> wx(pc->bw), wy(pc->bh);
  auto hann = [](int n, int i) { return n == 1 ? 1.0 : 0.5 - 0.5 * std::cos(2.0 * M_PI * i / (n - 1)); };
  for (int i = 0; i < pc->bw; ++i) wx[i] = (float)hann(pc->bw, i);
  for (int i = 0; i < pc->bh; ++i) wy[i] = (float)hann(pc->bh, i);
  auto twiddles = [](int n) {
    std::vector<float2> t(std::max(1, n / 2));
    for (int k = 0; k < n / 2; ++k) t[k] = make_float2((float)std::cos(-2.0 * M_PI * k / n), (float)std::sin(-2.0 * M_PI * k / n));
    return t;
  };
  const auto tm = twiddles(pc->M), tn = twiddles(pc->N);
  CHK(pc->wx.ensure(sizeof(float) * wx.size()));
  CHK(pc->wy.ensure(sizeof(float) * wy.size()));
  CHK(pc->tw_m.ensure(sizeof(float2) * tm.size()));
  CHK(pc->tw_n.ensure(sizeof(float2) * tn.size()));
  HIPCHK(hipMemcpyAsync(pc->wx.p, wx.data(), sizeof(float) * wx.size(), hipMemcpyHostToDevice, pc->stream));
  HIPCHK(hipMemcpyAsync(pc->wy.p, wy.data(), sizeof(float) * wy.size(), hipMemcpyHostToDevice, pc->stream));
  HIPCHK(hipMemcpyAsync(pc->tw_m.p, tm.data(), sizeof(float2) * tm.size(), hipMemcpyHostToDevice, pc->stream));
  HIPCHK(hipMemcpyAsync(pc->tw_n.p, tn.data(), sizeof(float2) * tn.size(), hipMemcpyHostToDevice, pc->stream));
  // frames per launch: two complex planes per frame, at most about 2 GiB in flight
  const size_t per = (size_t)pc->M * pc->N * sizeof(float2);
  pc->chunk = (int)std::max<size_t>(1, std::min<size_t>(64, ((size_t)2 << 30) / (2 * per)));
  CHK(pc->ref_spec.ensure(per));
  HIPCHK(hipStreamSynchronize(pc->stream));
  *out = pc.release();
  return SSG_OK;
}

extern "C" void ssg_phasecorr_destroy(ssg_phasecorr* pc) { delete pc; }

extern "C" int ssg_phasecorr_set_reference(ssg_phasecorr* pc, const ssg_image* ref) {
  if (!pc) return fail(SSG_E_INVALID_ARG, "null handle");
  CHK(check_image(ref, "reference"));
  if (ref->width != pc->w || ref->height != pc->h) return fail(SSG_E_INVALID_ARG, "reference size");
  if (ref->type != SSG_32FC1 && ref->type != SSG_16UC1)
    return fail(SSG_E_INVALID_ARG, "phase correlation takes SSG_32FC1 or SSG_16UC1 frames");
  hipStream_t s = pc->stream;
  DevBuf t;
  const void* p;
  size_t pb;
  CHK(device_view(ref, t, s, &p, &pb));
  const size_t per = (size_t)pc->M * pc->N * sizeof(float2);
  CHK(pc->buf.ensure(per));
  CHK(pc->tmp.ensure(per));
  CHK(pc->frame_ptrs.ensure(sizeof(void*)));
  HIPCHK(hipMemcpyAsync(pc->frame_ptrs.p, &p, sizeof(void*), hipMemcpyHostToDevice, s));
  const bool u16 = ref->type == SSG_16UC1;
  CHK(pc_spectra(pc, 1, pc->frame_ptrs.as<const void*>(), u16, (int)(pb / (u16 ? 2 : 4))));
  HIPCHK(hipMemcpyAsync(pc->ref_spec.p, pc->tmp.p, per, hipMemcpyDeviceToDevice, s));
  HIPCHK(hipStreamSynchronize(s));
  pc->have_ref = true;
  return SSG_OK;
}

extern "C" int ssg_phasecorr_estimate(ssg_phasecorr* pc, int n, const ssg_image* frames, double* out) {
  if (!pc) return fail(SSG_E_INVALID_ARG, "null handle");
  if (n <= 0) return SSG_OK;
  if (!frames || !out) return fail(SSG_E_INVALID_ARG, "null frames or out");
  if (!pc->have_ref) return fail(SSG_E_NOT_READY, "no reference frame set");
  const int type = frames[0].type;
  if (type != SSG_32FC1 && type != SSG_16UC1)
    return fail(SSG_E_INVALID_ARG, "phase correlation takes SSG_32FC1 or SSG_16UC1 frames");
  const int es = type == SSG_16UC1 ? 2 : 4;
  for (int f = 0; f < n; ++f) {
    CHK(check_image(&frames[f], "frame"));
    if (!frames[f].on_device) return fail(SSG_E_INVALID_ARG, "frame %d: batched frames must be device images", f);
    if (frames[f].width != pc->w || frames[f].height != pc->h || frames[f].type != type ||
        frames[f].step != frames[0].step)
      return fail(SSG_E_INVALID_ARG, "frame %d: frames share the configured size, type and step", f);
  }
  if (frames[0].step % es) return fail(SSG_E_INVALID_ARG, "frame step must be a multiple of the element size");
  hipStream_t s = pc->stream;
  const int chunk = std::min(pc->chunk, n);
  const long long per = (long long)pc->M * pc->N;
  CHK(pc->buf.ensure(sizeof(float2) * per * chunk));
  CHK(pc->tmp.ensure(sizeof(float2) * per * chunk));
  CHK(pc->frame_ptrs.ensure(sizeof(void*) * n));
  CHK(pc->out.ensure(sizeof(double) * 3 * n));
  std::vector<const void*> ptrs(n);
  for (int f = 0; f < n; ++f) ptrs[f] = frames[f].data;
  HIPCHK(hipMemcpyAsync(pc->frame_ptrs.p, ptrs.data(), sizeof(void*) * n, hipMemcpyHostToDevice, s));
  const int pitch = (int)(frames[0].step / es);
  pc->prof.start(s);
  for (int f0 = 0; f0 < n; f0 += chunk) {
    const int m = std::min(chunk, n - f0);
    CHK(pc_spectra(pc, m, pc->frame_ptrs.as<const void*>() + f0, type == SSG_16UC1, pitch));
    HIPCHK(launch_pc_cross(pc->tmp.as<float2>(), pc->ref_spec.as<float2>(), pc->M, pc->N, m, s));
    HIPCHK(launch_fft2(pc->tmp.as<float2>(), pc->buf.as<float2>(), pc->M, pc->N, m, pc->tw_m.as<float2>(),
                       pc->tw_n.as<float2>(), true, s));
    CHK(pc->pv.ensure(sizeof(float) * PC_PARTS_MAX * m));
    CHK(pc->pi.ensure(sizeof(int) * PC_PARTS_MAX * m));
    HIPCHK(launch_pc_peak(pc->buf.as<float2>(), pc->M, pc->N, pc->binning, m, pc->pv.as<float>(), pc->pi.as<int>(),
                          pc->out.as<double>() + 3 * f0, s));
  }
  pc->prof.stop(s, SSG_PROF_FFT);
  HIPCHK(hipMemcpyAsync(out, pc->out.p, sizeof(double) * 3 * n, hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  // algorithmic bytes: each frame read once (es B per pixel), plus the passes over the
  // complex planes: load write, 2 x (row FFT r+w, transpose r+w, row FFT r+w), cross
  // r+w + reference read, peak read -- 17 plane sweeps of 8 B per padded element
  pc->prof.add_bytes(SSG_PROF_FFT, (double)n * ((double)es * pc->w * pc->h + 17.0 * 8.0 * per));
  pc->prof.collect();
  return SSG_OK;
}

extern "C" int ssg_phasecorr_set_profiling(ssg_phasecorr* pc, int enable) {
  if (!pc) return fail(SSG_E_INVALID_ARG, "null handle");
  pc->prof.on = enable != 0;
  return SSG_OK;
}

extern "C" int ssg_phasecorr_get_profile(ssg_phasecorr* pc, ssg_profile* out, int reset) {
  if (!pc || !out) return fail(SSG_E_INVALID_ARG, "null argument");
  *out = pc->prof.acc;
  if (reset) pc->prof.acc = ssg_profile{};
  return SSG_OK;
}

extern "C" int ssg_frame_sharpness(int n, const ssg_image* frames, float* weights) {
  if (n <= 0) return SSG_OK;
  if (!frames || !weights) return fail(SSG_E_INVALID_ARG, "null frames or weights");
  for (int f = 0; f < n; ++f) {
    CHK(check_image(&frames[f], "frame"));
    if (!frames[f].on_device || frames[f].type != SSG_32FC1 || frames[f].width != frames[0].width ||
        frames[f].height != frames[0].height || frames[f].step != frames[0].step || frames[f].step % 4)
      return fail(SSG_E_INVALID_ARG, "frame %d: SSG_32FC1 device images of one size and step", f);
  }
  const int w = frames[0].width, h = frames[0].height;
  const int nb = (h + 15) / 16;
  hipStream_t s = 0;
  HIPCHK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
  struct StreamGuard {
    hipStream_t s;
    ~StreamGuard() { (void)hipStreamSynchronize(s); (void)hipStreamDestroy(s); }
  } guard{s};
  DevBuf ptrs, part;
  CHK(ptrs.ensure(sizeof(void*) * n));
  CHK(part.ensure(sizeof(double) * (size_t)n * nb));
  std::vector<const float*> hp(n);
  for (int f = 0; f < n; ++f) hp[f] = (const float*)frames[f].data;
  HIPCHK(hipMemcpyAsync(ptrs.p, hp.data(), sizeof(void*) * n, hipMemcpyHostToDevice, s));
  HIPCHK(launch_sharpness(ptrs.as<const float*>(), (int)(frames[0].step / 4), w, h, n, part.as<double>(), s));
  std::vector<double> hpart((size_t)n * nb);
  HIPCHK(hipMemcpyAsync(hpart.data(), part.p, sizeof(double) * hpart.size(), hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  for (int f = 0; f < n; ++f) {
    double t = 0;
    for (int b = 0; b < nb; ++b) t += hpart[(size_t)f * nb + b];
    weights[f] = (float)(t / ((double)w * h));
  }
  return SSG_OK;
}

// ---- allocation-free variant of ssg_frame_sharpness --------------------------------------------
// The scratch memory lives in a file-scope device array (one copy per GPU, held until the
// process ends).  Host code never passes `scratch` itself to a kernel or a copy: its device
// address comes from hipGetSymbolAddress, once per device, and is kept in g_scratch_ptr.
// Layout: [0, 16 KiB) the frame pointer table (2048 entries), [16 KiB, 24 KiB) the weights
// (2048 floats), [24 KiB, end) the per-block partial sums (doubles).
static __device__ __align__(256) unsigned char scratch[SSG_SHARPNESS_NOALLOC_SCRATCH_BYTES];

namespace {
constexpr size_t kTabOff = 0, kOutOff = 16384, kPartOff = 24576;
static_assert(SSG_SHARPNESS_NOALLOC_MAX_FRAMES * sizeof(void*) <= kOutOff - kTabOff, "pointer table");
static_assert(SSG_SHARPNESS_NOALLOC_MAX_FRAMES * sizeof(float) <= kPartOff - kOutOff, "weights");

std::mutex g_scratch_mu;
constexpr int kMaxDevices = 64;
void* g_scratch_ptr[kMaxDevices] = {};

// the device address of the current device's scratch array (caller holds g_scratch_mu)
int scratch_address(unsigned char** out) {
  int dev = -1;
  HIPCHK(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) return fail(SSG_E_UNSUPPORTED, "device %d: at most %d devices", dev, kMaxDevices);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    HIPCHK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) return fail(SSG_E_HIP, "hipGetSymbolAddress(scratch) returned a null address");
    g_scratch_ptr[dev] = p;
  }
  *out = (unsigned char*)g_scratch_ptr[dev];
  return SSG_OK;
}
}  // namespace

extern "C" int ssg_frame_sharpness_noalloc(int n, const ssg_image* frames, float* weights) {
  if (n <= 0) return SSG_OK;
  if (!frames || !weights) return fail(SSG_E_INVALID_ARG, "null frames or weights");
  for (int f = 0; f < n; ++f) {
    CHK(check_image(&frames[f], "frame"));
    if (!frames[f].on_device || frames[f].type != SSG_32FC1 || frames[f].width != frames[0].width ||
        frames[f].height != frames[0].height || frames[f].step != frames[0].step || frames[f].step % 4)
      return fail(SSG_E_INVALID_ARG, "frame %d: SSG_32FC1 device images of one size and step", f);
  }
  const int w = frames[0].width, h = frames[0].height;
  const int nb = (h + 15) / 16;
  const size_t need = kPartOff + sizeof(double) * (size_t)n * nb;
  if (n > SSG_SHARPNESS_NOALLOC_MAX_FRAMES || need > SSG_SHARPNESS_NOALLOC_SCRATCH_BYTES)
    return fail(SSG_E_INVALID_ARG,
                "ssg_frame_sharpness_noalloc: %d frames of height %d need %zu bytes of scratch; the maximum is %d "
                "frames and %d bytes (24576 + 8 * n * ceil(height / 16))",
                n, h, need, SSG_SHARPNESS_NOALLOC_MAX_FRAMES, SSG_SHARPNESS_NOALLOC_SCRATCH_BYTES);
  const float* tab[SSG_SHARPNESS_NOALLOC_MAX_FRAMES];
  for (int f = 0; f < n; ++f) tab[f] = (const float*)frames[f].data;
  // one call at a time per process: every call on a GPU shares its scratch array
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  unsigned char* base = nullptr;
  CHK(scratch_address(&base));
  hipStream_t s = nullptr;  // the null stream: no stream is created
  const float* const* dtab = (const float* const*)(base + kTabOff);
  float* dout = (float*)(base + kOutOff);
  double* dpart = (double*)(base + kPartOff);
  // every slot that is read is written first (table, partials, weights): nothing to clear
  HIPCHK(hipMemcpyAsync((void*)dtab, tab, sizeof(void*) * n, hipMemcpyHostToDevice, s));
  HIPCHK(launch_sharpness(dtab, (int)(frames[0].step / 4), w, h, n, dpart, s));
  HIPCHK(launch_sharpness_finish(dpart, nb, n, (double)w * h, dout, s));
  HIPCHK(hipMemcpyAsync(weights, dout, sizeof(float) * n, hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  return SSG_OK;
}

// =====================================================================================================
// dense ECC flow: c_eccflow (core/proc/image_registration/ecc2.cc:2103-2865)
// =====================================================================================================
namespace {
// Geometry tables, packed into one device buffer per handle.
struct TabBuild {
  std::vector<char> bytes;
  template <class T>
  size_t put(const std::vector<T>& v) {
    const size_t off = (bytes.size() + 15) / 16 * 16;
    bytes.resize(off + sizeof(T) * std::max<size_t>(1, v.size()));
    if (!v.empty()) memcpy(bytes.data() + off, v.data(), sizeof(T) * v.size());
    return off;
  }
};

// cv::resize INTER_AREA geometry (resize.cpp: is_area_fast, computeResizeAreaTab) in double
struct AreaHost {
  AreaGeom g{};
  size_t o[6] = {0, 0, 0, 0, 0, 0};  // xoff, xsi, xal, yoff, ysi, yal
};

bool area_fast_axis(int ssize, int dsize, int* iscale) {
  const double scale = 1.0 / ((double)dsize / ssize);
  const int is = (int)std::lrint(scale);
  *iscale = is;
  return std::abs(scale - is) < 2.220446049250313e-16;
}

void area_axis(int ssize, int dsize, std::vector<int>& off, std::vector<int>& si, std::vector<float>& al) {
  const double scale = 1.0 / ((double)dsize / ssize);
  off.assign(dsize + 1, 0);
  for (int dx = 0; dx < dsize; ++dx) {
    off[dx] = (int)si.size();
    const double fsx1 = dx * scale, fsx2 = fsx1 + scale;
    const double cell = std::min(scale, ssize - fsx1);
    int sx1 = (int)std::ceil(fsx1), sx2 = (int)std::floor(fsx2);
    sx2 = std::min(sx2, ssize - 1);
    sx1 = std::min(sx1, sx2);
    if (sx1 - fsx1 > 1e-3) {
      si.push_back(sx1 - 1);
      al.push_back((float)((sx1 - fsx1) / cell));
    }
    for (int sx = sx1; sx < sx2; ++sx) {
      si.push_back(sx);
      al.push_back((float)(1.0 / cell));
    }
    if (fsx2 - sx2 > 1e-3) {
      si.push_back(sx2);
      al.push_back((float)(std::min(std::min(fsx2 - sx2, 1.), cell) / cell));
    }
  }
  off[dsize] = (int)si.size();
}

AreaHost area_host(int sw, int sh, int dw, int dh, TabBuild& tb) {
  AreaHost a;
  a.g.sw = sw;
  a.g.sh = sh;
  a.g.dw = dw;
  a.g.dh = dh;
  const bool fx = area_fast_axis(sw, dw, &a.g.isx), fy = area_fast_axis(sh, dh, &a.g.isy);
  a.g.fast = fx && fy;
  if (!a.g.fast) {
    std::vector<int> xo, xs, yo, ys;
    std::vector<float> xa, ya;
    area_axis(sw, dw, xo, xs, xa);
    area_axis(sh, dh, yo, ys, ya);
    a.o[0] = tb.put(xo);
    a.o[1] = tb.put(xs);
    a.o[2] = tb.put(xa);
    a.o[3] = tb.put(yo);
    a.o[4] = tb.put(ys);
    a.o[5] = tb.put(ya);
  }
  return a;
}

void area_bind(AreaHost& a, const char* base) {
  if (a.g.fast) return;
  a.g.xoff = (const int*)(base + a.o[0]);
  a.g.xsi = (const int*)(base + a.o[1]);
  a.g.xal = (const float*)(base + a.o[2]);
  a.g.yoff = (const int*)(base + a.o[3]);
  a.g.ysi = (const int*)(base + a.o[4]);
  a.g.yal = (const float*)(base + a.o[5]);
}

// cv::resize INTER_CUBIC axis: floor of fx = (float)((d+0.5) scale - 0.5) and
// interpolateCubic(fx - floor) (resize.cpp)
struct CubicHost {
  int sw, sh, dw, dh;
  size_t o[4];  // xs, xa, ys, ya
};

void cubic_axis(int ssize, int dsize, std::vector<int>& s, std::vector<float4>& c) {
  const double scale = 1.0 / ((double)dsize / ssize);
  for (int d = 0; d < dsize; ++d) {
    float fx = (float)((d + 0.5) * scale - 0.5);
    const int sx = (int)std::floor(fx);
    fx -= (float)sx;
    float k[4];
    cubic_coeffs(fx, k);
    s.push_back(sx);
    c.push_back(make_float4(k[0], k[1], k[2], k[3]));
  }
}

CubicHost cubic_host(int sw, int sh, int dw, int dh, TabBuild& tb) {
  CubicHost h{sw, sh, dw, dh, {0, 0, 0, 0}};
  std::vector<int> xs, ys;
  std::vector<float4> xa, ya;
  cubic_axis(sw, dw, xs, xa);
  cubic_axis(sh, dh, ys, ya);
  h.o[0] = tb.put(xs);
  h.o[1] = tb.put(xa);
  h.o[2] = tb.put(ys);
  h.o[3] = tb.put(ya);
  return h;
}

FlowCubicArgs cubic_args(const CubicHost& h, const char* base) {
  FlowCubicArgs a{};
  a.sw = h.sw;
  a.sh = h.sh;
  a.dw = h.dw;
  a.dh = h.dh;
  a.xs = (const int*)(base + h.o[0]);
  a.xa = (const float4*)(base + h.o[1]);
  a.ys = (const int*)(base + h.o[2]);
  a.ya = (const float4*)(base + h.o[3]);
  a.rx = a.ry = 1.0;
  return a;
}

// ecc_differentiate (ecc2.cc:142-169): d5 = [1 -8 0 8 -1]/12 along one axis, [1 2 1]/4 along the other
SepKernel diff_sep(bool along_x) {
  const float d5[5] = {1.f / 12.f, -2.f / 3.f, 0.f, 2.f / 3.f, -1.f / 12.f};
  const float s3[3] = {0.25f, 0.5f, 0.25f};
  SepKernel K{};
  float* kd = along_x ? K.kx : K.ky;
  float* ks = along_x ? K.ky : K.kx;
  for (int i = 0; i < 5; ++i) kd[i] = d5[i];
  for (int i = 0; i < 3; ++i) ks[i] = s3[i];
  if (along_x) {
    K.nx = 5, K.tx = 2, K.ny = 3, K.ty = 1;
  } else {
    K.nx = 3, K.tx = 1, K.ny = 5, K.ty = 2;
  }
  K.delta = 0.f;
  return K;
}
}  // namespace

struct ssg_eccflow {
  ssg_eccflow_options o{};
  hipStream_t stream = nullptr;
  int L = 0, w0 = 0, h0 = 0;
  std::vector<int> lw, lh, cw, ch, src;  // level and cell sizes, source level of each level
  std::vector<int> fuse_k;               // cell size of levels that take the fused iteration, else 0
  std::vector<size_t> loff, coff;        // element offsets of the levels / cell grids when packed
  size_t lsum = 0, csum = 0;
  bool have_ref = false, ref_mask = false;
  DevBuf ref, rmask, Ix, Iy, D, tabs, ref_tmp, ref_ptr;
  std::vector<AreaHost> cell_geom, down_geom;
  std::vector<CubicHost> up_tab, lvl_tab;  // cells -> level; level l+1 -> level l
  CubicHost init_tab{};                   // level 0 -> coarsest
  int chunk = 1;
  DevBuf cur, cmask, uvA, uvB, itxy, cells, cuv, ptrs;
  Profiler prof;
  ~ssg_eccflow() {
    if (stream) {
      (void)hipStreamSynchronize(stream);
      (void)hipStreamDestroy(stream);
    }
  }
};

extern "C" void ssg_default_eccflow_options(ssg_eccflow_options* o) {
  if (!o) return;
  memset(o, 0, sizeof(*o));
  o->input_smooth_sigma = 0;
  o->reference_smooth_sigma = 0;
  o->update_multiplier = 1.5;
  o->scale_factor = 0.5;
  o->noise_level = -1;
  o->max_iterations = 1;
  o->support_scale = 5;
  o->min_image_size = 4;
  o->max_pyramid_level = -1;
  o->downscale = SSG_ECCFLOW_DOWNSCALE_RECURSIVE_RESIZE;
}

extern "C" int ssg_eccflow_create(const ssg_eccflow_options* o, ssg_eccflow** out) {
  if (!out) return fail(SSG_E_INVALID_ARG, "out is null");
  *out = nullptr;
  auto f = std::make_unique<ssg_eccflow>();
  if (o) f->o = *o;
  else ssg_default_eccflow_options(&f->o);
  const auto& op = f->o;
  if (op.downscale < SSG_ECCFLOW_DOWNSCALE_RECURSIVE_RESIZE || op.downscale > SSG_ECCFLOW_DOWNSCALE_PYRAMID)
    return fail(SSG_E_INVALID_ARG, "unknown downscale method %d", op.downscale);
  if (op.support_scale < 0 || op.support_scale > 12) return fail(SSG_E_INVALID_ARG, "support_scale %d", op.support_scale);
  if (op.max_iterations < 0) return fail(SSG_E_INVALID_ARG, "max_iterations %d", op.max_iterations);
  if (op.downscale != SSG_ECCFLOW_DOWNSCALE_PYRAMID && !(op.scale_factor > 0 && op.scale_factor < 1))
    return fail(SSG_E_INVALID_ARG, "scale_factor must be in (0, 1)");
  HIPCHK(hipStreamCreateWithFlags(&f->stream, hipStreamNonBlocking));
  *out = f.release();
  return SSG_OK;
}

extern "C" void ssg_eccflow_destroy(ssg_eccflow* f) { delete f; }

namespace {
// Level sizes of c_eccflow::set_reference_image (ecc2.cc:2530-2595)
void eccflow_levels(ssg_eccflow* f, int w0, int h0) {
  const auto& o = f->o;
  const int min_size = std::max(4, o.min_image_size);
  const bool big_aspect = std::max(w0, h0) / std::min(w0, h0) >= 2;
  f->lw.assign(1, w0);
  f->lh.assign(1, h0);
  f->src.assign(1, -1);
  for (int level = 0;; ++level) {
    if (o.max_pyramid_level >= 0 && level >= o.max_pyramid_level) break;
    const int pw = f->lw.back(), ph = f->lh.back();
    int nw, nh, s;
    if (o.downscale == SSG_ECCFLOW_DOWNSCALE_PYRAMID) {
      nw = std::max(o.min_image_size, (pw + 1) / 2);
      nh = std::max(o.min_image_size, (ph + 1) / 2);
      if ((nw == pw && nh == ph) || std::min(nw, nh) <= min_size) break;
      s = level;
    } else {
      nw = std::max(o.min_image_size, (int)((pw + 1) * o.scale_factor));
      nh = std::max(o.min_image_size, (int)((ph + 1) * o.scale_factor));
      if ((nw == pw && nh == ph) || std::max(nw, nh) <= min_size) break;
      s = (o.downscale == SSG_ECCFLOW_DOWNSCALE_FULL_RESIZE || (big_aspect && std::min(nw, nh) <= min_size + 1)) ? 0
                                                                                                                 : level;
    }
    f->lw.push_back(nw);
    f->lh.push_back(nh);
    f->src.push_back(s);
  }
  f->L = (int)f->lw.size();
  f->cw.resize(f->L);
  f->ch.resize(f->L);
  f->loff.resize(f->L);
  f->coff.resize(f->L);
  f->lsum = f->csum = 0;
  for (int l = 0; l < f->L; ++l) {
    int cw = f->lw[l], ch = f->lh[l];
    for (int i = 0; i < o.support_scale; ++i) cw = (cw + 1) / 2, ch = (ch + 1) / 2;
    f->cw[l] = cw;
    f->ch[l] = ch;
    f->loff[l] = f->lsum;
    f->coff[l] = f->csum;
    f->lsum += (size_t)f->lw[l] * f->lh[l];
    f->csum += (size_t)cw * ch;
  }
}

// level l of frame-strided pyramids from its source level (pyrDown or INTER_AREA; masks nearest)
int eccflow_downscale(ssg_eccflow* f, float* pyr, long long fs, uint8_t* mpyr, int l, int n) {
  hipStream_t s = f->stream;
  const int sl = f->src[l];
  const int sw = f->lw[sl], sh = f->lh[sl], dw = f->lw[l], dh = f->lh[l];
  if (f->o.downscale == SSG_ECCFLOW_DOWNSCALE_PYRAMID)
    HIPCHK(launch_pyrdown(pyr + f->loff[sl], fs, sw, sw, sh, pyr + f->loff[l], fs, dw, dw, dh, n, BORDER_REFLECT101, s));
  else
    HIPCHK(launch_area(1, pyr + f->loff[sl], fs, sw, pyr + f->loff[l], fs, dw, f->down_geom[l].g, n, s));
  if (mpyr) HIPCHK(launch_nearest_u8(mpyr + f->loff[sl], fs, sw, sh, mpyr + f->loff[l], fs, dw, dh, n, s));
  return SSG_OK;
}
}  // namespace

extern "C" int ssg_eccflow_set_reference_image(ssg_eccflow* f, const ssg_image* image, const ssg_image* mask) {
  if (!f) return fail(SSG_E_INVALID_ARG, "null handle");
  CHK(check_image(image, "reference image"));
  if (image->type != SSG_32FC1) return fail(SSG_E_INVALID_ARG, "reference image: SSG_32FC1 expected");
  if (mask) {
    CHK(check_image(mask, "reference mask"));
    if (mask->type != SSG_8UC1 || mask->width != image->width || mask->height != image->height)
      return fail(SSG_E_INVALID_ARG, "reference mask: SSG_8UC1 of the image size expected");
  }
  f->have_ref = false;
  hipStream_t s = f->stream;
  const int w0 = image->width, h0 = image->height;
  eccflow_levels(f, w0, h0);
  f->w0 = w0;
  f->h0 = h0;
  // tables
  TabBuild tb;
  f->cell_geom.clear();
  f->down_geom.clear();
  f->up_tab.clear();
  f->lvl_tab.clear();
  for (int l = 0; l < f->L; ++l) {
    f->cell_geom.push_back(area_host(f->lw[l], f->lh[l], f->cw[l], f->ch[l], tb));
    f->down_geom.push_back(l > 0 && f->o.downscale != SSG_ECCFLOW_DOWNSCALE_PYRAMID
                               ? area_host(f->lw[f->src[l]], f->lh[f->src[l]], f->lw[l], f->lh[l], tb)
                               : AreaHost{});
    f->up_tab.push_back(cubic_host(f->cw[l], f->ch[l], f->lw[l], f->lh[l], tb));
    f->lvl_tab.push_back(l + 1 < f->L ? cubic_host(f->lw[l + 1], f->lh[l + 1], f->lw[l], f->lh[l], tb) : CubicHost{});
  }
  f->init_tab = cubic_host(w0, h0, f->lw[f->L - 1], f->lh[f->L - 1], tb);
  // the fused iteration covers whole-factor cells of k = 4..32 px (a power of two) that tile the level
  f->fuse_k.assign(f->L, 0);
  for (int l = 0; l < f->L; ++l) {
    const AreaGeom& g = f->cell_geom[l].g;
    const int k = g.isx;
    if (g.fast && g.isx == g.isy && k >= 4 && k <= 32 && (k & (k - 1)) == 0 && f->lw[l] == f->cw[l] * k &&
        f->lh[l] == f->ch[l] * k)
      f->fuse_k[l] = k;
  }
  CHK(f->tabs.ensure(tb.bytes.size()));
  HIPCHK(hipMemcpyAsync(f->tabs.p, tb.bytes.data(), tb.bytes.size(), hipMemcpyHostToDevice, s));
  const char* base = (const char*)f->tabs.p;
  for (auto& g : f->cell_geom) area_bind(g, base);
  for (auto& g : f->down_geom) area_bind(g, base);
  // reference pyramid
  f->ref_mask = mask != nullptr;
  CHK(f->ref.ensure(sizeof(float) * f->lsum));
  CHK(f->Ix.ensure(sizeof(float) * f->lsum));
  CHK(f->Iy.ensure(sizeof(float) * f->lsum));
  CHK(f->D.ensure(sizeof(float4) * f->csum));
  if (mask) CHK(f->rmask.ensure(f->lsum));
  {
    DevBuf t;
    const void* p;
    size_t pb;
    CHK(device_view(image, t, s, &p, &pb));
    HIPCHK(hipMemcpy2DAsync(f->ref.p, sizeof(float) * w0, p, pb, sizeof(float) * w0, h0, hipMemcpyDeviceToDevice, s));
    if (mask) {
      DevBuf tm;
      CHK(device_view(mask, tm, s, &p, &pb));
      HIPCHK(hipMemcpy2DAsync(f->rmask.p, w0, p, pb, w0, h0, hipMemcpyDeviceToDevice, s));
    }
    HIPCHK(hipStreamSynchronize(s));
  }
  for (int l = 1; l < f->L; ++l)
    CHK(eccflow_downscale(f, f->ref.as<float>(), 0, mask ? f->rmask.as<uint8_t>() : nullptr, l, 1));
  // per level: Ix, Iy, D
  const size_t n0 = (size_t)w0 * h0, c0 = (size_t)f->cw[0] * f->ch[0];
  DevBuf P2, P1, A2, A1;
  CHK(P2.ensure(sizeof(float2) * n0));
  CHK(P1.ensure(sizeof(float) * n0));
  CHK(A2.ensure(sizeof(float2) * c0));
  CHK(A1.ensure(sizeof(float) * c0));
  CHK(f->ref_ptr.ensure(sizeof(void*) * f->L));
  std::vector<const float*> lp(f->L);
  for (int l = 0; l < f->L; ++l) lp[l] = f->ref.as<float>() + f->loff[l];
  HIPCHK(hipMemcpyAsync(f->ref_ptr.p, lp.data(), sizeof(void*) * f->L, hipMemcpyHostToDevice, s));
  const double noise = f->o.noise_level >= 0 ? f->o.noise_level : 1e-3;
  const SepKernel Kx = diff_sep(true), Ky = diff_sep(false);
  for (int l = 0; l < f->L; ++l) {
    const int w = f->lw[l], h = f->lh[l];
    const float* const* sp = f->ref_ptr.as<const float*>() + l;
    HIPCHK(launch_sepfilter(sp, w, f->Ix.as<float>() + f->loff[l], 0, w, w, h, 1, Kx, BORDER_REPLICATE, s));
    HIPCHK(launch_sepfilter(sp, w, f->Iy.as<float>() + f->loff[l], 0, w, w, h, 1, Ky, BORDER_REPLICATE, s));
    HIPCHK(launch_flow_products(f->Ix.as<float>() + f->loff[l], f->Iy.as<float>() + f->loff[l], w * h,
                                P2.as<float2>(), P1.as<float>(), s));
    const AreaGeom& g = f->cell_geom[l].g;
    HIPCHK(launch_area(2, P2.as<float>(), 0, 2 * w, A2.as<float>(), 0, 2 * f->cw[l], g, 1, s));
    HIPCHK(launch_area(1, P1.as<float>(), 0, w, A1.as<float>(), 0, f->cw[l], g, 1, s));
    const float reg = noise > 0 ? (float)std::pow(1e-5 * noise / (1 << l), 4) : 0.f;
    HIPCHK(launch_flow_D(A2.as<float>(), A1.as<float>(), f->cw[l], f->ch[l], (float)f->o.update_multiplier, reg,
                         f->D.as<float4>() + f->coff[l], s));
  }
  HIPCHK(hipStreamSynchronize(s));
  // frames per launch: workspace of about 2 GiB at most
  const size_t per = sizeof(float) * f->lsum + f->lsum + 3 * sizeof(float2) * n0 + 2 * sizeof(float2) * c0;
  f->chunk = (int)std::max<size_t>(1, std::min<size_t>(64, ((size_t)2 << 30) / per));
  f->have_ref = true;
  return SSG_OK;
}

extern "C" int ssg_eccflow_num_levels(const ssg_eccflow* f, int32_t* n) {
  if (!f || !n) return fail(SSG_E_INVALID_ARG, "null argument");
  if (!f->have_ref) return fail(SSG_E_NOT_READY, "no reference image set");
  *n = f->L;
  return SSG_OK;
}

extern "C" int ssg_eccflow_level_size(const ssg_eccflow* f, int32_t level, int32_t* width, int32_t* height) {
  if (!f || !width || !height) return fail(SSG_E_INVALID_ARG, "null argument");
  if (!f->have_ref) return fail(SSG_E_NOT_READY, "no reference image set");
  if (level < 0 || level >= f->L) return fail(SSG_E_INVALID_ARG, "level %d of %d", level, f->L);
  *width = f->lw[level];
  *height = f->lh[level];
  return SSG_OK;
}

extern "C" int ssg_eccflow_reference_plane(const ssg_eccflow* f, int32_t level, int32_t which, ssg_image* out) {
  if (!f || !out) return fail(SSG_E_INVALID_ARG, "null argument");
  if (!f->have_ref) return fail(SSG_E_NOT_READY, "no reference image set");
  if (level < 0 || level >= f->L) return fail(SSG_E_INVALID_ARG, "level %d of %d", level, f->L);
  if (which < 0 || which > 6) return fail(SSG_E_INVALID_ARG, "plane %d", which);
  CHK(check_image(out, "out"));
  const bool cellp = which >= 3;
  const int w = cellp ? f->cw[level] : f->lw[level], h = cellp ? f->ch[level] : f->lh[level];
  if (out->type != SSG_32FC1 || out->width != w || out->height != h)
    return fail(SSG_E_INVALID_ARG, "out: SSG_32FC1 of %dx%d expected", w, h);
  const hipMemcpyKind kind = out->on_device ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost;
  if (!cellp) {
    const DevBuf& b = which == 0 ? f->ref : (which == 1 ? f->Ix : f->Iy);
    HIPCHK(hipMemcpy2DAsync(out->data, (size_t)out->step, b.as<float>() + f->loff[level], sizeof(float) * w,
                            sizeof(float) * w, h, kind, f->stream));
  } else {
    // component (which - 3) of the float4 cells, one row of w cells at a time
    const float* p = (const float*)(f->D.as<float4>() + f->coff[level]) + (which - 3);
    for (int y = 0; y < h; ++y)
      HIPCHK(hipMemcpy2DAsync((char*)out->data + (size_t)y * out->step, sizeof(float), p + (size_t)y * w * 4,
                              sizeof(float4), sizeof(float), w, kind, f->stream));
  }
  HIPCHK(hipStreamSynchronize(f->stream));
  return SSG_OK;
}

namespace {
// The level loop of c_eccflow::compute_uv (ecc2.cc:2770-2835) over n device frames given by
// host arrays of device pointers (pitches in elements).
int eccflow_run(ssg_eccflow* f, int n, const float* const* in, int in_pitch, const uint8_t* const* masks,
                int mask_pitch, const float2* const* rin, int rin_pitch, float2* const* rout, int rout_pitch) {
  hipStream_t s = f->stream;
  const int L = f->L, w0 = f->w0, h0 = f->h0;
  const size_t n0 = (size_t)w0 * h0, c0 = (size_t)f->cw[0] * f->ch[0];
  const int chunk = std::min(f->chunk, n);
  const bool hm = masks != nullptr;
  CHK(f->cur.ensure(sizeof(float) * f->lsum * chunk));
  if (hm) CHK(f->cmask.ensure(f->lsum * chunk));
  CHK(f->uvA.ensure(sizeof(float2) * n0 * chunk));
  CHK(f->uvB.ensure(sizeof(float2) * n0 * chunk));
  CHK(f->itxy.ensure(sizeof(float2) * n0 * chunk));
  CHK(f->cells.ensure(sizeof(float2) * c0 * chunk));
  CHK(f->cuv.ensure(sizeof(float2) * c0 * chunk));
  CHK(f->ptrs.ensure(sizeof(void*) * 4 * (size_t)n));
  std::vector<const void*> hp(4 * (size_t)n, nullptr);
  for (int i = 0; i < n; ++i) {
    hp[i] = in[i];
    hp[n + i] = hm ? (const void*)masks[i] : nullptr;
    hp[2 * n + i] = rin ? (const void*)rin[i] : nullptr;
    hp[3 * n + i] = rout[i];
  }
  HIPCHK(hipMemcpyAsync(f->ptrs.p, hp.data(), sizeof(void*) * hp.size(), hipMemcpyHostToDevice, s));
  const void* const* dp = f->ptrs.as<const void*>();
  const char* base = (const char*)f->tabs.p;
  const long long fs = (long long)f->lsum;
  const long long ufs = (long long)n0, cfs = (long long)c0;
  f->prof.start(s);
  for (int f0 = 0; f0 < n; f0 += chunk) {
    const int m = std::min(chunk, n - f0);
    float* cur = f->cur.as<float>();
    uint8_t* cm = hm ? f->cmask.as<uint8_t>() : nullptr;
    // current pyramid (setup_input_image, ecc2.cc:2674-2767)
    HIPCHK(launch_copy_frames((const float* const*)(dp + f0), in_pitch, cur, fs, w0, w0, h0, m, s));
    if (hm)
      HIPCHK(launch_mask_prepare((const uint8_t* const*)(dp + n + f0), mask_pitch, cm, fs, w0, w0, h0, w0, h0, m, s));
    for (int l = 1; l < L; ++l) CHK(eccflow_downscale(f, cur, fs, cm, l, m));
    bool mapped = false;
    // initial flow at the coarsest level
    float2* ua = f->uvA.as<float2>();
    float2* ub = f->uvB.as<float2>();
    if (rin) {
      FlowCubicArgs a = cubic_args(f->init_tab, base);
      a.maps = (const float2* const*)(dp + 2 * n + f0);
      a.map_pitch = rin_pitch;
      a.dst = ua;
      a.dfs = ufs;
      a.rx = (double)f->lw[L - 1] / (double)w0;
      a.ry = (double)f->lh[L - 1] / (double)h0;
      HIPCHK(launch_flow_cubic(CUBIC_INIT, a, m, s));
    } else {
      HIPCHK(hipMemsetAsync(ua, 0, sizeof(float2) * n0 * m, s));
    }
    for (int l = L - 1; l >= 0; --l) {
      const int w = f->lw[l], h = f->lh[l];
      if (l < L - 1) {
        // cv::multiply(uv, size_ratio, uv) at the coarser level, then the INTER_CUBIC upscale
        HIPCHK(launch_flow_scale(ua, ufs, f->lw[l + 1], f->lh[l + 1], (double)w / (double)f->lw[l + 1],
                                 (double)h / (double)f->lh[l + 1], m, s));
        FlowCubicArgs a = cubic_args(f->lvl_tab[l], base);
        a.src = ua;
        a.sfs = ufs;
        a.dst = ub;
        a.dfs = ufs;
        HIPCHK(launch_flow_cubic(CUBIC_PLAIN, a, m, s));
        std::swap(ua, ub);
      }
      FlowLevelArgs la{};
      la.cur = cur + f->loff[l];
      la.cfs = fs;
      la.cmask = hm ? cm + f->loff[l] : nullptr;
      la.mfs = fs;
      la.ref = f->ref.as<float>() + f->loff[l];
      la.Ix = f->Ix.as<float>() + f->loff[l];
      la.Iy = f->Iy.as<float>() + f->loff[l];
      la.rmask = f->ref_mask ? f->rmask.as<uint8_t>() + f->loff[l] : nullptr;
      la.w = w;
      la.h = h;
      FlowCubicArgs up = cubic_args(f->up_tab[l], base);
      up.src = f->cuv.as<float2>();
      up.sfs = cfs;
      up.dst = ua;
      up.dfs = ufs;
      if (f->fuse_k[l] && f->o.max_iterations > 0) {
        FlowIterArgs ia{};
        ia.L = la;
        ia.uv = ua;
        ia.ufs = ufs;
        ia.cuv = f->cuv.as<float2>();
        ia.cfs = cfs;
        ia.cw = f->cw[l];
        ia.ch = f->ch[l];
        ia.k = f->fuse_k[l];
        ia.xs = up.xs;
        ia.xa = up.xa;
        ia.ys = up.ys;
        ia.ya = up.ya;
        ia.cells = f->cells.as<float>();
        ia.cellfs = 2 * cfs;
        ia.maps = (float2* const*)(dp + 3 * n + f0);
        ia.map_pitch = rout_pitch;
        for (int it = 0; it < f->o.max_iterations; ++it) {
          HIPCHK(launch_flow_iter(it > 0, true, false, ia, m, s));
          HIPCHK(launch_flow_solve(f->cells.as<float>(), 2 * cfs, f->D.as<float4>() + f->coff[l], f->cw[l], f->ch[l],
                                   f->cuv.as<float2>(), cfs, m, s));
        }
        // the last update; at level 0 it goes straight into the caller's maps
        HIPCHK(launch_flow_iter(true, false, l == 0, ia, m, s));
        if (l == 0) mapped = true;
        continue;
      }
      for (int it = 0; it < f->o.max_iterations; ++it) {
        HIPCHK(launch_flow_residual(la, ua, ufs, f->itxy.as<float2>(), ufs, m, s));
        HIPCHK(launch_area(2, f->itxy.as<float>(), 2 * ufs, 2 * w, f->cells.as<float>(), 2 * cfs, 2 * f->cw[l],
                           f->cell_geom[l].g, m, s));
        HIPCHK(launch_flow_solve(f->cells.as<float>(), 2 * cfs, f->D.as<float4>() + f->coff[l], f->cw[l], f->ch[l],
                                 f->cuv.as<float2>(), cfs, m, s));
        HIPCHK(launch_flow_cubic(CUBIC_ADD, up, m, s));
      }
    }
    if (!mapped) HIPCHK(launch_flow_to_remap(ua, ufs, w0, h0, (float2* const*)(dp + 3 * n + f0), rout_pitch, m, s));
  }
  f->prof.stop(s, SSG_PROF_ECCFLOW);
  if (f->prof.on) {
    // algorithmic bytes per frame: level-0 copy (8 N0) and the downscales (4 N_src + 4 N_l),
    // per level and iteration the fused minimum 32 B/px (uv r+w, current, reference, Ix, Iy),
    // level transitions (8 N_{l+1} + 8 N_l), the initial flow and the final remap
    double b = 8.0 * n0;
    for (int l = 1; l < L; ++l) b += 4.0 * f->lw[f->src[l]] * f->lh[f->src[l]] + 4.0 * f->lw[l] * f->lh[l];
    for (int l = 0; l < L; ++l) {
      const double nl = (double)f->lw[l] * f->lh[l];
      b += f->o.max_iterations * 32.0 * nl;
      if (l < L - 1) b += 8.0 * f->lw[l + 1] * f->lh[l + 1] + 8.0 * nl;
    }
    b += (rin ? 8.0 * n0 : 0.0) + 8.0 * f->lw[L - 1] * f->lh[L - 1] + 16.0 * n0;
    f->prof.add_bytes(SSG_PROF_ECCFLOW, b * n);
  }
  return SSG_OK;
}
}  // namespace

extern "C" int ssg_eccflow_compute_batch(ssg_eccflow* f, int n, const ssg_image* inputs, const ssg_image* input_masks,
                                         const ssg_image* rmaps_in, ssg_image* rmaps_out) {
  if (!f) return fail(SSG_E_INVALID_ARG, "null handle");
  if (n <= 0) return SSG_OK;
  if (!inputs || !rmaps_out) return fail(SSG_E_INVALID_ARG, "null inputs or rmaps_out");
  if (!f->have_ref) return fail(SSG_E_NOT_READY, "no reference image set");
  auto uniform = [&](const ssg_image* im, int type, int es, const char* what) -> int {
    for (int i = 0; i < n; ++i) {
      CHK(check_image(&im[i], what));
      if (!im[i].on_device || im[i].type != type || im[i].width != f->w0 || im[i].height != f->h0 ||
          im[i].step != im[0].step || im[i].step % es)
        return fail(SSG_E_INVALID_ARG, "%s %d: device images of the reference size, one type and step", what, i);
    }
    return SSG_OK;
  };
  CHK(uniform(inputs, SSG_32FC1, 4, "input"));
  if (input_masks) CHK(uniform(input_masks, SSG_8UC1, 1, "input mask"));
  if (rmaps_in) CHK(uniform(rmaps_in, SSG_32FC2, 8, "rmap_in"));
  CHK(uniform(rmaps_out, SSG_32FC2, 8, "rmap_out"));
  std::vector<const float*> in(n);
  std::vector<const uint8_t*> mk(input_masks ? n : 0);
  std::vector<const float2*> ri(rmaps_in ? n : 0);
  std::vector<float2*> ro(n);
  for (int i = 0; i < n; ++i) {
    in[i] = (const float*)inputs[i].data;
    if (input_masks) mk[i] = (const uint8_t*)input_masks[i].data;
    if (rmaps_in) ri[i] = (const float2*)rmaps_in[i].data;
    ro[i] = (float2*)rmaps_out[i].data;
  }
  CHK(eccflow_run(f, n, in.data(), (int)(inputs[0].step / 4), input_masks ? mk.data() : nullptr,
                  input_masks ? (int)input_masks[0].step : 0, rmaps_in ? ri.data() : nullptr,
                  rmaps_in ? (int)(rmaps_in[0].step / 8) : 0, ro.data(), (int)(rmaps_out[0].step / 8)));
  HIPCHK(hipStreamSynchronize(f->stream));
  f->prof.collect();
  return SSG_OK;
}

extern "C" int ssg_eccflow_compute(ssg_eccflow* f, const ssg_image* input, const ssg_image* input_mask,
                                   const ssg_image* rmap_in, ssg_image* rmap_out) {
  if (!f) return fail(SSG_E_INVALID_ARG, "null handle");
  if (!f->have_ref) return fail(SSG_E_NOT_READY, "no reference image set");
  CHK(check_image(input, "input"));
  CHK(check_image(rmap_out, "rmap_out"));
  auto same = [&](const ssg_image* im, int type, const char* what) -> int {
    CHK(check_image(im, what));
    if (im->type != type || im->width != f->w0 || im->height != f->h0)
      return fail(SSG_E_INVALID_ARG, "%s: type %d of the reference size %dx%d expected", what, type, f->w0, f->h0);
    if (im->on_device && im->step % elem_size(type)) return fail(SSG_E_INVALID_ARG, "%s: step", what);
    return SSG_OK;
  };
  CHK(same(input, SSG_32FC1, "input"));
  if (input_mask) CHK(same(input_mask, SSG_8UC1, "input mask"));
  if (rmap_in) CHK(same(rmap_in, SSG_32FC2, "rmap_in"));
  CHK(same(rmap_out, SSG_32FC2, "rmap_out"));
  hipStream_t s = f->stream;
  DevBuf ti, tm, tr, to;
  const void *pi, *pm = nullptr, *pr = nullptr;
  size_t bi, bm = 0, br = 0;
  CHK(device_view(input, ti, s, &pi, &bi));
  if (input_mask) CHK(device_view(input_mask, tm, s, &pm, &bm));
  if (rmap_in) CHK(device_view(rmap_in, tr, s, &pr, &br));
  float2* po;
  size_t bo;
  if (rmap_out->on_device) {
    po = (float2*)rmap_out->data;
    bo = (size_t)rmap_out->step;
  } else {
    bo = sizeof(float2) * f->w0;
    CHK(to.ensure(bo * f->h0));
    po = to.as<float2>();
  }
  const float* in1 = (const float*)pi;
  const uint8_t* mk1 = (const uint8_t*)pm;
  const float2* ri1 = (const float2*)pr;
  CHK(eccflow_run(f, 1, &in1, (int)(bi / 4), input_mask ? &mk1 : nullptr, (int)bm, rmap_in ? &ri1 : nullptr,
                  (int)(br / 8), &po, (int)(bo / 8)));
  if (!rmap_out->on_device)
    HIPCHK(hipMemcpy2DAsync(rmap_out->data, (size_t)rmap_out->step, po, bo, bo, f->h0, hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  f->prof.collect();
  return SSG_OK;
}

extern "C" int ssg_eccflow_set_profiling(ssg_eccflow* f, int enable) {
  if (!f) return fail(SSG_E_INVALID_ARG, "null handle");
  f->prof.on = enable != 0;
  return SSG_OK;
}

extern "C" int ssg_eccflow_get_profile(ssg_eccflow* f, ssg_profile* out, int reset) {
  if (!f || !out) return fail(SSG_E_INVALID_ARG, "null argument");
  *out = f->prof.acc;
  if (reset) f->prof.acc = ssg_profile{};
  return SSG_OK;
}

// =====================================================================================================
// registration with the eccflow refine
// =====================================================================================================
namespace {
int refine_current_remap(ssg_registration* r, const ssg_image* image, const ssg_image* mask) {
  const size_t pitch = sizeof(float2) * r->ref_w;
  CHK(r->cur_map.ensure(pitch * r->ref_h));
  const MapCoef mc = map_coef(r->o.motion_type, r->last.params, nullptr);
  HIPCHK(launch_create_map(r->o.motion_type, mc, r->cur_map.as<float>(), (int)(pitch / 4), r->ref_w, r->ref_h,
                           r->stream));
  HIPCHK(hipStreamSynchronize(r->stream));
  ssg_image m{r->cur_map.p, r->ref_w, r->ref_h, (int64_t)pitch, SSG_32FC2, 1};
  return ssg_eccflow_compute(r->flow, image, mask, &m, &m);
}
}  // namespace

extern "C" int ssg_registration_current_remap(ssg_registration* r, ssg_image* map) {
  if (!r) return fail(SSG_E_INVALID_ARG, "null handle");
  if (!r->have_status || !r->last.ok) return fail(SSG_E_NOT_READY, "no successfully registered frame");
  if (!r->flow) return ssg_registration_create_remap(r, &r->last, map);
  CHK(check_image(map, "map"));
  if (map->type != SSG_32FC2 || map->width != r->ref_w || map->height != r->ref_h)
    return fail(SSG_E_INVALID_ARG, "map must be SSG_32FC2 of the reference size %dx%d", r->ref_w, r->ref_h);
  const size_t row = sizeof(float2) * r->ref_w;
  HIPCHK(hipMemcpy2DAsync(map->data, (size_t)map->step, r->cur_map.p, row, row, r->ref_h,
                          map->on_device ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost, r->stream));
  HIPCHK(hipStreamSynchronize(r->stream));
  return SSG_OK;
}

extern "C" int ssg_registration_register_batch_remaps(ssg_registration* r, int n, const ssg_image* frames,
                                                      const ssg_image* masks, ssg_registration_status* status,
                                                      ssg_image* remaps) {
  if (!r) return fail(SSG_E_INVALID_ARG, "null handle");
  if (n < 0) return fail(SSG_E_INVALID_ARG, "n < 0");
  if (n == 0) return SSG_OK;
  if (!remaps) return fail(SSG_E_INVALID_ARG, "null remaps");
  for (int f = 0; f < n; ++f) {
    CHK(check_image(&remaps[f], "remap"));
    if (!remaps[f].on_device || remaps[f].type != SSG_32FC2 || remaps[f].width != r->ref_w ||
        remaps[f].height != r->ref_h || remaps[f].step != remaps[0].step || remaps[f].step % 8)
      return fail(SSG_E_INVALID_ARG, "remap %d: device SSG_32FC2 maps of the reference size and one step", f);
  }
  DevBuf gb;
  std::vector<ssg_image> gv;
  CHK(ecc_channel_views(r, n, frames, gb, gv, &frames));
  if (r->flow) {
    for (int f = 0; f < n; ++f) {
      if (!frames[f].on_device || frames[f].step != frames[0].step || frames[f].step % 4)
        return fail(SSG_E_INVALID_ARG, "frame %d: the eccflow refine takes device frames of one step", f);
      if (masks && (!masks[f].on_device || masks[f].step != masks[0].step))
        return fail(SSG_E_INVALID_ARG, "mask %d: the eccflow refine takes device masks of one step", f);
    }
  }
  r->have_status = false;
  CHK(register_n(r, n, frames, masks, status, false));
  hipStream_t s = r->stream;
  const int mpitch = (int)(remaps[0].step / 4);
  for (int f = 0; f < n; ++f) {
    const MapCoef mc = map_coef(r->o.motion_type, status[f].params, nullptr);
    HIPCHK(launch_create_map(r->o.motion_type, mc, (float*)remaps[f].data, mpitch, r->ref_w, r->ref_h, s));
  }
  HIPCHK(hipStreamSynchronize(s));
  if (!r->flow) return SSG_OK;
  // refine the maps of the accepted frames in place
  std::vector<const float*> in;
  std::vector<const uint8_t*> mk;
  std::vector<const float2*> ri;
  std::vector<float2*> ro;
  for (int f = 0; f < n; ++f) {
    if (!status[f].ok) continue;
    in.push_back((const float*)frames[f].data);
    if (masks) mk.push_back((const uint8_t*)masks[f].data);
    ri.push_back((const float2*)remaps[f].data);
    ro.push_back((float2*)remaps[f].data);
  }
  if (in.empty()) return SSG_OK;
  CHK(eccflow_run(r->flow, (int)in.size(), in.data(), (int)(frames[0].step / 4), masks ? mk.data() : nullptr,
                  masks ? (int)masks[0].step : 0, ri.data(), (int)(remaps[0].step / 8), ro.data(),
                  (int)(remaps[0].step / 8)));
  HIPCHK(hipStreamSynchronize(r->flow->stream));
  r->flow->prof.collect();
  return SSG_OK;
}

extern "C" int ssg_registration_remap(ssg_registration* r, const ssg_image* src, const ssg_image* src_mask,
                                      ssg_image* dst, ssg_image* dst_mask, int32_t interpolation, int32_t border_mode,
                                      double border_value) {
  if (!r) return fail(SSG_E_INVALID_ARG, "null handle");
  if (!r->have_status || !r->last.ok) return fail(SSG_E_NOT_READY, "no successfully registered frame");
  if (interpolation < 0) interpolation = r->o.interpolation;
  if (border_mode < 0) {
    border_mode = r->o.border_mode;
    border_value = r->o.border_value[0];
  }
  if (r->flow) {
    ssg_image m{r->cur_map.p, r->ref_w, r->ref_h, (int64_t)(sizeof(float2) * r->ref_w), SSG_32FC2, 1};
    return ssg_remap(src, src_mask, &m, nullptr, r->o.motion_type, nullptr, dst, dst_mask, interpolation, border_mode,
                     border_value);
  }
  return ssg_remap(src, src_mask, nullptr, r->last.matrix, r->o.motion_type, r->last.params, dst, dst_mask,
                   interpolation, border_mode, border_value);
}

#include "ssg_input_api.inc"
#include "ssg_stack_api.inc"
#include "ssg_finish_api.inc"

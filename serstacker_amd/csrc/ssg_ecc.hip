// ssg_ecc.hip -- batched forward-additive ECC for gfx950.
//
// One launch = one ECC iteration (ecc2.cc:1295-1362) for every active frame
// of a batch.  A persistent grid walks (frame, tile) items.  Per 64x64 tile a
// block:
//   1. samples the current level image at W(x;p) into LDS with a 2-px halo
//      (cv::remap INTER_LINEAR, BORDER_REPLICATE, 1/32-px positions),
//   2. differentiates the warped tile (ecc_differentiate: d5 x s3 separable,
//      BORDER_REPLICATE of the warped image),
//   3. evaluates the warped-mask predicate (remap(mask) >= 255, exact u8
//      fixed point) and the reference mask,
//   4. accumulates the affine moments of the steepest-descent images:
//      sum (gx^2, gx gy, gy^2) x {1,x,y,x^2,xy,y^2}, sum (gx,gy) x (g', f', 1)
//      x {1,x,y}, and the masked first/second moments of f and g, with
//      f' = f - cf and g' = g - cg centred to limit cancellation,
//   5. reduces the 41 moments across the wave by recursive halving, across
//      waves through LDS, and stores the tile's moments in global
//      coordinates (FP64).
// A second launch (one block per frame) sums each frame's tiles in a fixed
// order (deterministic), projects the affine moments onto the parameters of
// the transform (J = M J_affine for translation and euclidean), and solves
// the normal equations in FP64 (Cholesky; failure rule of cv::invert).  It
// updates the parameters, eps and the frame's convergence state and appends
// the frames still iterating to the next active list.  The kernel boundary
// between the two launches makes the partials visible: no block waits on
// another and no device-scope fence runs in the hot loop.
#include <float.h>

#include <algorithm>

#include "../../include/ssg.h"
#include "ssg_engine.hpp"

namespace ssg {

constexpr float D5_K1 = 2.f / 3.f;    // d5[3] of ecc_differentiate (ecc2.cc:149)
constexpr float D5_K2 = -1.f / 12.f;  // d5[4]

__device__ __forceinline__ int motion_nparams(int m) {
  return m == MOTION_TRANSLATION ? 2 : (m == MOTION_EUCLIDEAN ? 3 : (m == MOTION_SCALED_EUCLIDEAN ? 4 : 6));
}

// map coefficients of a parameter vector (see warp_xy)
__device__ void update_map(int motion, FrameXf& X) {
  if (motion == MOTION_TRANSLATION) {
    X.mc.c[0] = X.p[0];
    X.mc.c[1] = X.p[1];
  } else if (motion == MOTION_AFFINE) {
    for (int i = 0; i < 6; ++i) X.mc.c[i] = X.p[i];
  } else {
    const float ang = X.p[2];
    const float s = motion == MOTION_SCALED_EUCLIDEAN ? X.p[3] : 1.f;
    X.mc.c[0] = (float)sin((double)ang);
    X.mc.c[1] = (float)cos((double)ang);
    X.mc.c[2] = s;
    X.mc.c[3] = X.p[0];
    X.mc.c[4] = X.p[1];
    X.mc.c[5] = X.C[0];
    X.mc.c[6] = X.C[1];
  }
}

// c_image_transform::scale_transfrom (c_image_transform.cc:130-134, 502-507, 911-915)
__device__ void scale_xf(int motion, FrameXf& X, double f) {
  if (motion == MOTION_AFFINE) {
    X.p[2] = (float)((double)X.p[2] * f);
    X.p[5] = (float)((double)X.p[5] * f);
  } else {
    X.p[0] = (float)((double)X.p[0] * f);
    X.p[1] = (float)((double)X.p[1] * f);
    if (motion != MOTION_TRANSLATION) {
      X.C[0] = (float)((double)X.C[0] * f);
      X.C[1] = (float)((double)X.C[1] * f);
    }
  }
}

// eps() of the transforms (c_image_transform.cc:136-139, 509-523, 917-924).
// Euclidean with a fixed scale reads ds from the transform's scale (1), as the
// reference does, so its eps never drops below max(W,H).
__device__ float xf_eps(int motion, const float* dp, int w, int h) {
  if (motion == MOTION_TRANSLATION) return (float)sqrt((double)dp[0] * dp[0] + (double)dp[1] * dp[1]);
  if (motion == MOTION_AFFINE) {
    const double a = (double)w * dp[0], b = (double)h * dp[1], c = dp[2];
    const double d = (double)w * dp[3], e = (double)h * dp[4], f = dp[5];
    return (float)sqrt(a * a + b * b + c * c + d * d + e * e + f * f);
  }
  const double ds = motion == MOTION_SCALED_EUCLIDEAN ? (double)dp[3] : 1.0;
  const double sa = sin((double)dp[2]);
  const double m = (double)(w > h ? w : h);
  const double ws = w * sa, hs = h * sa, ms = m * ds;
  return (float)sqrt((double)dp[0] * dp[0] + (double)dp[1] * dp[1] + ws * ws + hs * hs + ms * ms);
}

// global-coordinate moment k of a tile from its tile-local moments
__device__ double expand_moment(int k, const float* loc, double X0, double Y0) {
  if (k < 18) {
    const int b = (k / 6) * 6, i = k - b;
    const double s0 = loc[b], sx = loc[b + 1], sy = loc[b + 2];
    switch (i) {
      case 0: return s0;
      case 1: return X0 * s0 + sx;
      case 2: return Y0 * s0 + sy;
      case 3: return X0 * X0 * s0 + 2.0 * X0 * sx + (double)loc[b + 3];
      case 4: return X0 * Y0 * s0 + X0 * sy + Y0 * sx + (double)loc[b + 4];
      default: return Y0 * Y0 * s0 + 2.0 * Y0 * sy + (double)loc[b + 5];
    }
  }
  if (k < 36) {
    const int b = 18 + ((k - 18) / 3) * 3, i = k - b;
    const double s0 = loc[b];
    if (i == 0) return s0;
    if (i == 1) return X0 * s0 + (double)loc[b + 1];
    return Y0 * s0 + (double)loc[b + 2];
  }
  return loc[k];
}

// Solve of one ECC iteration (ecc2.cc:1320-1362) from the frame's moments.
// Runs on one thread; every matrix lives in the LDS scratch sm (the moments G
// occupy sm[0..63]) so the hot path of the kernel keeps its register budget.
template <int MOTION, bool FULL>
__device__ void ecc_solve(double* sm, FrameXf& X, FrameRes& R, const EccPassArgs& a) {
  const double* G = sm;
  double* Ha = sm + 64;   // 6x6
  double* H = sm + 100;   // NP x NP
  double* L = sm + 136;   // NP x NP
  double* T = sm + 172;   // NP x 6
  double* M = sm + 208;   // NP x 6
  double* ea = sm + 244;  // 6
  double* e = sm + 250;   // NP
  double* z = sm + 256;   // NP
  // ---- masked statistics (cv::meanStdDev, population stdev)
  double n, sf, sff, sg, sgg;
  if (FULL) {
    n = G[36]; sf = G[37]; sff = G[38]; sg = G[39]; sgg = G[40];
  } else {
    n = G[9]; sf = G[10]; sff = G[11]; sg = G[12]; sgg = G[13];
  }
  const double inv = n > 0 ? 1.0 / n : 0.0;
  const double mf = sf * inv, mg = sg * inv;  // means of f', g'
  const double vf = sff * inv - mf * mf, vg = sgg * inv - mg * mg;
  const double stdf = sqrt(vf > 0 ? vf : 0.0), stdg = sqrt(vg > 0 ? vg : 0.0);
  const double ratio = stdg / stdf;
  const double kc = mg - ratio * mf;  // (mu_g - cg) - ratio (mu_f - cf)

  constexpr int NP = MOTION == MOTION_TRANSLATION ? 2
                     : (MOTION == MOTION_EUCLIDEAN ? 3 : (MOTION == MOTION_SCALED_EUCLIDEAN ? 4 : 6));
  if (!FULL) {
    // J = [gx, gy]
    H[0] = G[0]; H[1] = G[1]; H[2] = G[1]; H[3] = G[2];
    e[0] = G[3] - ratio * G[5] - kc * G[7];
    e[1] = G[4] - ratio * G[6] - kc * G[8];
  } else {
    // affine moments, J_aff = [gx X, gx Y, gx, gy X, gy Y, gy]; A, B, C = {1, X, Y, XX, XY, YY}
    const double *A = G, *B = G + 6, *Cc = G + 12;
    const double up[21] = {A[3], A[4], A[1], B[3], B[4], B[1],
                                 A[5], A[2], B[4], B[5], B[2],
                                       A[0], B[1], B[2], B[0],
                                             Cc[3], Cc[4], Cc[1],
                                                    Cc[5], Cc[2],
                                                           Cc[0]};
    for (int i = 0, q = 0; i < 6; ++i)
      for (int j = i; j < 6; ++j, ++q) Ha[i * 6 + j] = Ha[j * 6 + i] = up[q];
    const double *P = G + 18, *Q = G + 21, *Rr = G + 24, *S = G + 27, *GX = G + 30, *GY = G + 33;
    for (int i = 0; i < 6; ++i) {
      const int k = (i % 3 == 0) ? 1 : ((i % 3 == 1) ? 2 : 0);  // X, Y, 1
      const double bg = i < 3 ? P[k] : Q[k];
      const double bf = i < 3 ? Rr[k] : S[k];
      const double b1 = i < 3 ? GX[k] : GY[k];
      ea[i] = bg - ratio * bf - kc * b1;
    }
    if (MOTION == MOTION_AFFINE) {
      for (int i = 0; i < 36; ++i) H[i] = Ha[i];
      for (int i = 0; i < 6; ++i) e[i] = ea[i];
    } else {
      // J = M J_aff (c_image_transform.cc:587-668 rewritten on the affine basis)
      for (int i = 0; i < NP * 6; ++i) M[i] = 0;
      M[0 * 6 + 2] = 1;
      M[1 * 6 + 5] = 1;
      {
        const double sa = X.mc.c[0], ca = X.mc.c[1], s = X.mc.c[2], Cx = X.C[0], Cy = X.C[1];
        M[2 * 6 + 0] = -s * sa; M[2 * 6 + 1] = -s * ca; M[2 * 6 + 2] = s * (sa * Cx + ca * Cy);
        M[2 * 6 + 3] = s * ca;  M[2 * 6 + 4] = -s * sa; M[2 * 6 + 5] = s * (-ca * Cx + sa * Cy);
        if (NP > 3) {
          M[3 * 6 + 0] = ca; M[3 * 6 + 1] = -sa; M[3 * 6 + 2] = -(ca * Cx - sa * Cy);
          M[3 * 6 + 3] = sa; M[3 * 6 + 4] = ca;  M[3 * 6 + 5] = -(sa * Cx + ca * Cy);
        }
      }
      for (int i = 0; i < NP; ++i)
        for (int j = 0; j < 6; ++j) {
          double t = 0;
          for (int k = 0; k < 6; ++k) t += M[i * 6 + k] * Ha[k * 6 + j];
          T[i * 6 + j] = t;
        }
      for (int i = 0; i < NP; ++i) {
        double t = 0;
        for (int k = 0; k < 6; ++k) t += M[i * 6 + k] * ea[k];
        e[i] = t;
        for (int j = 0; j < NP; ++j) {
          double u = 0;
          for (int k = 0; k < 6; ++k) u += T[i * 6 + k] * M[j * 6 + k];
          H[i * NP + j] = u;
        }
      }
    }
  }

  // ---- cv::invert(DECOMP_CHOLESKY) failure rule: n <= 3 inverts through the determinant and
  // fails only when it is 0 (a NaN Hessian goes on, as it does in the reference);
  // n > 3 runs the Cholesky factorisation and fails on a pivot below FLT_EPSILON
  bool failed = false;
  if constexpr (NP <= 3) {
    double inv[9];
    if (NP == 2) {
      const double d = H[0] * H[3] - H[1] * H[2];
      failed = d == 0.0;
      const double r = 1.0 / d;
      inv[0] = H[3] * r; inv[1] = -H[1] * r; inv[2] = -H[2] * r; inv[3] = H[0] * r;
    } else {
      const double d = H[0] * (H[4] * H[8] - H[5] * H[7]) - H[1] * (H[3] * H[8] - H[5] * H[6]) +
                       H[2] * (H[3] * H[7] - H[4] * H[6]);
      failed = d == 0.0;
      const double r = 1.0 / d;
      inv[0] = (H[4] * H[8] - H[5] * H[7]) * r;
      inv[1] = (H[2] * H[7] - H[1] * H[8]) * r;
      inv[2] = (H[1] * H[5] - H[2] * H[4]) * r;
      inv[3] = (H[5] * H[6] - H[3] * H[8]) * r;
      inv[4] = (H[0] * H[8] - H[2] * H[6]) * r;
      inv[5] = (H[2] * H[3] - H[0] * H[5]) * r;
      inv[6] = (H[3] * H[7] - H[4] * H[6]) * r;
      inv[7] = (H[1] * H[6] - H[0] * H[7]) * r;
      inv[8] = (H[0] * H[4] - H[1] * H[3]) * r;
    }
    for (int i = 0; i < NP; ++i) {
      double t = 0;
      for (int j = 0; j < NP; ++j) t += inv[i * NP + j] * e[j];
      z[i] = t;
    }
  } else {
    for (int i = 0; i < NP && !failed; ++i) {
      for (int j = 0; j < i; ++j) {
        double s = H[i * NP + j];
        for (int k = 0; k < j; ++k) s -= L[i * NP + k] * L[j * NP + k];
        L[i * NP + j] = s / L[j * NP + j];
      }
      double s = H[i * NP + i];
      for (int k = 0; k < i; ++k) s -= L[i * NP + k] * L[i * NP + k];
      if (s < (double)FLT_EPSILON) failed = true;
      else L[i * NP + i] = sqrt(s);
    }
    if (!failed) {
      for (int i = 0; i < NP; ++i) {
        double s = e[i];
        for (int j = 0; j < i; ++j) s -= L[i * NP + j] * z[j];
        z[i] = s / L[i * NP + i];
      }
      for (int i = NP - 1; i >= 0; --i) {
        double s = z[i];
        for (int j = i + 1; j < NP; ++j) s -= L[j * NP + i] * z[j];
        z[i] = s / L[i * NP + i];
      }
    }
  }
  X.iters += 1;
  const int k = X.iters;
  if (failed) {
    R.level_iters[a.pass][a.level] = k;
    R.level_failed[a.pass][a.level] = 1;
    X.flags &= ~XF_ACTIVE;
    return;
  }
  float dp[6] = {0, 0, 0, 0, 0, 0};
  for (int i = 0; i < NP; ++i) {
    dp[i] = (float)(-a.update_step_scale * z[i]);
    X.p[i] = X.p[i] + dp[i];
  }
  update_map(MOTION, X);
  X.cg = (float)((double)X.cg + mg);
  const float eps = xf_eps(MOTION, dp, a.w, a.h);
  X.eps = eps;
  R.eps = eps;
  if (eps < a.max_eps) {
    R.level_iters[a.pass][a.level] = k;
    X.flags &= ~XF_ACTIVE;
  } else if (k >= a.max_iterations) {
    // while (_num_iterations++ < _max_iterations): exhaustion leaves max + 1
    R.level_iters[a.pass][a.level] = k + 1;
    X.flags &= ~XF_ACTIVE;
  }
}

// deterministic block-wide compaction of active frames (ascending order)
__device__ void compact_active(const FrameXf* xf, int nframes, int* out, int* out_count) {
  __shared__ int s_wave[16];
  __shared__ int s_base;
  if (threadIdx.x == 0) s_base = 0;
  __syncthreads();
  const int lane = threadIdx.x & (WAVE - 1), wv = threadIdx.x / WAVE, nw = blockDim.x / WAVE;
  for (int base = 0; base < nframes; base += blockDim.x) {
    const int f = base + threadIdx.x;
    const bool act = f < nframes && (xf[f].flags & XF_ACTIVE);
    const unsigned long long bal = __ballot(act);
    const int before = __popcll(bal & ((1ull << lane) - 1ull));
    if (lane == 0) s_wave[wv] = __popcll(bal);
    __syncthreads();
    int off = s_base;
    for (int i = 0; i < wv; ++i) off += s_wave[i];
    if (act) out[off + before] = f;
    __syncthreads();
    if (threadIdx.x == 0) {
      int t = 0;
      for (int i = 0; i < nw; ++i) t += s_wave[i];
      s_base += t;
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) *out_count = s_base;
}

// ------------------------------------------------------------------------------------------------
#include "ssg_ecc_pass.inc"

// One block per frame of the batch: sums the frame's tile moments in tile
// order (deterministic), solves, and appends the frame to the next active
// list while it keeps iterating.  Runs after k_ecc_pass, so the kernel
// boundary makes every tile's partial visible without device-scope fences.
template <int MOTION, bool FULL>
__global__ void __launch_bounds__(256) k_ecc_solve(EccPassArgs a) {
  constexpr int NM = FULL ? NMOM_FULL : NMOM_TRANS;
  __shared__ double dred[ECC_WAVES][64];
  __shared__ double s_solve[272];
  const int f = blockIdx.x;
  if (!(a.xf[f].flags & XF_ACTIVE)) return;
  const int tid = threadIdx.x, lane = tid & (WAVE - 1), wv = tid / WAVE;
  {
    const int k = lane, j = wv;
    double s = 0;
    if (k < NM) {
      const double* base = reinterpret_cast<const double*>(a.partials) + (size_t)f * a.ntiles * NM;
      s = ordered_strided_sum(base + k, j, a.ntiles, ECC_WAVES, NM);
    }
    dred[j][k] = s;
  }
  __syncthreads();
  if (tid < NM) s_solve[tid] = ((dred[0][tid] + dred[1][tid]) + dred[2][tid]) + dred[3][tid];
  __syncthreads();
  if (tid == 0) {
    ecc_solve<MOTION, FULL>(s_solve, a.xf[f], a.res[f], a);
    if (a.xf[f].flags & XF_ACTIVE) {
      // the list order does not affect any frame's result
      const int at = atomicAdd(a.active_count_out, 1);
      a.active_out[at] = f;
    }
  }
}

static int device_cus() {
  static const int cus = [] {
    int dev = 0, n = 0;
    (void)hipGetDevice(&dev);
    if (hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || n <= 0) n = 256;
    return n;
  }();
  return cus;
}

// upper bound of the persistent grids (8 blocks per CU); each pass kernel
// launches at most the blocks that are resident at once at its own VGPR/LDS use
int ecc_persistent_grid() { return device_cus() * tune_value(SSG_TUNE_ECC_GRID); }

template <int MOTION, bool F, bool CM, bool RM>
static int pass_resident_grid() {
  static const int g = [] {
    int nb = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_ecc_pass<MOTION, F, CM, RM>, ECC_THREADS, 0) !=
            hipSuccess || nb <= 0)
      nb = 4;
    return (device_cus() * nb + 7) / 8 * 8;
  }();
  return g;
}

template <int MOTION>
static hipError_t launch_pass_m(const EccPassArgs& a, bool full, int grid, hipStream_t s) {
  const bool cm = a.cmask != nullptr, rm = a.rmask != nullptr;
#define SSG_LAUNCH(F, CM, RM)                                                                       \
  hipLaunchKernelGGL((k_ecc_pass<MOTION, F, CM, RM>), dim3(std::min(grid, pass_resident_grid<MOTION, F, CM, RM>())), \
                     dim3(ECC_THREADS), 0, s, a)
  if (full) {
    if (cm && rm) SSG_LAUNCH(true, true, true);
    else if (cm) SSG_LAUNCH(true, true, false);
    else if (rm) SSG_LAUNCH(true, false, true);
    else SSG_LAUNCH(true, false, false);
  } else {
    if (cm && rm) SSG_LAUNCH(false, true, true);
    else if (cm) SSG_LAUNCH(false, true, false);
    else if (rm) SSG_LAUNCH(false, false, true);
    else SSG_LAUNCH(false, false, false);
  }
#undef SSG_LAUNCH
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  return hipSuccess;
}

template <int MOTION>
static hipError_t launch_solve_m(const EccPassArgs& a, bool full, hipStream_t s) {
  if (full) hipLaunchKernelGGL((k_ecc_solve<MOTION, true>), dim3(a.nframes), dim3(256), 0, s, a);
  else hipLaunchKernelGGL((k_ecc_solve<MOTION, false>), dim3(a.nframes), dim3(256), 0, s, a);
  return hipGetLastError();
}

hipError_t launch_ecc_solve(int motion, const EccPassArgs& a, hipStream_t s) {
  switch (motion) {
    case MOTION_TRANSLATION: return launch_solve_m<MOTION_TRANSLATION>(a, false, s);
    case MOTION_EUCLIDEAN: return launch_solve_m<MOTION_EUCLIDEAN>(a, true, s);
    case MOTION_SCALED_EUCLIDEAN: return launch_solve_m<MOTION_SCALED_EUCLIDEAN>(a, true, s);
    case MOTION_AFFINE: return launch_solve_m<MOTION_AFFINE>(a, true, s);
  }
  return hipErrorInvalidValue;
}

hipError_t launch_ecc_pass(int motion, const EccPassArgs& a, int grid, hipStream_t s) {
  switch (motion) {
    case MOTION_TRANSLATION: return launch_pass_m<MOTION_TRANSLATION>(a, false, grid, s);
    case MOTION_EUCLIDEAN: return launch_pass_m<MOTION_EUCLIDEAN>(a, true, grid, s);
    case MOTION_SCALED_EUCLIDEAN: return launch_pass_m<MOTION_SCALED_EUCLIDEAN>(a, true, grid, s);
    case MOTION_AFFINE: return launch_pass_m<MOTION_AFFINE>(a, true, grid, s);
  }
  return hipErrorInvalidValue;
}

// ------------------------------------------------------------------------------------------------
// Level start (c_ecch::align, ecc2.cc:1133-1176): initial parameters at the
// coarsest level (scaled by the width ratio), rescale between levels unless
// the finer level's predecessor failed, reset iteration state, build the
// active list.  One block.
__global__ void __launch_bounds__(1024) k_level_begin(LevelBeginArgs a, int use_translation) {
  const int mmot = a.pass == 0 ? MOTION_TRANSLATION : a.main_motion;
  for (int f = threadIdx.x; f < a.nframes; f += blockDim.x) {
    FrameRes& R = a.res[f];
    FrameXf& X = a.pass == 0 ? a.xf_t[f] : a.xf_m[f];
    if (R.rejected) {
      X.flags = 0;
      continue;
    }
    if (a.level == a.nlevels - 1) {
      for (int i = 0; i < 6; ++i) X.p[i] = 0.f;
      X.C[0] = X.C[1] = 0.f;
      if (mmot == MOTION_AFFINE) {
        X.p[0] = 1.f;
        X.p[4] = 1.f;
      } else if (mmot == MOTION_SCALED_EUCLIDEAN) {
        X.p[3] = 1.f;
      }
      if (a.pass == 1 && use_translation) {
        // set_translation from the translation-first estimate (c_frame_registration.cc:850)
        const float t0 = a.xf_t[f].p[0], t1 = a.xf_t[f].p[1];
        if (mmot == MOTION_AFFINE) {
          X.p[2] = t0;
          X.p[5] = t1;
        } else {
          X.p[0] = t0;
          X.p[1] = t1;
        }
      }
      if (a.nlevels > 1) scale_xf(mmot, X, a.ratio);
      for (int l = 0; l < MAX_LEVELS; ++l) {
        R.level_iters[a.pass][l] = 0;
        R.level_failed[a.pass][l] = 0;
      }
    } else if (!R.level_failed[a.pass][a.level + 1]) {
      scale_xf(mmot, X, a.ratio);
    }
    X.iters = 0;
    X.flags = XF_ACTIVE;
    X.cg = a.cf;
    X.eps = FLT_MAX;
    update_map(mmot, X);
    if (a.lm) {
      // the aligner's local state at the start of its align() (ecc2.cc:1546-1560, 1731-1747, 1980-2000)
      FrameLm& S = a.lm[f];
      S.lambda = a.method == ECC_LM ? 0.1 : 0.001;
      S.err = 0;
      S.rmsold = (double)FLT_MAX;
      S.phase = 0;
      S.iteration = a.method == ECC_IC ? 1 : 0;  // the plain loop counts its first check before the first pass
      S.dpe = 0.f;
      // J of the inverse-compositional aligners at the parameters the level starts from
      S.jc[0] = X.mc.c[0]; S.jc[1] = X.mc.c[1]; S.jc[2] = X.mc.c[2];
      S.jc[3] = X.C[0]; S.jc[4] = X.C[1];
    }
  }
  __syncthreads();
  if (a.lm && a.jrec && mmot != MOTION_TRANSLATION && mmot != MOTION_AFFINE) {
    // J kept across frames: the first frame aligned after the reference was set fixes it
    JRec& J = a.jrec[a.level];
    if (threadIdx.x == 0 && !J.valid) {
      for (int f = 0; f < a.nframes; ++f) {
        if (a.res[f].rejected) continue;
        for (int i = 0; i < 5; ++i) J.c[i] = a.lm[f].jc[i];
        J.valid = 1;
        break;
      }
    }
    __syncthreads();
    for (int f = threadIdx.x; f < a.nframes; f += blockDim.x)
      for (int i = 0; i < 5; ++i) a.lm[f].jc[i] = J.c[i];
    __syncthreads();
  }
  compact_active(a.pass == 0 ? a.xf_t : a.xf_m, a.nframes, a.active_out, a.active_count_out);
}

// The accumulation inputs of a registered batch, formed on the device: the main pass's parameters
// scaled back by 1/ecc.scale when the ECC images were halved, turned into map coefficients with
// the centre at the origin (the host's map_coef of the reported parameters), and the frames that
// are added -- not rejected by a rho gate, with a positive weight (c_frame_accumulation.cc:112).
__global__ void __launch_bounds__(256) k_xf_to_acc(const FrameXf* __restrict__ xf, const FrameRes* __restrict__ res,
                                                   const float* __restrict__ w, int n, int motion, double inv_scale,
                                                   MapCoef* __restrict__ maps, int* __restrict__ valid) {
  const int f = blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= n) return;
  FrameXf X = xf[f];
  if (inv_scale != 1.0) scale_xf(motion, X, inv_scale);
  X.C[0] = X.C[1] = 0.f;
  update_map(motion, X);
  maps[f] = X.mc;
  valid[f] = res[f].rejected == 0 && w[f] > 0.f;
}

hipError_t launch_xf_to_acc(const FrameXf* xf, const FrameRes* res, const float* w, int n, int motion,
                            double inv_scale, MapCoef* maps, int* valid, hipStream_t s) {
  hipLaunchKernelGGL(k_xf_to_acc, dim3((n + 255) / 256), dim3(256), 0, s, xf, res, w, n, motion, inv_scale, maps,
                     valid);
  return hipGetLastError();
}

hipError_t launch_level_begin(const LevelBeginArgs& a, hipStream_t s) {
  const int use_t = a.pass == 1 && a.xf_t != nullptr;
  LevelBeginArgs b = a;
  if (!b.xf_t) b.xf_t = b.xf_m;
  hipLaunchKernelGGL(k_level_begin, dim3(1), dim3(1024), 0, s, b, use_t);
  return hipGetLastError();
}

#include "ssg_ecc_lm.inc"

// ------------------------------------------------------------------------------------------------
// compute_correlation (ecc2.cc:65-137) for a batch: remap(cur, INTER_LINEAR,
// BORDER_CONSTANT 0), mask = remap(mask) >= 254 & reference mask, Pearson
// correlation over the mask; then the rho gate of register_frame.
template <int MOTION, bool HAS_CMASK, bool HAS_RMASK>
__global__ void __launch_bounds__(ECC_THREADS) k_corr_pass(CorrPassArgs a) {
  __shared__ double dred[ECC_WAVES][8];
  __shared__ float fred[ECC_WAVES][4];
  __shared__ float4 wtab[1024];  // bilinear weights of the 32 x 32 sub-pixel positions
  const int tid = threadIdx.x, lane = tid & (WAVE - 1), wv = tid / WAVE;
  fill_bilinear_table(wtab, tid, ECC_THREADS);
  __syncthreads();
  const int items = a.nlist * a.ntiles;
  const int G = gridDim.x, b = blockIdx.x;
  const int vb = (G % 8 == 0) ? (b % 8) * (G / 8) + b / 8 : b;
  for (int item = vb; item < items; item += G) {
    const int fi = item / a.ntiles;
    const int tile = item - fi * a.ntiles;
    const int f = __builtin_amdgcn_readfirstlane(a.frames[fi]);
    if (a.res[f].rejected) continue;
    const int tx = tile % a.tiles_x, ty = tile / a.tiles_x;
    const int x0 = tx * CORR_TW, y0 = ty * CORR_TH;
    MapCoef mc;
#pragma unroll
    for (int i = 0; i < 8; ++i) mc.c[i] = a.xf[f].mc.c[i];
    const float* __restrict__ cur = a.cur + f * a.cur_fstride;
    const uint8_t* __restrict__ cmask = HAS_CMASK ? a.cmask + f * a.cmask_fstride : nullptr;
    const int x = x0 + lane;
    float n = 0, F1 = 0, G1 = 0, F2 = 0, G2 = 0, FG = 0;
    // extremes of f' and g' over the mask: a constant image is recognised exactly
    float fmn = INFINITY, fmx = -INFINITY, gmn = INFINITY, gmx = -INFINITY;
    // Interior strips (the warped corners of the wave's clipped 16-row strip keep every tap
    // inside the current image, no current mask): the mask test always passes and the sample
    // needs no border handling, so four rows' loads go out together; the sums keep the row order.
    bool fast = false;
    const int ys0 = y0 + wv * ECC_RPT;
    if (!HAS_CMASK && ys0 < a.h) {
      const int ys1 = min(ys0 + ECC_RPT - 1, a.h - 1);
      const int cx[4] = {x0, min(x0 + CORR_TW - 1, a.w - 1), x0, min(x0 + CORR_TW - 1, a.w - 1)};
      const int cy[4] = {ys0, ys0, ys1, ys1};
      float mnx = 1e30f, mxx = -1e30f, mny = 1e30f, mxy = -1e30f;
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        float sx, sy;
        warp_xy<MOTION>(mc, cx[i], cy[i], sx, sy);
        mnx = fminf(mnx, sx); mxx = fmaxf(mxx, sx); mny = fminf(mny, sy); mxy = fmaxf(mxy, sy);
        if (sx != sx || sy != sy) mnx = -1e30f;
      }
      fast = mnx >= 0.25f && mny >= 0.25f && mxx <= (float)a.cw - 1.25f && mxy <= (float)a.ch - 1.25f;
    }
    if (fast) {
      // every position lies in [0.25, size - 1.25]: quantised without the clamp, the taps read with
      // buffer loads at 32-bit offsets, the weights from the table
      const int xc = min(x, a.w - 1);
      const uint32_t cpb = (uint32_t)a.cur_pitch * 4u, rpb = (uint32_t)a.ref_pitch * 4u;
      const rsrc_t rc = make_rsrc(cur, (uint32_t)a.ch * cpb);
      const rsrc_t rr = make_rsrc(a.ref, (uint32_t)a.h * rpb);
      // a translation: the x position of a column is the same in every row and the y position of
      // a row the same in every column -- the column's tap offset and fraction once per lane, the
      // row's (formed by lane j for row j of the strip) read with readlane; the same samples
      int txo = 0, tfx = 0, tro = 0, tqy = 0;
      if constexpr (MOTION == MOTION_TRANSLATION) {
        const int X = quant32_inrange((float)xc + mc.c[0]);
        txo = (X >> 5) * 4;
        tfx = X & 31;
        const int Y = quant32_inrange((float)min(y0 + wv * ECC_RPT + (lane & (ECC_RPT - 1)), a.h - 1) + mc.c[1]);
        tro = (Y >> 5) * (int)cpb;
        tqy = (Y & 31) << 5;
      }
#pragma nounroll
      for (int dy0 = 0; dy0 < ECC_RPT; dy0 += 4) {
        float t[4][4], fv[4];
        int q[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          const int yc = min(y0 + wv * ECC_RPT + dy0 + j, a.h - 1);
          if constexpr (MOTION == MOTION_TRANSLATION) {
            const uint32_t r = (uint32_t)__builtin_amdgcn_readlane(tro, dy0 + j);
            t[j][0] = bload_f32(rc, (uint32_t)txo, r);
            t[j][1] = bload_f32(rc, (uint32_t)txo + 4u, r);
            t[j][2] = bload_f32(rc, (uint32_t)txo, r + cpb);
            t[j][3] = bload_f32(rc, (uint32_t)txo + 4u, r + cpb);
            q[j] = __builtin_amdgcn_readlane(tqy, dy0 + j) | tfx;
          } else {
            float sx, sy;
            warp_xy<MOTION>(mc, xc, yc, sx, sy);
            const int X = quant32_inrange(sx), Y = quant32_inrange(sy);
            const uint32_t o = (uint32_t)(Y >> 5) * cpb + (uint32_t)(X >> 5) * 4u;
            t[j][0] = bload_f32(rc, o, 0u);
            t[j][1] = bload_f32(rc, o + 4u, 0u);
            t[j][2] = bload_f32(rc, o, cpb);
            t[j][3] = bload_f32(rc, o + 4u, cpb);
            q[j] = ((Y & 31) << 5) | (X & 31);
          }
          fv[j] = bload_f32(rr, (uint32_t)xc * 4u, (uint32_t)yc * rpb);
        }
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          const int y = y0 + wv * ECC_RPT + dy0 + j;
          if (x >= a.w || y >= a.h) continue;
          if (HAS_RMASK && a.rmask[(size_t)y * a.rmask_pitch + x] == 0) continue;
          const float4 w = wtab[q[j]];
          const float gv = bilinear_sum(t[j][0], t[j][1], t[j][2], t[j][3], w.x, w.y, w.z, w.w);
          const float fp = fv[j] - a.cf, gp = gv - a.cf;
          n += 1.f;
          F1 += fp; G1 += gp;
          F2 = fmaf(fp, fp, F2); G2 = fmaf(gp, gp, G2); FG = fmaf(fp, gp, FG);
          fmn = fminf(fmn, fp); fmx = fmaxf(fmx, fp); gmn = fminf(gmn, gp); gmx = fmaxf(gmx, gp);
        }
      }
    }
    for (int dy = 0; dy < (fast ? 0 : ECC_RPT); ++dy) {
      const int y = y0 + wv * ECC_RPT + dy;
      if (x >= a.w || y >= a.h) continue;
      float sx, sy;
      warp_xy<MOTION>(mc, x, y, sx, sy);
      const int X = quant32(sx), Y = quant32(sy);
      bool m = remap_mask_u8(cmask, a.cmask_pitch, a.cw, a.ch, X, Y) >= 254;
      if (HAS_RMASK) m = m && a.rmask[(size_t)y * a.rmask_pitch + x] != 0;
      if (!m) continue;
      float gv;
      sample_linear(cur, a.cur_pitch, a.cw, a.ch, X, Y, BORDER_CONSTANT, 0.f, gv);
      const float fp = a.ref[(size_t)y * a.ref_pitch + x] - a.cf, gp = gv - a.cf;
      n += 1.f;
      F1 += fp; G1 += gp;
      F2 = fmaf(fp, fp, F2); G2 = fmaf(gp, gp, G2); FG = fmaf(fp, gp, FG);
      fmn = fminf(fmn, fp); fmx = fmaxf(fmx, fp); gmn = fminf(gmn, gp); gmx = fmaxf(gmx, gp);
    }
    {
      // wave totals by recursive halving in the butterfly's mask order (the sums of wave_sum, bit
      // for bit); the maxima as minima of the negated values (exact), all four in one halving
      const double v[6] = {n, F1, G1, F2, G2, FG};
      const double tot = ReduceScatterDesc<6>::run(v, lane, [](double p, double q) { return p + q; });
      const int s6 = ReduceScatterDesc<6>::slot_of(lane);
      if (s6 >= 0) dred[wv][s6] = tot;
      const float e[4] = {fmn, -fmx, gmn, -gmx};
      const float m = ReduceScatterDesc<4>::run(e, lane, [](float p, float q) { return fminf(p, q); });
      const int s4 = ReduceScatterDesc<4>::slot_of(lane);
      if (s4 >= 0) fred[wv][s4] = (s4 & 1) ? -m : m;
    }
    __syncthreads();
    double* part = reinterpret_cast<double*>(a.partials) + ((size_t)f * a.ntiles + tile) * NMOM_CORR;
    if (tid < 6) part[tid] = ((dred[0][tid] + dred[1][tid]) + dred[2][tid]) + dred[3][tid];
    if (tid >= 6 && tid < NMOM_CORR) {
      const int q = tid - 6;
      float e = fred[0][q];
      for (int w = 1; w < ECC_WAVES; ++w) e = (q & 1) ? fmaxf(e, fred[w][q]) : fminf(e, fred[w][q]);
      part[tid] = (double)e;
    }
    __syncthreads();  // dred is rewritten by the next item
  }
}

// One block per listed frame: tile sums in tile order, rho, and the gate.
__global__ void __launch_bounds__(256) k_corr_finalize(CorrPassArgs a) {
  __shared__ double red[ECC_WAVES][8];
  __shared__ double ered[ECC_WAVES][4];
  const int f = a.frames[blockIdx.x];
  if (a.res[f].rejected) return;
  const int tid = threadIdx.x, lane = tid & (WAVE - 1), wv = tid / WAVE;
  const double* base = reinterpret_cast<const double*>(a.partials) + (size_t)f * a.ntiles * NMOM_CORR;
  double S[6] = {0, 0, 0, 0, 0, 0};
  double E[4] = {INFINITY, -INFINITY, INFINITY, -INFINITY};  // min f', max f', min g', max g'
  for (int t = tid; t < a.ntiles; t += blockDim.x) {
    for (int i = 0; i < 6; ++i) S[i] += base[(size_t)t * NMOM_CORR + i];
    for (int i = 0; i < 4; ++i) {
      const double e = base[(size_t)t * NMOM_CORR + 6 + i];
      E[i] = (i & 1) ? fmax(E[i], e) : fmin(E[i], e);
    }
  }
  for (int i = 0; i < 6; ++i) S[i] = wave_sum(S[i]);
#pragma unroll
  for (int o = 32; o > 0; o >>= 1)
    for (int i = 0; i < 4; ++i) {
      const double e = __shfl_xor(E[i], o);
      E[i] = (i & 1) ? fmax(E[i], e) : fmin(E[i], e);
    }
  if (lane == 0) {
    for (int i = 0; i < 6; ++i) red[wv][i] = S[i];
    for (int i = 0; i < 4; ++i) ered[wv][i] = E[i];
  }
  __syncthreads();
  if (tid == 0) {
    for (int i = 0; i < 6; ++i) S[i] = ((red[0][i] + red[1][i]) + red[2][i]) + red[3][i];
    for (int w = 1; w < ECC_WAVES; ++w)
      for (int i = 0; i < 4; ++i) ered[0][i] = (i & 1) ? fmax(ered[0][i], ered[w][i]) : fmin(ered[0][i], ered[w][i]);
    // the image or the reference constant over the mask: the reference's centred covariance
    // is exactly 0 (its float mean is that constant) and its rho 0 / (s1 s2)
    const bool flat = ered[0][0] == ered[0][1] || ered[0][2] == ered[0][3];
    const double inv = S[0] > 0 ? 1.0 / S[0] : 0.0;
    const double mf = S[1] * inv, mg = S[2] * inv;
    const double vf = S[3] * inv - mf * mf, vg = S[4] * inv - mg * mg;
    const double sf = sqrt(vf > 0 ? vf : 0.0), sg = sqrt(vg > 0 ? vg : 0.0);
    const double cov = S[5] * inv - mf * mg;
    // a constant image over the mask (or an empty mask): the raw moments here can leave a
    // rounding residue in cov and the variance, so the reference's value is set explicitly --
    // 0 / 0 = NaN, its result whenever its double sums of the constant are exact (0 and
    // short-mantissa values; other constants make its own variance a rounding residue)
    const double rho = (flat || S[0] == 0) ? __builtin_nan("") : cov / (sf * sg);
    FrameRes& R = a.res[f];
    R.rho[a.pass] = rho;
    // `rho < thr` as in the reference: a NaN rho passes the gate
    const double thr = a.pass == 0 ? 0.75 * a.min_rho : a.min_rho;
    if (rho < thr) R.rejected = a.pass == 0 ? 3 : 4;
  }
}

template <int MOTION>
static hipError_t launch_corr_m(const CorrPassArgs& a, int grid, hipStream_t s) {
  const bool cm = a.cmask != nullptr, rm = a.rmask != nullptr;
#define SSG_LAUNCH(CM, RM) hipLaunchKernelGGL((k_corr_pass<MOTION, CM, RM>), dim3(grid), dim3(ECC_THREADS), 0, s, a)
  if (cm && rm) SSG_LAUNCH(true, true);
  else if (cm) SSG_LAUNCH(true, false);
  else if (rm) SSG_LAUNCH(false, true);
  else SSG_LAUNCH(false, false);
#undef SSG_LAUNCH
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_corr_finalize, dim3(a.nlist), dim3(256), 0, s, a);
  return hipGetLastError();
}

hipError_t launch_corr_pass(int motion, const CorrPassArgs& a, int grid, hipStream_t s) {
  switch (motion) {
    case MOTION_TRANSLATION: return launch_corr_m<MOTION_TRANSLATION>(a, grid, s);
    case MOTION_EUCLIDEAN: return launch_corr_m<MOTION_EUCLIDEAN>(a, grid, s);
    case MOTION_SCALED_EUCLIDEAN: return launch_corr_m<MOTION_SCALED_EUCLIDEAN>(a, grid, s);
    case MOTION_AFFINE: return launch_corr_m<MOTION_AFFINE>(a, grid, s);
  }
  return hipErrorInvalidValue;
}

}  // namespace ssg

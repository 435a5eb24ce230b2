/*
 * ssg_oracle.c -- CPU oracle (test infrastructure only; see ssg_oracle.h).
 *
 * Reference citations are relative to the SerStacker source tree.
 * OpenCV semantics restated here follow OpenCV 4.5.x (the version the
 * reference's Ubuntu 22.04 build uses, dists/setup_build_deps_ubuntu_22.sh:24).
 */
#include "ssg_oracle.h"

#include <float.h>
#include <limits.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <omp.h>

#if SSG_DOUBLE
#define RFMA fma
#define RSQRT sqrt
#define RLRINT lrint
#define RSIN sin
#define RCOS cos
#define RATAN2 atan2
#else
#define RFMA fmaf
#define RSQRT sqrtf
#define RLRINT lrintf
#define RSIN sinf
#define RCOS cosf
#define RATAN2 atan2f
#endif

typedef ora_real real;

static int g_threads = 0;

int ora_real_size(void) { return (int)sizeof(real); }

void ora_set_threads(int n) {
  g_threads = n;
  if (n > 0) omp_set_num_threads(n);
}

static void *xcalloc(size_t n, size_t sz) {
  void *p = calloc(n ? n : 1, sz);
  if (!p) abort();
  return p;
}

/* ------------------------------------------------------------------------ */
/* cv::borderInterpolate                                                     */
/* ------------------------------------------------------------------------ */
int ora_border_interpolate(int p, int len, int border) {
  if ((unsigned)p < (unsigned)len) return p;
  if (border == ORA_BORDER_REPLICATE) return p < 0 ? 0 : len - 1;
  if (border == ORA_BORDER_REFLECT || border == ORA_BORDER_REFLECT101) {
    int delta = border == ORA_BORDER_REFLECT101;
    if (len == 1) return 0;
    do {
      if (p < 0) p = -p - 1 + delta;
      else p = len - 1 - (p - len) - delta;
    } while ((unsigned)p >= (unsigned)len);
    return p;
  }
  if (border == ORA_BORDER_WRAP) {
    if (p < 0) p -= ((p - len + 1) / len) * len;
    if (p >= len) p %= len;
    return p;
  }
  return -1; /* CONSTANT / TRANSPARENT */
}

/* ------------------------------------------------------------------------ */
/* cv::getGaussianKernel(n, sigma, CV_64F) followed by the conversion to the  */
/* filter's working type (createSeparableLinearFilter converts kernels to the */
/* float buffer depth).                                                      */
/* ------------------------------------------------------------------------ */
void ora_gaussian_kernel(int n, double sigma, real *out) {
  static const float small_tab[4][7] = {
      {1.f},
      {0.25f, 0.5f, 0.25f},
      {0.0625f, 0.25f, 0.375f, 0.25f, 0.0625f},
      {0.03125f, 0.109375f, 0.21875f, 0.28125f, 0.21875f, 0.109375f, 0.03125f}};
  const float *fixed = (n % 2 == 1 && n <= 7 && sigma <= 0) ? small_tab[n >> 1] : NULL;
  double sigmaX = sigma > 0 ? sigma : ((n - 1) * 0.5 - 1) * 0.3 + 0.8;
  double scale2X = -0.5 / (sigmaX * sigmaX);
  double *cd = (double *)xcalloc(n, sizeof(double));
  double sum = 0;
  for (int i = 0; i < n; i++) {
    double x = i - (n - 1) * 0.5;
    double t = fixed ? (double)fixed[i] : exp(scale2X * x * x);
    cd[i] = t;
    sum += cd[i];
  }
  sum = 1. / sum;
  for (int i = 0; i < n; i++) out[i] = (real)(cd[i] * sum);
  free(cd);
}

/* cv::getKernelType: 1 = symmetric, 2 = antisymmetric, 0 = neither */
static int kernel_symmetry(const real *k, int n) {
  int sym = 1, asym = 1;
  for (int i = 0; i < n; i++) {
    double a = k[i], b = k[n - 1 - i];
    if (a != b) sym = 0;
    if (a != -b) asym = 0;
  }
  return sym ? 1 : (asym ? 2 : 0);
}

/* ------------------------------------------------------------------------ */
/* cv::sepFilter2D: row filter then column filter, correlation, centred      */
/* anchor.  Canonical fma order (shared with the HIP kernels):               */
/*   row, symmetric n<=5  : s = S0*k0; s = fma(S-1+S1, k1, s); ...           */
/*   row, antisym  n<=5   : s = (S1-S-1)*k1; s = fma(S2-S-2, k2, s)          */
/*   row, other           : s = 0; s = fma(S[i], k[i], s) for i = 0..n-1      */
/*   column, symmetric    : s = fma(S0, k0, delta); s = fma(S-k+Sk, kk, s)   */
/*   column, antisym      : s = delta; s = fma(Sk-S-k, kk, s)                */
/*   column, other        : s = delta; s = fma(S[i], k[i], s)                */
/* These follow OpenCV's SymmRowSmallFilter / RowFilter / SymmColumnFilter   */
/* split (filter.simd.hpp) with its multiply-adds fused.                     */
/* ------------------------------------------------------------------------ */
static real row_filter_at(const real *S, int w, int x, const real *k, int n, int ktype,
                          int border, real cval) {
#define SRC(o) ({ int _p = ora_border_interpolate(x + (o), w, border); _p < 0 ? cval : S[_p]; })
  const int a = n / 2;
  real s;
  if (ktype == 1 && n <= 5) {
    s = SRC(0) * k[a];
    if (n >= 3) s = RFMA(SRC(-1) + SRC(1), k[a + 1], s);
    if (n >= 5) s = RFMA(SRC(-2) + SRC(2), k[a + 2], s);
  } else if (ktype == 2 && n <= 5 && n >= 3) {
    s = (SRC(1) - SRC(-1)) * k[a + 1];
    if (n >= 5) s = RFMA(SRC(2) - SRC(-2), k[a + 2], s);
  } else {
    s = 0;
    for (int i = 0; i < n; i++) s = RFMA(SRC(i - a), k[i], s);
  }
  return s;
#undef SRC
}

static real col_filter_at(real *const *R, const real *k, int n, int ktype, real delta) {
  /* R[i] points to the row-filtered value for row offset (i - n/2) */
  const int a = n / 2;
  real s;
  if (ktype == 1) {
    s = RFMA(*R[a], k[a], delta);
    for (int i = 1; i <= a; i++) s = RFMA(*R[a - i] + *R[a + i], k[a + i], s);
  } else if (ktype == 2) {
    s = delta;
    for (int i = 1; i <= a; i++) s = RFMA(*R[a + i] - *R[a - i], k[a + i], s);
  } else {
    s = delta;
    for (int i = 0; i < n; i++) s = RFMA(*R[i], k[i], s);
  }
  return s;
}

void ora_sep_filter(const real *src, int w, int h, real *dst, const real *kx, int nx,
                    const real *ky, int ny, int border, real delta) {
  const int ay = ny / 2;
  const int kxt = kernel_symmetry(kx, nx), kyt = kernel_symmetry(ky, ny);
  const int nrows = h + 2 * ay;
  real *rows = (real *)xcalloc((size_t)nrows * w, sizeof(real));
  real *zero_row = (real *)xcalloc((size_t)w, sizeof(real));
#pragma omp parallel for schedule(static)
  for (int r = 0; r < nrows; r++) {
    int sr = ora_border_interpolate(r - ay, h, border);
    const real *S = sr < 0 ? zero_row : src + (size_t)sr * w;
    real *out = rows + (size_t)r * w;
    for (int x = 0; x < w; x++) out[x] = row_filter_at(S, w, x, kx, nx, kxt, border, 0);
  }
#pragma omp parallel for schedule(static)
  for (int y = 0; y < h; y++) {
    real *R[64];
    for (int x = 0; x < w; x++) {
      for (int i = 0; i < ny; i++) R[i] = rows + (size_t)(y + i) * w + x;
      dst[(size_t)y * w + x] = col_filter_at(R, ky, ny, kyt, delta);
    }
  }
  free(rows);
  free(zero_row);
}

/* ------------------------------------------------------------------------ */
/* cv::pyrDown (float), [1 4 6 4 1]/16 separable at even source pixels.      */
/* Canonical: r = fma(s[-1]+s[1], 4, fma(s[0], 6, s[-2]+s[2]));              */
/*            d = fma(r1+r3, 4, fma(r2, 6, r0+r4)) * (1/256)                 */
/* ------------------------------------------------------------------------ */
void ora_pyr_down(const real *src, int sw, int sh, real *dst, int dw, int dh, int border) {
  real *rows = (real *)xcalloc((size_t)sh * dw, sizeof(real));
#pragma omp parallel for schedule(static)
  for (int r = 0; r < sh; r++) {
    const real *S = src + (size_t)r * sw;
    for (int x = 0; x < dw; x++) {
      int c = 2 * x;
      real sm2 = S[ora_border_interpolate(c - 2, sw, border)];
      real sm1 = S[ora_border_interpolate(c - 1, sw, border)];
      real s0 = S[ora_border_interpolate(c, sw, border)];
      real sp1 = S[ora_border_interpolate(c + 1, sw, border)];
      real sp2 = S[ora_border_interpolate(c + 2, sw, border)];
      rows[(size_t)r * dw + x] = RFMA(sm1 + sp1, (real)4, RFMA(s0, (real)6, sm2 + sp2));
    }
  }
  const real scale = (real)(1. / 256);
#pragma omp parallel for schedule(static)
  for (int y = 0; y < dh; y++) {
    int c = 2 * y;
    const real *r0 = rows + (size_t)ora_border_interpolate(c - 2, sh, border) * dw;
    const real *r1 = rows + (size_t)ora_border_interpolate(c - 1, sh, border) * dw;
    const real *r2 = rows + (size_t)ora_border_interpolate(c, sh, border) * dw;
    const real *r3 = rows + (size_t)ora_border_interpolate(c + 1, sh, border) * dw;
    const real *r4 = rows + (size_t)ora_border_interpolate(c + 2, sh, border) * dw;
    for (int x = 0; x < dw; x++)
      dst[(size_t)y * dw + x] =
          RFMA(r1[x] + r3[x], (real)4, RFMA(r2[x], (real)6, r0[x] + r4[x])) * scale;
  }
  free(rows);
}

/* cv::pyrDown on CV_8U: integer taps, (sum + 128) >> 8, BORDER_REFLECT_101 */
void ora_pyr_down_u8(const uint8_t *src, int sw, int sh, uint8_t *dst, int dw, int dh) {
  int *rows = (int *)xcalloc((size_t)sh * dw, sizeof(int));
  const int b = ORA_BORDER_REFLECT101;
  for (int r = 0; r < sh; r++) {
    const uint8_t *S = src + (size_t)r * sw;
    for (int x = 0; x < dw; x++) {
      int c = 2 * x;
      rows[(size_t)r * dw + x] = S[ora_border_interpolate(c, sw, b)] * 6 +
                                 (S[ora_border_interpolate(c - 1, sw, b)] +
                                  S[ora_border_interpolate(c + 1, sw, b)]) * 4 +
                                 S[ora_border_interpolate(c - 2, sw, b)] +
                                 S[ora_border_interpolate(c + 2, sw, b)];
    }
  }
  for (int y = 0; y < dh; y++) {
    int c = 2 * y;
    const int *r0 = rows + (size_t)ora_border_interpolate(c - 2, sh, b) * dw;
    const int *r1 = rows + (size_t)ora_border_interpolate(c - 1, sh, b) * dw;
    const int *r2 = rows + (size_t)ora_border_interpolate(c, sh, b) * dw;
    const int *r3 = rows + (size_t)ora_border_interpolate(c + 1, sh, b) * dw;
    const int *r4 = rows + (size_t)ora_border_interpolate(c + 2, sh, b) * dw;
    for (int x = 0; x < dw; x++) {
      int v = (r0[x] + r4[x] + (r1[x] + r3[x]) * 4 + r2[x] * 6 + 128) >> 8;
      dst[(size_t)y * dw + x] = (uint8_t)(v < 0 ? 0 : (v > 255 ? 255 : v));
    }
  }
  free(rows);
}

/* ------------------------------------------------------------------------ */
/* cv::pyrUp (float): even taps [1 6 1]/8, odd taps [4 4]/8 per axis.         */
/* Left/top border reflect-101, right/bottom border duplicates the edge,     */
/* odd destination sizes repeat the last computed column/row.                */
/* Canonical: even = fma(s0, 6, s-1 + s1), odd = (s0 + s1) * 4, then /64.    */
/* ------------------------------------------------------------------------ */
void ora_pyr_up(const real *src, int sw, int sh, real *dst, int dw, int dh) {
  const int bw = 2 * sw + 2;
  real *rows = (real *)xcalloc((size_t)(sh + 2) * bw, sizeof(real));
  /* horizontally upsampled rows for source rows -1 .. sh */
  for (int rr = -1; rr <= sh; rr++) {
    int r = rr < 0 ? (sh > 1 ? 1 : 0) : (rr >= sh ? sh - 1 : rr);
    const real *S = src + (size_t)r * sw;
    real *out = rows + (size_t)(rr + 1) * bw;
    if (sw == 1) {
      out[0] = out[1] = S[0] * (real)8;
      continue;
    }
    out[0] = RFMA(S[0], (real)6, S[1] + S[1]);
    out[1] = (S[0] + S[1]) * (real)4;
    for (int x = 1; x < sw - 1; x++) {
      out[2 * x] = RFMA(S[x], (real)6, S[x - 1] + S[x + 1]);
      out[2 * x + 1] = (S[x] + S[x + 1]) * (real)4;
    }
    out[2 * (sw - 1)] = RFMA(S[sw - 1], (real)6, S[sw - 2] + S[sw - 1]);
    out[2 * (sw - 1) + 1] = S[sw - 1] * (real)8;
    out[2 * sw] = out[2 * sw - 1];
  }
  const real scale = (real)(1. / 64);
  for (int y = 0; y < sh; y++) {
    const real *r0 = rows + (size_t)(y)*bw;     /* source row y-1 */
    const real *r1 = rows + (size_t)(y + 1) * bw; /* source row y */
    const real *r2 = rows + (size_t)(y + 2) * bw; /* source row y+1 */
    int y0 = 2 * y, y1 = 2 * y + 1 < dh - 1 ? 2 * y + 1 : dh - 1;
    for (int x = 0; x < dw; x++) {
      real t1 = (r1[x] + r2[x]) * (real)4 * scale;
      real t0 = RFMA(r1[x], (real)6, r0[x] + r2[x]) * scale;
      if (y1 < dh) dst[(size_t)y1 * dw + x] = t1;
      if (y0 < dh) dst[(size_t)y0 * dw + x] = t0;
    }
  }
  if (dh > 2 * sh) memcpy(dst + (size_t)(2 * sh) * dw, dst + (size_t)(2 * sh - 2) * dw, sizeof(real) * dw);
  free(rows);
}

/* cv::resize INTER_NEAREST: sx = min(floor(x * sw/dw), sw-1) */
void ora_resize_nearest_u8(const uint8_t *src, int sw, int sh, uint8_t *dst, int dw, int dh) {
  double ifx = (double)sw / dw, ify = (double)sh / dh;
  for (int y = 0; y < dh; y++) {
    int sy = (int)floor(y * ify);
    if (sy > sh - 1) sy = sh - 1;
    for (int x = 0; x < dw; x++) {
      int sx = (int)floor(x * ifx);
      if (sx > sw - 1) sx = sw - 1;
      dst[(size_t)y * dw + x] = src[(size_t)sy * sw + sx];
    }
  }
}

/* ------------------------------------------------------------------------ */
/* cv::remap with a float map (imgwarp.cpp RemapInvoker).  Positions are      */
/* quantised to 1/32 px: X = cvRound(sx*32), ix = X>>5, fx = X&31.            */
/* Linear float weights w = {(1-ty)(1-tx), (1-ty)tx, ty(1-tx), ty tx}.        */
/* Canonical sum: v = S00*w0; v = fma(S01,w1,v); v = fma(S10,w2,v);           */
/*                v = fma(S11,w3,v).                                          */
/* Cubic: Keys a=-0.75 table at t=i/32; per-row fma chains added row by row. */
/* ------------------------------------------------------------------------ */
static inline int quant32(real v) {
  real s = v * (real)32;
  if (!(s > (real)-1073741824.0)) return -1073741824;
  if (!(s < (real)1073741824.0)) return 1073741824;
  return (int)RLRINT(s);
}

static inline int sat_short(int v) { return v < -32768 ? -32768 : (v > 32767 ? 32767 : v); }

static void cubic_coeffs(float x, float *c) {
  const float A = -0.75f;
  c[0] = ((A * (x + 1) - 5 * A) * (x + 1) + 8 * A) * (x + 1) - 4 * A;
  c[1] = ((A + 2) * x - (A + 3)) * x * x + 1;
  c[2] = ((A + 2) * (1 - x) - (A + 3)) * (1 - x) * (1 - x) + 1;
  c[3] = 1.f - c[0] - c[1] - c[2];
}

static inline real tap(const real *src, int sw, int sh, int x, int y, int border, real cval,
                       int *missing) {
  int xx = ora_border_interpolate(x, sw, border);
  int yy = ora_border_interpolate(y, sh, border);
  if (xx < 0 || yy < 0) {
    *missing = 1;
    return cval;
  }
  return src[(size_t)yy * sw + xx];
}

void ora_remap(const real *src, int sw, int sh, const real *map, int dw, int dh, real *dst,
               int interp, int border, real cval) {
  float ctab[32][4];
  for (int i = 0; i < 32; i++) cubic_coeffs(i * (1.f / 32), ctab[i]);
#pragma omp parallel for schedule(static)
  for (int y = 0; y < dh; y++) {
    for (int x = 0; x < dw; x++) {
      size_t o = (size_t)y * dw + x;
      real sx = map[2 * o], sy = map[2 * o + 1];
      if (interp == ORA_INTER_NEAREST) {
        int ix = sat_short((int)RLRINT(sx)), iy = sat_short((int)RLRINT(sy));
        int miss = 0;
        real v = tap(src, sw, sh, ix, iy, border, cval, &miss);
        if (miss && border == ORA_BORDER_TRANSPARENT) continue;
        dst[o] = v;
        continue;
      }
      int X = quant32(sx), Y = quant32(sy);
      int ix = sat_short(X >> 5), iy = sat_short(Y >> 5);
      int fx = X & 31, fy = Y & 31;
      if (interp == ORA_INTER_LINEAR) {
        if (border == ORA_BORDER_CONSTANT && (ix >= sw || ix + 1 < 0 || iy >= sh || iy + 1 < 0)) {
          dst[o] = cval;
          continue;
        }
        real tx = (real)(float)(fx * (1.f / 32)), ty = (real)(float)(fy * (1.f / 32));
        real w0 = (1 - ty) * (1 - tx), w1 = (1 - ty) * tx, w2 = ty * (1 - tx), w3 = ty * tx;
        int miss = 0;
        real s00 = tap(src, sw, sh, ix, iy, border, cval, &miss);
        real s01 = tap(src, sw, sh, ix + 1, iy, border, cval, &miss);
        real s10 = tap(src, sw, sh, ix, iy + 1, border, cval, &miss);
        real s11 = tap(src, sw, sh, ix + 1, iy + 1, border, cval, &miss);
        if (miss && border == ORA_BORDER_TRANSPARENT) continue;
        real v = s00 * w0;
        v = RFMA(s01, w1, v);
        v = RFMA(s10, w2, v);
        v = RFMA(s11, w3, v);
        dst[o] = v;
      } else { /* cubic (remapBicubic) */
        int bx = ix - 1, by = iy - 1;
        const float *cx = ctab[fx], *cy = ctab[fy];
        if ((unsigned)bx < (unsigned)(sw > 3 ? sw - 3 : 0) && (unsigned)by < (unsigned)(sh > 3 ? sh - 3 : 0)) {
          /* interior: per-row fma chains added row by row */
          real sum = 0;
          for (int r = 0; r < 4; r++) {
            const real *S = src + (size_t)(by + r) * sw + bx;
            real t = S[0] * (real)(cy[r] * cx[0]);
            for (int c = 1; c < 4; c++) t = RFMA(S[c], (real)(cy[r] * cx[c]), t);
            sum = r == 0 ? t : sum + t;
          }
          dst[o] = sum;
          continue;
        }
        if (border == ORA_BORDER_TRANSPARENT &&
            ((unsigned)(bx + 1) >= (unsigned)sw || (unsigned)(by + 1) >= (unsigned)sh))
          continue;
        if (border == ORA_BORDER_CONSTANT && (bx >= sw || bx + 4 <= 0 || by >= sh || by + 4 <= 0)) {
          dst[o] = cval;
          continue;
        }
        /* border path: sum = cval + sum over in-image taps of (S - cval) * w */
        int xs[4], ys[4];
        for (int i = 0; i < 4; i++) {
          xs[i] = ora_border_interpolate(bx + i, sw, border);
          ys[i] = ora_border_interpolate(by + i, sh, border);
        }
        real sum = cval;
        for (int r = 0; r < 4; r++) {
          if (ys[r] < 0) continue;
          const real *S = src + (size_t)ys[r] * sw;
          for (int c = 0; c < 4; c++)
            if (xs[c] >= 0) sum = RFMA(S[xs[c]] - cval, (real)(cy[r] * cx[c]), sum);
        }
        dst[o] = sum;
      }
    }
  }
}

/* cv::remap on CV_8U, INTER_LINEAR: 15-bit fixed-point weights from
 * BilinearTab_i.  The table entry for fx = fy = 0 is {32767, 0, 0, 1}: the
 * float weight 1.0 saturates to SHRT_MAX and initInterTab2D puts the missing
 * unit on the bottom-right tap.  Result = (sum + (1<<14)) >> 15. */
static inline void bilinear_itab(int fx, int fy, int w[4]) {
  w[0] = (32 - fx) * (32 - fy) * 32;
  w[1] = fx * (32 - fy) * 32;
  w[2] = (32 - fx) * fy * 32;
  w[3] = fx * fy * 32;
  if (fx == 0 && fy == 0) {
    w[0] = 32767;
    w[3] = 1;
  }
}

void ora_remap_u8(const uint8_t *src, int sw, int sh, const real *map, int dw, int dh,
                  uint8_t *dst, int border, uint8_t cval) {
#pragma omp parallel for schedule(static)
  for (int y = 0; y < dh; y++) {
    for (int x = 0; x < dw; x++) {
      size_t o = (size_t)y * dw + x;
      int X = quant32(map[2 * o]), Y = quant32(map[2 * o + 1]);
      int ix = sat_short(X >> 5), iy = sat_short(Y >> 5);
      int fx = X & 31, fy = Y & 31;
      if (border == ORA_BORDER_CONSTANT && (ix >= sw || ix + 1 < 0 || iy >= sh || iy + 1 < 0)) {
        dst[o] = cval;
        continue;
      }
      int w[4];
      bilinear_itab(fx, fy, w);
      int vals[4], k = 0, miss = 0;
      for (int j = 0; j < 2; j++)
        for (int i = 0; i < 2; i++, k++) {
          int xx = ora_border_interpolate(ix + i, sw, border);
          int yy = ora_border_interpolate(iy + j, sh, border);
          if (xx < 0 || yy < 0) {
            vals[k] = cval;
            miss = 1;
          } else
            vals[k] = src[(size_t)yy * sw + xx];
        }
      if (miss && border == ORA_BORDER_TRANSPARENT) continue;
      int s = vals[0] * w[0] + vals[1] * w[1] + vals[2] * w[2] + vals[3] * w[3];
      int v = (s + (1 << 14)) >> 15;
      dst[o] = (uint8_t)(v < 0 ? 0 : (v > 255 ? 255 : v));
    }
  }
}

/* BicubicTab_i (imgwarp.cpp initInterTab2D): 15-bit weights of the float
 * products cy[k1]*cx[k2], rounded; when they do not sum to 32768 the
 * difference goes to the largest (or smallest) entry among taps (2..3, 2..3),
 * the range OpenCV's adjustment loop scans (ksize2 = ksize/2). */
static void bicubic_itab(int fx, int fy, int w[16]) {
  float cx[4], cy[4];
  cubic_coeffs(fx * (1.f / 32), cx);
  cubic_coeffs(fy * (1.f / 32), cy);
  int isum = 0;
  for (int k1 = 0; k1 < 4; k1++)
    for (int k2 = 0; k2 < 4; k2++) {
      float v = cy[k1] * cx[k2];
      long r = lrintf(v * 32768.f);
      int iv = r < -32768 ? -32768 : (r > 32767 ? 32767 : (int)r);
      w[k1 * 4 + k2] = iv;
      isum += iv;
    }
  if (isum != 32768) {
    int diff = isum - 32768, Mk1 = 2, Mk2 = 2, mk1 = 2, mk2 = 2;
    for (int k1 = 2; k1 < 4; k1++)
      for (int k2 = 2; k2 < 4; k2++) {
        if (w[k1 * 4 + k2] < w[mk1 * 4 + mk2]) { mk1 = k1; mk2 = k2; }
        else if (w[k1 * 4 + k2] > w[Mk1 * 4 + Mk2]) { Mk1 = k1; Mk2 = k2; }
      }
    if (diff < 0) w[Mk1 * 4 + Mk2] -= diff;
    else w[mk1 * 4 + mk2] -= diff;
  }
}

/* cv::remap on CV_8U, INTER_CUBIC (remapBicubic with FixedPtCast<int,uchar,15>) */
void ora_remap_u8_cubic(const uint8_t *src, int sw, int sh, const real *map, int dw, int dh, uint8_t *dst,
                        int border, uint8_t cval) {
  static int tab[32][32][16];
  static int init = 0;
  if (!init) {
    for (int fy = 0; fy < 32; fy++)
      for (int fx = 0; fx < 32; fx++) bicubic_itab(fx, fy, tab[fy][fx]);
    init = 1;
  }
#pragma omp parallel for schedule(static)
  for (int y = 0; y < dh; y++) {
    for (int x = 0; x < dw; x++) {
      size_t o = (size_t)y * dw + x;
      int X = quant32(map[2 * o]), Y = quant32(map[2 * o + 1]);
      int bx = sat_short(X >> 5) - 1, by = sat_short(Y >> 5) - 1;
      const int *w = tab[Y & 31][X & 31];
      int sum;
      if ((unsigned)bx < (unsigned)(sw > 3 ? sw - 3 : 0) && (unsigned)by < (unsigned)(sh > 3 ? sh - 3 : 0)) {
        sum = 0;
        for (int r = 0; r < 4; r++)
          for (int c = 0; c < 4; c++) sum += src[(size_t)(by + r) * sw + bx + c] * w[r * 4 + c];
      } else {
        if (border == ORA_BORDER_TRANSPARENT &&
            ((unsigned)(bx + 1) >= (unsigned)sw || (unsigned)(by + 1) >= (unsigned)sh))
          continue;
        if (border == ORA_BORDER_CONSTANT && (bx >= sw || bx + 4 <= 0 || by >= sh || by + 4 <= 0)) {
          dst[o] = cval;
          continue;
        }
        sum = cval * 32768;
        for (int r = 0; r < 4; r++) {
          int yy = ora_border_interpolate(by + r, sh, border);
          if (yy < 0) continue;
          for (int c = 0; c < 4; c++) {
            int xx = ora_border_interpolate(bx + c, sw, border);
            if (xx >= 0) sum += (src[(size_t)yy * sw + xx] - cval) * w[r * 4 + c];
          }
        }
      }
      int v = (sum + (1 << 14)) >> 15;
      dst[o] = (uint8_t)(v < 0 ? 0 : (v > 255 ? 255 : v));
    }
  }
}

void ora_bicubic_itab(int fx, int fy, int *w) { bicubic_itab(fx, fy, w); }

/* cv::erode with a kw x kh rectangle, anchor at the centre */
void ora_erode_rect_u8(const uint8_t *src, int w, int h, int kw, int kh, uint8_t *dst, int border,
                       uint8_t cval) {
  const int ax = kw / 2, ay = kh / 2;
  uint8_t *tmp = (uint8_t *)xcalloc((size_t)w * h, 1);
  for (int y = 0; y < h; y++) {
    for (int x = 0; x < w; x++) {
      uint8_t m = 255;
      for (int j = -ay; j < kh - ay; j++)
        for (int i = -ax; i < kw - ax; i++) {
          int xx = ora_border_interpolate(x + i, w, border);
          int yy = ora_border_interpolate(y + j, h, border);
          uint8_t v = (xx < 0 || yy < 0) ? cval : src[(size_t)yy * w + xx];
          if (v < m) m = v;
        }
      tmp[(size_t)y * w + x] = m;
    }
  }
  memcpy(dst, tmp, (size_t)w * h);
  free(tmp);
}

/* cv::meanStdDev with an optional mask: double sums, population stdev */
void ora_mean_stddev(const real *src, const uint8_t *mask, int n, double *mean, double *stdev) {
  double s = 0, sq = 0;
  long cnt = 0;
  for (int i = 0; i < n; i++) {
    if (mask && !mask[i]) continue;
    double v = src[i];
    s += v;
    sq += v * v;
    cnt++;
  }
  double scale = cnt ? 1. / cnt : 0.;
  s *= scale;
  *mean = s;
  double var = sq * scale - s * s;
  *stdev = sqrt(var > 0 ? var : 0.);
}

/* cv::invert(DECOMP_CHOLESKY) on the parameter-space Hessian.
 * For n <= 3 OpenCV inverts through the determinant (fails iff det == 0);
 * for larger n it runs hal::Cholesky (CholImpl: double accumulation, a pivot
 * below FLT_EPSILON fails).  Returns 1 on success and zeroes A on failure. */
int ora_cholesky_invert(real *A, int n) {
#define M(i, j) A[(i)*n + (j)]
  if (n <= 3) {
    double d;
    if (n == 1) {
      d = M(0, 0);
      if (d == 0.) goto fail;
      M(0, 0) = (real)(1. / d);
      return 1;
    }
    if (n == 2) {
      d = (double)M(0, 0) * M(1, 1) - (double)M(0, 1) * M(1, 0);
      if (d == 0.) goto fail;
      d = 1. / d;
      double t0 = M(0, 0) * d, t1 = M(1, 1) * d;
      double t2 = -M(0, 1) * d, t3 = -M(1, 0) * d;
      M(1, 1) = (real)t0;
      M(0, 0) = (real)t1;
      M(0, 1) = (real)t2;
      M(1, 0) = (real)t3;
      return 1;
    }
    d = M(0, 0) * ((double)M(1, 1) * M(2, 2) - (double)M(1, 2) * M(2, 1)) -
        M(0, 1) * ((double)M(1, 0) * M(2, 2) - (double)M(1, 2) * M(2, 0)) +
        M(0, 2) * ((double)M(1, 0) * M(2, 1) - (double)M(1, 1) * M(2, 0));
    if (d == 0.) goto fail;
    d = 1. / d;
    double t[9];
    t[0] = ((double)M(1, 1) * M(2, 2) - (double)M(1, 2) * M(2, 1)) * d;
    t[1] = ((double)M(0, 2) * M(2, 1) - (double)M(0, 1) * M(2, 2)) * d;
    t[2] = ((double)M(0, 1) * M(1, 2) - (double)M(0, 2) * M(1, 1)) * d;
    t[3] = ((double)M(1, 2) * M(2, 0) - (double)M(1, 0) * M(2, 2)) * d;
    t[4] = ((double)M(0, 0) * M(2, 2) - (double)M(0, 2) * M(2, 0)) * d;
    t[5] = ((double)M(0, 2) * M(1, 0) - (double)M(0, 0) * M(1, 2)) * d;
    t[6] = ((double)M(1, 0) * M(2, 1) - (double)M(1, 1) * M(2, 0)) * d;
    t[7] = ((double)M(0, 1) * M(2, 0) - (double)M(0, 0) * M(2, 1)) * d;
    t[8] = ((double)M(0, 0) * M(1, 1) - (double)M(0, 1) * M(1, 0)) * d;
    for (int i = 0; i < 9; i++) A[i] = (real)t[i];
    return 1;
  }
  {
    real L[64], B[64];
    memcpy(L, A, sizeof(real) * n * n);
    for (int i = 0; i < n * n; i++) B[i] = 0;
    for (int i = 0; i < n; i++) B[i * n + i] = 1;
    double s;
    int i, j, k;
    for (i = 0; i < n; i++) {
      for (j = 0; j < i; j++) {
        s = L[i * n + j];
        for (k = 0; k < j; k++) s -= (double)L[i * n + k] * L[j * n + k];
        L[i * n + j] = (real)(s * L[j * n + j]);
      }
      s = L[i * n + i];
      for (k = 0; k < j; k++) {
        double t = L[i * n + k];
        s -= t * t;
      }
      if (s < FLT_EPSILON) goto fail;
      L[i * n + i] = (real)(1. / sqrt(s));
    }
    for (i = 0; i < n; i++)
      for (j = 0; j < n; j++) {
        s = B[i * n + j];
        for (k = 0; k < i; k++) s -= (double)L[i * n + k] * B[k * n + j];
        B[i * n + j] = (real)(s * L[i * n + i]);
      }
    for (i = n - 1; i >= 0; i--)
      for (j = 0; j < n; j++) {
        s = B[i * n + j];
        for (k = n - 1; k > i; k--) s -= (double)L[k * n + i] * B[k * n + j];
        B[i * n + j] = (real)(s * L[i * n + i]);
      }
    memcpy(A, B, sizeof(real) * n * n);
    return 1;
  }
fail:
  for (int i = 0; i < n * n; i++) A[i] = 0;
  return 0;
#undef M
}

/* ecc_differentiate (ecc2.cc:142-169): gx = sepFilter2D(d5 along x, s3 along
 * y), gy = sepFilter2D(s3 along x, d5 along y), BORDER_REPLICATE */
void ora_ecc_differentiate(const real *src, int w, int h, real *gx, real *gy) {
  const real d5[5] = {(real)(1.f / 12.f), (real)(-2.f / 3.f), 0, (real)(2.f / 3.f), (real)(-1.f / 12.f)};
  const real s3[3] = {(real)0.25f, (real)0.5f, (real)0.25f};
  ora_sep_filter(src, w, h, gx, d5, 5, s3, 3, ORA_BORDER_REPLICATE, 0);
  ora_sep_filter(src, w, h, gy, s3, 3, d5, 5, ORA_BORDER_REPLICATE, 0);
}

/* ------------------------------------------------------------------------ */
/* Image transforms (c_image_transform.cc).                                  */
/* Canonical map arithmetic (shared with the HIP kernels):                   */
/*   translation : sx = x + t0, sy = y + t1                  (:150-170)      */
/*   euclidean   : xx = x-Cx, yy = y-Cy; u = fma(ca,xx,-(sa*yy));             */
/*                 v = fma(sa,xx,ca*yy); sx = fma(s,u,Tx); sy = fma(s,v,Ty)   */
/*                 with sa,ca = (real)sin/cos((double)angle)      (:587-614)  */
/*   affine      : sx = fma(a01,y,a00*x) + a02; sy = fma(a11,y,a10*x) + a12   */
/*                                                                (:926-946)  */
/* ------------------------------------------------------------------------ */
static int motion_nparams(int motion) {
  switch (motion) {
    case ORA_MOTION_TRANSLATION: return 2;
    case ORA_MOTION_EUCLIDEAN: return 3;
    case ORA_MOTION_SCALED_EUCLIDEAN: return 4;
    default: return 6;
  }
}

void ora_xform_reset(ora_xform *t, int motion) {
  memset(t, 0, sizeof(*t));
  t->motion = motion;
  t->nparams = motion_nparams(motion);
  t->fixed_scale = 1;
  if (motion == ORA_MOTION_SCALED_EUCLIDEAN) t->p[3] = 1;
  if (motion == ORA_MOTION_AFFINE) {
    t->p[0] = 1;
    t->p[4] = 1;
  }
}

static void euclid_params(const ora_xform *t, const real *p, real *Tx, real *Ty, real *a, real *s) {
  *Tx = p[0];
  *Ty = p[1];
  *a = p[2];
  /* c_euclidean_image_transform::get_parameters (:376-414): a fixed scale
   * reads the transform's own _scale, not an entry of p */
  *s = t->motion == ORA_MOTION_SCALED_EUCLIDEAN ? p[3] : t->fixed_scale;
}

void ora_xform_create_map(const ora_xform *t, int w, int h, real *map) {
  if (t->motion == ORA_MOTION_TRANSLATION) {
    const real t0 = t->p[0], t1 = t->p[1];
#pragma omp parallel for schedule(static)
    for (int y = 0; y < h; y++)
      for (int x = 0; x < w; x++) {
        real *mp = map + 2 * ((size_t)y * w + x);
        mp[0] = (real)x + t0;
        mp[1] = (real)y + t1;
      }
  } else if (t->motion == ORA_MOTION_AFFINE) {
    const real *a = t->p;
#pragma omp parallel for schedule(static)
    for (int y = 0; y < h; y++)
      for (int x = 0; x < w; x++) {
        real *mp = map + 2 * ((size_t)y * w + x);
        mp[0] = RFMA(a[1], (real)y, a[0] * (real)x) + a[2];
        mp[1] = RFMA(a[4], (real)y, a[3] * (real)x) + a[5];
      }
  } else {
    real Tx, Ty, ang, s;
    euclid_params(t, t->p, &Tx, &Ty, &ang, &s);
    const real sa = (real)sin((double)ang), ca = (real)cos((double)ang);
    const real Cx = t->C[0], Cy = t->C[1];
#pragma omp parallel for schedule(static)
    for (int y = 0; y < h; y++) {
      const real yy = (real)y - Cy;
      for (int x = 0; x < w; x++) {
        const real xx = (real)x - Cx;
        real *mp = map + 2 * ((size_t)y * w + x);
        real u = RFMA(ca, xx, -(sa * yy));
        real v = RFMA(sa, xx, ca * yy);
        mp[0] = RFMA(s, u, Tx);
        mp[1] = RFMA(s, v, Ty);
      }
    }
  }
}

/* scale_transfrom: translation scales T (:130-134), euclidean scales T and C
 * (:502-507), affine scales a02 and a12 (:911-915) */
void ora_xform_scale(ora_xform *t, double f) {
  if (t->motion == ORA_MOTION_TRANSLATION) {
    t->p[0] = (real)(t->p[0] * f);
    t->p[1] = (real)(t->p[1] * f);
  } else if (t->motion == ORA_MOTION_AFFINE) {
    t->p[2] = (real)(t->p[2] * f);
    t->p[5] = (real)(t->p[5] * f);
  } else {
    t->p[0] = (real)(t->p[0] * f);
    t->p[1] = (real)(t->p[1] * f);
    t->C[0] = (real)(t->C[0] * f);
    t->C[1] = (real)(t->C[1] * f);
  }
}

/* eps(): the convergence metric of each transform.  Euclidean with a fixed
 * scale takes ds from the transform's own scale (1), so its eps never falls
 * below max(W,H) -- the reference behaviour (:509-523) is kept. */
ora_real ora_xform_eps(const ora_xform *t, const real *dp, int w, int h) {
  if (t->motion == ORA_MOTION_TRANSLATION) {
    return (real)sqrt((double)dp[0] * dp[0] + (double)dp[1] * dp[1]);
  }
  if (t->motion == ORA_MOTION_AFFINE) {
    double a = (double)w * dp[0], b = (double)h * dp[1], c = dp[2];
    double d = (double)w * dp[3], e = (double)h * dp[4], f = dp[5];
    return (real)sqrt(a * a + b * b + c * c + d * d + e * e + f * f);
  }
  real dTx, dTy, da, ds;
  euclid_params(t, dp, &dTx, &dTy, &da, &ds);
  double sa = sin((double)da);
  double m = (double)(w > h ? w : h);
  double ws = w * sa, hs = h * sa, ms = m * ds;
  return (real)sqrt((double)dTx * dTx + (double)dTy * dTy + ws * ws + hs * hs + ms * ms);
}

void ora_xform_translation(const ora_xform *t, real T[2]) {
  if (t->motion == ORA_MOTION_AFFINE) {
    T[0] = t->p[2];
    T[1] = t->p[5];
  } else {
    T[0] = t->p[0];
    T[1] = t->p[1];
  }
}

void ora_xform_set_translation(ora_xform *t, const real T[2]) {
  if (t->motion == ORA_MOTION_AFFINE) {
    t->p[2] = T[0];
    t->p[5] = T[1];
  } else {
    t->p[0] = T[0];
    t->p[1] = T[1];
  }
}

/* create_steepest_descent_images: translation (:262-269), euclidean
 * (:587-668), affine (:1045-1068).  Canonical:
 *   A = fma(sa,xx,ca*yy), B = fma(ca,xx,-(sa*yy))
 *   J_angle = s * fma(gy, B, -(gx*A)),  J_scale = fma(gx, B, gy*A) */
void ora_xform_sd_images(const ora_xform *t, const real *gx, const real *gy, int w, int h, real **J) {
  const int np = t->nparams;
  if (t->motion == ORA_MOTION_TRANSLATION) {
    memcpy(J[0], gx, sizeof(real) * w * h);
    memcpy(J[1], gy, sizeof(real) * w * h);
    return;
  }
  if (t->motion == ORA_MOTION_AFFINE) {
#pragma omp parallel for schedule(static)
    for (int y = 0; y < h; y++)
      for (int x = 0; x < w; x++) {
        size_t o = (size_t)y * w + x;
        J[0][o] = gx[o] * (real)x;
        J[1][o] = gx[o] * (real)y;
        J[2][o] = gx[o];
        J[3][o] = gy[o] * (real)x;
        J[4][o] = gy[o] * (real)y;
        J[5][o] = gy[o];
      }
    return;
  }
  real Tx, Ty, ang, s;
  euclid_params(t, t->p, &Tx, &Ty, &ang, &s);
  const real sa = (real)sin((double)ang), ca = (real)cos((double)ang);
  const real Cx = t->C[0], Cy = t->C[1];
#pragma omp parallel for schedule(static)
  for (int y = 0; y < h; y++) {
    const real yy = (real)y - Cy;
    for (int x = 0; x < w; x++) {
      size_t o = (size_t)y * w + x;
      const real xx = (real)x - Cx;
      real A = RFMA(sa, xx, ca * yy), B = RFMA(ca, xx, -(sa * yy));
      J[0][o] = gx[o];
      J[1][o] = gy[o];
      J[2][o] = s * RFMA(gy[o], B, -(gx[o] * A));
      if (np > 3) J[3][o] = RFMA(gx[o], B, gy[o] * A);
    }
  }
}

/* the equivalent 2x3 map x' = m0 x + m1 y + m2, y' = m3 x + m4 y + m5 */
void ora_xform_matrix(const ora_xform *t, double m[6]) {
  if (t->motion == ORA_MOTION_TRANSLATION) {
    m[0] = 1; m[1] = 0; m[2] = t->p[0];
    m[3] = 0; m[4] = 1; m[5] = t->p[1];
  } else if (t->motion == ORA_MOTION_AFFINE) {
    for (int i = 0; i < 6; i++) m[i] = t->p[i];
  } else {
    real Tx, Ty, ang, s;
    euclid_params(t, t->p, &Tx, &Ty, &ang, &s);
    double sa = (real)sin((double)ang), ca = (real)cos((double)ang);
    m[0] = s * ca; m[1] = -s * sa; m[2] = Tx - s * (ca * t->C[0] - sa * t->C[1]);
    m[3] = s * sa; m[4] = s * ca; m[5] = Ty - s * (sa * t->C[0] + ca * t->C[1]);
  }
}

/* ------------------------------------------------------------------------ */
/* Forward-additive ECC, one pyramid level (ecc2.cc:1247-1365).              */
/* ------------------------------------------------------------------------ */
static double dot_rows(const real *a, const real *b, int w, int h) {
  double *rs = (double *)xcalloc(h, sizeof(double));
#pragma omp parallel for schedule(static)
  for (int y = 0; y < h; y++) {
    double s = 0;
    const real *pa = a + (size_t)y * w, *pb = b + (size_t)y * w;
    for (int x = 0; x < w; x++) s += (double)(pa[x] * pb[x]);
    rs[y] = s;
  }
  double s = 0;
  for (int y = 0; y < h; y++) s += rs[y];
  free(rs);
  return s;
}

int ora_ecc_fa_align(const real *ref, const uint8_t *ref_mask, int w, int h, const real *cur,
                     const uint8_t *cur_mask, int cw, int chh, ora_xform *t, int interp,
                     int max_iterations, double max_eps, double update_step_scale,
                     ora_ecc_level_result *res) {
  const size_t N = (size_t)w * h;
  const int np = t->nparams;
  real *map = (real *)xcalloc(2 * N, sizeof(real));
  real *gw = (real *)xcalloc(N, sizeof(real));
  real *gx = (real *)xcalloc(N, sizeof(real));
  real *gy = (real *)xcalloc(N, sizeof(real));
  real *rhs = (real *)xcalloc(N, sizeof(real));
  uint8_t *wmask = (uint8_t *)xcalloc(N, 1);
  real *J[6];
  for (int i = 0; i < np; i++) J[i] = (real *)xcalloc(N, sizeof(real));
  uint8_t *cmask_all = NULL;
  if (!cur_mask) {
    cmask_all = (uint8_t *)xcalloc((size_t)cw * chh, 1);
    memset(cmask_all, 255, (size_t)cw * chh);
    cur_mask = cmask_all;
  }
  if (max_eps <= 0) max_eps = 1e-3;

  int failed = 0, it = 0;
  double eps = FLT_MAX;
  while (it++ < max_iterations) {
    ora_xform_create_map(t, w, h, map);
    ora_remap(cur, cw, chh, map, w, h, gw, interp, ORA_BORDER_REPLICATE, 0);
    ora_ecc_differentiate(gw, w, h, gx, gy);
    ora_remap_u8(cur_mask, cw, chh, map, w, h, wmask, ORA_BORDER_CONSTANT, 0);
    for (size_t i = 0; i < N; i++) {
      uint8_t m = wmask[i] >= 255 ? 255 : 0;
      if (ref_mask && !ref_mask[i]) m = 0;
      wmask[i] = m;
      if (!m) gx[i] = gy[i] = 0;
    }
    double fMean, fStd, gMean, gStd;
    ora_mean_stddev(ref, wmask, (int)N, &fMean, &fStd);
    ora_mean_stddev(gw, wmask, (int)N, &gMean, &gStd);
    double ratio = gStd / fStd;

    ora_xform_sd_images(t, gx, gy, w, h, J);
    real H[36];
    for (int i = 0; i < np; i++)
      for (int j = 0; j <= i; j++) H[i * np + j] = (real)dot_rows(J[i], J[j], w, h);
    for (int i = 0; i < np; i++)
      for (int j = i + 1; j < np; j++) H[i * np + j] = H[j * np + i];
    if (!ora_cholesky_invert(H, np)) {
      failed = 1;
      break;
    }
    const real a = (real)(-ratio);
    const real c = (real)(gMean - ratio * fMean);
#pragma omp parallel for schedule(static)
    for (size_t i = 0; i < N; i++) {
      real r = RFMA(ref[i], a, gw[i]) - c;
      rhs[i] = wmask[i] ? r : 0;
    }
    real ep[6];
    for (int i = 0; i < np; i++) ep[i] = (real)dot_rows(J[i], rhs, w, h);
    real dp[6];
    for (int i = 0; i < np; i++) {
      double s = 0;
      for (int k = 0; k < np; k++) s += (double)H[i * np + k] * ep[k];
      dp[i] = (real)(-update_step_scale * s);
    }
    for (int i = 0; i < np; i++) t->p[i] = t->p[i] + dp[i];
    eps = ora_xform_eps(t, dp, w, h);
    if (eps < max_eps) break;
  }
  if (res) {
    res->num_iterations = it;
    res->failed = failed;
    res->eps = eps;
  }
  free(map); free(gw); free(gx); free(gy); free(rhs); free(wmask); free(cmask_all);
  for (int i = 0; i < np; i++) free(J[i]);
  return !failed;
}

/* ------------------------------------------------------------------------ */
/* invert_and_compose.  translation: p - dp (c_image_transform.h:164-167);   */
/* affine: inv(inv(M(p)) + M(dp)) (c_image_transform.h:312-318); euclidean:  */
/* M(p) * inv(M(dp)) with the increment's scale 1 + ds (1 when the scale is  */
/* fixed), parameters read back from the product (c_image_transform.cc:      */
/* 736-833).  cv::invertAffineTransform works in double on the stored        */
/* entries and stores its result; Matx products sum k = 0..2 in order.       */
/* ------------------------------------------------------------------------ */
static void inv_affine23(const real *M, real *iM) {
  double D = (double)M[0] * M[4] - (double)M[1] * M[3];
  D = D != 0 ? 1. / D : 0.;
  const double A11 = M[4] * D, A22 = M[0] * D, A12 = -M[1] * D, A21 = -M[3] * D;
  const double b1 = -A11 * M[2] - A12 * M[5];
  const double b2 = -A21 * M[2] - A22 * M[5];
  iM[0] = (real)A11; iM[1] = (real)A12; iM[2] = (real)b1;
  iM[3] = (real)A21; iM[4] = (real)A22; iM[5] = (real)b2;
}

static void euclid_mat(real tX, real tY, real ang, real scl, real cX, real cY, real *M) {
  const real sa = RSIN(ang), ca = RCOS(ang);
  M[0] = scl * ca; M[1] = -scl * sa; M[2] = tX - scl * ca * cX + scl * sa * cY;
  M[3] = scl * sa; M[4] = scl * ca;  M[5] = tY - scl * sa * cX - scl * ca * cY;
  M[6] = 0; M[7] = 0; M[8] = 1;
}

void ora_xform_invert_and_compose(const ora_xform *t, const real *dp, real *out) {
  if (t->motion == ORA_MOTION_TRANSLATION) {
    out[0] = t->p[0] - dp[0];
    out[1] = t->p[1] - dp[1];
    return;
  }
  if (t->motion == ORA_MOTION_AFFINE) {
    real a[6];
    inv_affine23(t->p, a);
    for (int i = 0; i < 6; i++) a[i] = a[i] + dp[i];
    inv_affine23(a, out);
    return;
  }
  real Tx, Ty, ang, s;
  euclid_params(t, t->p, &Tx, &Ty, &ang, &s);
  const real Cx = t->C[0], Cy = t->C[1];
  const int fix_scale = t->motion != ORA_MOTION_SCALED_EUCLIDEAN;
  const real sdp = fix_scale ? (real)1 : (real)1 + dp[3];
  real Mp[9], Md[9], Mi[9], R[9];
  euclid_mat(Tx, Ty, ang, s, Cx, Cy, Mp);
  euclid_mat(dp[0], dp[1], dp[2], sdp, Cx, Cy, Md);
  inv_affine23(Md, Mi);
  Mi[6] = 0; Mi[7] = 0; Mi[8] = 1;
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) {
      real acc = 0;
      for (int k = 0; k < 3; k++) acc = acc + Mp[i * 3 + k] * Mi[k * 3 + j];
      R[i * 3 + j] = acc;
    }
  const real m00 = R[0], m10 = R[3];
  const real rs = fix_scale ? s : RSQRT(m00 * m00 + m10 * m10);
  const real ra = RATAN2(m10, m00);
  const real rca = RCOS(ra), rsa = RSIN(ra);
  out[0] = R[2] + rs * rca * Cx - rs * rsa * Cy;
  out[1] = R[5] + rs * rsa * Cx + rs * rca * Cy;
  out[2] = ra;
  if (!fix_scale) out[3] = rs;
}

/* cv::solve(H, v, DECOMP_CHOLESKY) of the normal equations: the failure rule
 * of ora_cholesky_invert (determinant for n <= 3, float Cholesky pivots
 * otherwise); a failed solve yields dp = 0.  Returns 1 on success. */
static int solve_normal(const real *H, const real *v, int n, real *dp) {
  real A[36];
  memcpy(A, H, sizeof(real) * n * n);
  if (!ora_cholesky_invert(A, n)) {
    for (int i = 0; i < n; i++) dp[i] = 0;
    return 0;
  }
  for (int i = 0; i < n; i++) {
    double s = 0;
    for (int k = 0; k < n; k++) s += (double)A[i * n + k] * v[k];
    dp[i] = (real)s;
  }
  return 1;
}

static void xform_with(const ora_xform *t, const real *p, ora_xform *out) {
  *out = *t;
  memcpy(out->p, p, sizeof(real) * 6);
}

/* H = J^T J (lower triangle mirrored) and v = J^T rhs from n SD images */
static void normal_eqs(real **J, const real *rhs, int n, int w, int h, real *H, real *v) {
  for (int i = 0; i < n; i++) {
    if (v) v[i] = (real)dot_rows(J[i], rhs, w, h);
    for (int j = 0; j <= i; j++) H[i * n + j] = (real)dot_rows(J[i], J[j], w, h);
  }
  for (int i = 0; i < n; i++)
    for (int j = i + 1; j < n; j++) H[i * n + j] = H[j * n + i];
}

/* the gemm(Hp, deltap, -1, v, 2) of the LM step quality: dS = dp.(2v - Hp dp) */
static double lm_predicted(const real *Hp, const real *dp, const real *v, int n) {
  double dS = 0;
  for (int i = 0; i < n; i++) {
    double s = 0;
    for (int k = 0; k < n; k++) s += (double)Hp[i * n + k] * dp[k];
    const real td = (real)(-s + 2.0 * (double)v[i]);
    dS += (double)dp[i] * td;
  }
  return dS;
}

/* ------------------------------------------------------------------------ */
/* c_ecclm (ecc2.cc:1371-1650): forward-additive Levenberg-Marquardt.        */
/* compute_remap (:1440-1473): g = remap(cur, LINEAR, REPLICATE), mask =     */
/* remap(cur mask or 255, LINEAR, CONSTANT 0) >= 250 (ecc2.cc:178-213) and    */
/* the 5x5-eroded reference mask; rhs = g - f, 0 off the mask.               */
/* compute_jac (:1480-1523): gradients of g zeroed off the mask, J at the    */
/* evaluated parameters, H = J^T J, v = J^T rhs.                             */
/* ------------------------------------------------------------------------ */
typedef struct lm_ws {
  real *map, *gw, *rhs, *gx, *gy;
  uint8_t *m, *all;
  real *J[6];
} lm_ws;

static void lm_ws_init(lm_ws *ws, size_t N, int np, size_t cn) {
  ws->map = (real *)xcalloc(2 * N, sizeof(real));
  ws->gw = (real *)xcalloc(N, sizeof(real));
  ws->rhs = (real *)xcalloc(N, sizeof(real));
  ws->gx = (real *)xcalloc(N, sizeof(real));
  ws->gy = (real *)xcalloc(N, sizeof(real));
  ws->m = (uint8_t *)xcalloc(N, 1);
  ws->all = (uint8_t *)xcalloc(cn, 1);
  memset(ws->all, 255, cn);
  for (int i = 0; i < 6; i++) ws->J[i] = i < np ? (real *)xcalloc(N, sizeof(real)) : NULL;
}

static void lm_ws_free(lm_ws *ws) {
  free(ws->map); free(ws->gw); free(ws->rhs); free(ws->gx); free(ws->gy); free(ws->m); free(ws->all);
  for (int i = 0; i < 6; i++) free(ws->J[i]);
}

static double lm_remap(const ora_xform *t, const real *params, const real *ref, const uint8_t *ref_mask, int w,
                       int h, const real *cur, const uint8_t *cur_mask, int cw, int chh, lm_ws *ws) {
  const size_t N = (size_t)w * h;
  ora_xform tp;
  xform_with(t, params, &tp);
  ora_xform_create_map(&tp, w, h, ws->map);
  ora_remap(cur, cw, chh, ws->map, w, h, ws->gw, ORA_INTER_LINEAR, ORA_BORDER_REPLICATE, 0);
  ora_remap_u8(cur_mask ? cur_mask : ws->all, cw, chh, ws->map, w, h, ws->m, ORA_BORDER_CONSTANT, 0);
  for (size_t i = 0; i < N; i++) {
    const int ok = ws->m[i] >= 250 && (!ref_mask || ref_mask[i]);
    ws->m[i] = ok ? 255 : 0;
    ws->rhs[i] = ok ? ws->gw[i] - ref[i] : 0;
  }
  return dot_rows(ws->rhs, ws->rhs, w, h);
}

int ora_ecc_lm_align(const real *ref, const uint8_t *ref_mask, int w, int h, const real *cur,
                     const uint8_t *cur_mask, int cw, int chh, ora_xform *t, int max_iterations,
                     double max_eps, double update_step_scale, ora_ecc_level_result *res) {
  const size_t N = (size_t)w * h;
  const int M = t->nparams;
  const double epse = 1e-4; /* c_ecclm::_max_epse (ecc2.h:145) */
  lm_ws ws;
  lm_ws_init(&ws, N, M, (size_t)cw * chh);
  real params[6], newp[6], H[36], Hp[36], v[6], dp[6];
  memcpy(params, t->p, sizeof(params));
  double lambda = 0.1, rms = 0, eps = FLT_MAX;
  int iteration = 0, converged = 0, recompute = 1;
  while (iteration < max_iterations) {
    if (recompute) rms = lm_remap(t, params, ref, ref_mask, w, h, cur, cur_mask, cw, chh, &ws);
    ora_ecc_differentiate(ws.gw, w, h, ws.gx, ws.gy);
    for (size_t i = 0; i < N; i++)
      if (!ws.m[i]) ws.gx[i] = ws.gy[i] = 0;
    ora_xform tp;
    xform_with(t, params, &tp);
    ora_xform_sd_images(&tp, ws.gx, ws.gy, w, h, ws.J);
    normal_eqs(ws.J, ws.rhs, M, w, h, H, v);
    const double err = rms;
    if (err < 1) {
      converged = 1;
      break;
    }
    memcpy(Hp, H, sizeof(H));
    while (iteration++ < max_iterations) {
      recompute = 1;
      for (int i = 0; i < M; i++) H[i * M + i] = (real)((1 + lambda) * Hp[i * M + i]);
      solve_normal(H, v, M, dp);
      for (int i = 0; i < M; i++) newp[i] = RFMA(dp[i], (real)(-update_step_scale), params[i]);
      for (int i = M; i < 6; i++) newp[i] = params[i];
      eps = ora_xform_eps(t, dp, w, h);
      if (eps <= max_eps) {
        memcpy(params, newp, sizeof(params));
        converged = 1;
        break;
      }
      const double newerr = lm_remap(t, newp, ref, ref_mask, w, h, cur, cur_mask, cw, chh, &ws);
      if (newerr > err) {
        if (lambda > 1e6) break;
        lambda *= 10.0f;
        continue;
      }
      memcpy(params, newp, sizeof(params));
      recompute = 0;
      rms = newerr;
      const double diff = err - newerr;
      if (diff < err * epse) {
        converged = 1;
        break;
      }
      const double dS = lm_predicted(Hp, dp, v, M);
      const double rho = fabs(dS) > 1e-9f ? diff / fabs(dS) : diff;
      if (rho > 0.25) lambda = fmax(1e-8, 0.2 * lambda);
      else if (rho < 0.1) lambda = lambda < 1.0 ? 1.0 : lambda * 10.0;
      break;
    }
    if (converged) break;
  }
  memcpy(t->p, params, sizeof(params));
  if (res) {
    res->num_iterations = iteration;
    res->failed = !converged;
    res->eps = eps;
  }
  lm_ws_free(&ws);
  return converged;
}

/* ------------------------------------------------------------------------ */
/* Inverse-compositional ECC.  Both variants precompute J on the reference  */
/* (gradients zeroed where the reference mask is 0, J at the parameters of  */
/* the transform `jx` holds when the precompute runs) and H = J^T J.         */
/* ------------------------------------------------------------------------ */
static void ic_precompute(const real *ref, const uint8_t *ref_mask, int w, int h, const ora_xform *jx, lm_ws *ws,
                          real *Hp) {
  const size_t N = (size_t)w * h;
  ora_ecc_differentiate(ref, w, h, ws->gx, ws->gy);
  if (ref_mask)
    for (size_t i = 0; i < N; i++)
      if (!ref_mask[i]) ws->gx[i] = ws->gy[i] = 0;
  ora_xform_sd_images(jx, ws->gx, ws->gy, w, h, ws->J);
  normal_eqs(ws->J, NULL, jx->nparams, w, h, Hp, NULL);
}

static double count_nonzero(const uint8_t *m, size_t n) {
  double c = 0;
  for (size_t i = 0; i < n; i++) c += m[i] != 0;
  return c;
}

/* c_ecc_inverse_compositional::align (ecc2.cc:1690-1792): rhs = remap(cur,
 * LINEAR, REPLICATE) - f, zeroed off remap(cur mask or 255, LINEAR,
 * CONSTANT 0) >= 250; rms = |rhs|^2 RMA^2/CMA^2; dp = H^-1 J^T rhs RMA/CMA;
 * stop when rms stops decreasing or eps(dp) < max_eps; p = p o (s dp)^-1. */
int ora_ecc_ic_align(const real *ref, const uint8_t *ref_mask, int w, int h, const real *cur,
                     const uint8_t *cur_mask, int cw, int chh, ora_xform *t, const ora_xform *jx,
                     int max_iterations, double max_eps, double update_step_scale,
                     ora_ecc_level_result *res) {
  const size_t N = (size_t)w * h;
  const int M = t->nparams;
  lm_ws ws;
  lm_ws_init(&ws, N, M, (size_t)cw * chh);
  real H[36], v[6], dp[6], sdp[6], newp[6];
  ic_precompute(ref, ref_mask, w, h, jx, &ws, H);
  const double RMA = ref_mask ? count_nonzero(ref_mask, N) : (double)N;
  double rmsold = FLT_MAX, eps = FLT_MAX;
  int it = 0;
  while (it++ < max_iterations) {
    ora_xform_create_map(t, w, h, ws.map);
    ora_remap(cur, cw, chh, ws.map, w, h, ws.gw, ORA_INTER_LINEAR, ORA_BORDER_REPLICATE, 0);
    ora_remap_u8(cur_mask ? cur_mask : ws.all, cw, chh, ws.map, w, h, ws.m, ORA_BORDER_CONSTANT, 0);
    double CMA = 0;
    for (size_t i = 0; i < N; i++) {
      const int ok = ws.m[i] >= 250;
      ws.rhs[i] = ok ? ws.gw[i] - ref[i] : 0;
      CMA += ok;
    }
    const double rmsnew = dot_rows(ws.rhs, ws.rhs, w, h) * (RMA * RMA) / (CMA * CMA);
    for (int i = 0; i < M; i++) v[i] = (real)((real)dot_rows(ws.J[i], ws.rhs, w, h) * (RMA / CMA));
    solve_normal(H, v, M, dp);
    if (rmsnew >= rmsold) break;
    for (int i = 0; i < M; i++) sdp[i] = (real)(update_step_scale * dp[i]);
    ora_xform_invert_and_compose(t, sdp, newp);
    rmsold = rmsnew;
    memcpy(t->p, newp, sizeof(real) * M);
    eps = ora_xform_eps(t, dp, w, h);
    if (eps < max_eps) break;
  }
  if (res) {
    res->num_iterations = it;
    res->failed = 0;
    res->eps = eps;
  }
  lm_ws_free(&ws);
  return 1;
}

/* c_ecclm_inverse_compositional (ecc2.cc:1796-2086).  compute_rhs
 * (:1892-1917): bad = remap(~cur mask, NEAREST, CONSTANT 255) | ~ref mask,
 * rhs = g - f zeroed at bad, CMA = area - bad, rms = |rhs|^2 RMA^2/CMA^2;
 * compute_v (:1919-1924): v = J^T rhs RMA/CMA. */
static double iclm_rhs(const ora_xform *t, const real *params, const real *ref, const uint8_t *ref_mask, int w, int h,
                       const real *cur, const uint8_t *cur_mask, int cw, int chh, double RMA, lm_ws *ws,
                       double *CMA) {
  const size_t N = (size_t)w * h;
  ora_xform tp;
  xform_with(t, params, &tp);
  ora_xform_create_map(&tp, w, h, ws->map);
  ora_remap(cur, cw, chh, ws->map, w, h, ws->gw, ORA_INTER_LINEAR, ORA_BORDER_REPLICATE, 0);
  double bad = 0;
  for (size_t i = 0; i < N; i++) {
    long X = RLRINT(ws->map[2 * i]), Y = RLRINT(ws->map[2 * i + 1]);
    X = X < SHRT_MIN ? SHRT_MIN : (X > SHRT_MAX ? SHRT_MAX : X);
    Y = Y < SHRT_MIN ? SHRT_MIN : (Y > SHRT_MAX ? SHRT_MAX : Y);
    int b = (X < 0 || X >= cw || Y < 0 || Y >= chh) ? 1 : (cur_mask && !cur_mask[(size_t)Y * cw + X]);
    if (ref_mask && !ref_mask[i]) b = 1;
    ws->m[i] = b ? 0 : 255;
    ws->rhs[i] = b ? 0 : ws->gw[i] - ref[i];
    bad += b;
  }
  *CMA = (double)N - bad;
  return dot_rows(ws->rhs, ws->rhs, w, h) * (RMA * RMA) / (*CMA * *CMA);
}

int ora_ecc_iclm_align(const real *ref, const uint8_t *ref_mask, int w, int h, const real *cur,
                       const uint8_t *cur_mask, int cw, int chh, ora_xform *t, const ora_xform *jx,
                       int max_iterations, double max_eps, double update_step_scale,
                       ora_ecc_level_result *res) {
  const size_t N = (size_t)w * h;
  const int M = t->nparams;
  lm_ws ws;
  lm_ws_init(&ws, N, M, (size_t)cw * chh);
  real Hp[36], H[36], v[6], dp[6], sdp[6], params[6], newp[6];
  ic_precompute(ref, ref_mask, w, h, jx, &ws, Hp);
  const double RMA = ref_mask ? count_nonzero(ref_mask, N) : (double)N;
  memcpy(params, t->p, sizeof(params));
  double lambda = 0.001, dpe = 0, last_rms = 0, CMA = 0, err = 0, newerr = 0;
  int it = 0, recompute = 1;
  while (it < max_iterations) {
    if (recompute) last_rms = iclm_rhs(t, params, ref, ref_mask, w, h, cur, cur_mask, cw, chh, RMA, &ws, &CMA);
    for (int i = 0; i < M; i++) v[i] = (real)((real)dot_rows(ws.J[i], ws.rhs, w, h) * (RMA / CMA));
    memcpy(H, Hp, sizeof(H));
    err = last_rms;
    do {
      ++it;
      recompute = 1;
      for (int i = 0; i < M; i++) H[i * M + i] = (real)((1 + lambda) * Hp[i * M + i]);
      solve_normal(H, v, M, dp);
      for (int i = 0; i < M; i++) sdp[i] = (real)(update_step_scale * dp[i]);
      ora_xform tp;
      xform_with(t, params, &tp);
      ora_xform_invert_and_compose(&tp, sdp, newp);
      for (int i = M; i < 6; i++) newp[i] = params[i];
      newerr = last_rms = iclm_rhs(t, newp, ref, ref_mask, w, h, cur, cur_mask, cw, chh, RMA, &ws, &CMA);
      if ((dpe = ora_xform_eps(t, dp, w, h)) < max_eps) break;
      const double dS = lm_predicted(Hp, dp, v, M);
      const double rho = (err - newerr) / (fabs(dS) > DBL_EPSILON ? dS : 1);
      if (rho > 0.25) {
        if (lambda > 1e-6) lambda = fmax(1e-6, lambda / 5);
      } else if (rho > 0.1) {
      } else if (lambda < 1) {
        lambda = 1;
      } else {
        lambda *= 10;
      }
      if (newerr < err) break;
    } while (it < max_iterations);
    if (newerr < err) {
      err = newerr;
      recompute = 0;
      memcpy(params, newp, sizeof(params));
    }
    if (dpe < max_eps) break;
  }
  memcpy(t->p, params, sizeof(params));
  if (res) {
    res->num_iterations = it;
    res->failed = 0;
    res->eps = dpe;
  }
  lm_ws_free(&ws);
  return 1;
}

/* compute_correlation (ecc2.cc:65-137): remap (LINEAR, CONSTANT 0), mask =
 * remap(mask) >= 254 & reference mask, Pearson correlation over the mask */
double ora_compute_correlation(const real *cur, const uint8_t *cur_mask, int cw, int chh,
                               const real *ref, const uint8_t *ref_mask, int w, int h,
                               const real *map) {
  const size_t N = (size_t)w * h;
  real *g = (real *)xcalloc(N, sizeof(real));
  uint8_t *m = (uint8_t *)xcalloc(N, 1);
  uint8_t *all = NULL;
  ora_remap(cur, cw, chh, map, w, h, g, ORA_INTER_LINEAR, ORA_BORDER_CONSTANT, 0);
  if (!cur_mask) {
    all = (uint8_t *)xcalloc((size_t)cw * chh, 1);
    memset(all, 255, (size_t)cw * chh);
    cur_mask = all;
  }
  ora_remap_u8(cur_mask, cw, chh, map, w, h, m, ORA_BORDER_CONSTANT, 0);
  long npix = 0;
  for (size_t i = 0; i < N; i++) {
    uint8_t v = m[i] >= 254 ? 255 : 0;
    if (ref_mask && !ref_mask[i]) v = 0;
    m[i] = v;
    npix += v != 0;
  }
  double m1, s1, m2, s2;
  ora_mean_stddev(ref, m, (int)N, &m1, &s1);
  ora_mean_stddev(g, m, (int)N, &m2, &s2);
  const real fm1 = (real)m1, fm2 = (real)m2;
  double covar = 0;
  for (size_t i = 0; i < N; i++) {
    if (!m[i]) continue;
    real a = ref[i] - fm1, b = g[i] - fm2;
    covar += (double)(a * b);
  }
  covar /= npix;
  free(g); free(m); free(all);
  return covar / (s1 * s2);
}

/* ------------------------------------------------------------------------ */
/* c_ecch pyramid + c_frame_registration::register_frame                    */
/* ------------------------------------------------------------------------ */
#define ORA_MAX_LEVELS 16

typedef struct level_img {
  int w, h;
  real *img;
  uint8_t *mask; /* eroded (or NULL) */
} level_img;

struct ora_reg {
  ora_reg_options o;
  int nlevels;
  int ref_w, ref_h;  /* full-resolution reference size */
  level_img ref[ORA_MAX_LEVELS];
  /* inverse-compositional J cache per level: the transform state of the
   * precompute, kept until the reference or the aligner's transform changes */
  ora_xform jx[ORA_MAX_LEVELS];
  int jvalid[ORA_MAX_LEVELS];
};

static void next_size(int w, int h, int *nw, int *nh) {
  /* c_ecch::compute_next_pyramid_layer_size (ecc2.h:290-293) */
  *nw = ((w + 1) >> 1) & ~1;
  *nh = ((h + 1) >> 1) & ~1;
}

static void smooth(real *img, int w, int h, double sigma) {
  /* c_ecch::set_reference_image / set_current_image smoothing (ecc2.cc:998-1006) */
  int ksize = 2 * ((int)(3 * sigma)) + 1;
  if (ksize < 3) ksize = 3;
  real k[64];
  ora_gaussian_kernel(ksize, sigma, k);
  real *tmp = (real *)xcalloc((size_t)w * h, sizeof(real));
  ora_sep_filter(img, w, h, tmp, k, ksize, k, ksize, ORA_BORDER_REPLICATE, 0);
  memcpy(img, tmp, sizeof(real) * w * h);
  free(tmp);
}

static int count_all(const uint8_t *m, size_t n) {
  for (size_t i = 0; i < n; i++)
    if (!m[i]) return 0;
  return 1;
}

/* build the level images; masks follow c_ecc_forward_additive's rules:
 * reference masks that are all-nonzero are dropped, partial ones eroded 5x5
 * (ecc2.cc:1187-1211); current masks default to 255 and partial ones are
 * eroded (ecc2.cc:1214-1234).  The mask passed down the pyramid is the
 * un-eroded one resized INTER_NEAREST (ecc2.cc:972-982). */
static void build_levels(const real *img0, const uint8_t *mask0, int w, int h, int nlev,
                         double sigma, int is_ref, int method, level_img *L) {
  /* erosion of partial masks: forward additive erodes both (ecc2.cc:1187-1234),
   * c_ecclm only the reference (ecc2.cc:1387-1400), the inverse-compositional
   * aligners neither (base c_ecc_align::set_*_image, ecc2.cc:585-630) */
  const int erode = method == ORA_ECC_FORWARD_ADDITIVE || (method == ORA_ECC_LM && is_ref);
  real *img = (real *)xcalloc((size_t)w * h, sizeof(real));
  memcpy(img, img0, sizeof(real) * w * h);
  uint8_t *mask = NULL;
  if (mask0) {
    mask = (uint8_t *)xcalloc((size_t)w * h, 1);
    memcpy(mask, mask0, (size_t)w * h);
  }
  if (sigma > 0) smooth(img, w, h, sigma);
  int cw = w, chh = h;
  for (int l = 0; l < nlev; l++) {
    L[l].w = cw;
    L[l].h = chh;
    L[l].img = (real *)xcalloc((size_t)cw * chh, sizeof(real));
    memcpy(L[l].img, img, sizeof(real) * cw * chh);
    L[l].mask = NULL;
    const size_t n = (size_t)cw * chh;
    if (mask && !(is_ref && count_all(mask, n))) {
      L[l].mask = (uint8_t *)xcalloc(n, 1);
      if (count_all(mask, n) || !erode) {
        memcpy(L[l].mask, mask, n);
      } else {
        ora_erode_rect_u8(mask, cw, chh, 5, 5, L[l].mask, ORA_BORDER_REPLICATE, 255);
      }
    }
    if (l < nlev - 1) {
      int nw, nh;
      next_size(cw, chh, &nw, &nh);
      real *nimg = (real *)xcalloc((size_t)nw * nh, sizeof(real));
      ora_pyr_down(img, cw, chh, nimg, nw, nh, ORA_BORDER_REFLECT101);
      free(img);
      img = nimg;
      if (mask) {
        uint8_t *nm = (uint8_t *)xcalloc((size_t)nw * nh, 1);
        ora_resize_nearest_u8(mask, cw, chh, nm, nw, nh);
        free(mask);
        mask = nm;
      }
      cw = nw;
      chh = nh;
    }
  }
  free(img);
  free(mask);
}

static void free_levels(level_img *L, int n) {
  for (int i = 0; i < n; i++) {
    free(L[i].img);
    free(L[i].mask);
    L[i].img = NULL;
    L[i].mask = NULL;
  }
}

ora_reg *ora_reg_create(const ora_reg_options *opts) {
  ora_reg *r = (ora_reg *)xcalloc(1, sizeof(ora_reg));
  r->o = *opts;
  return r;
}

void ora_reg_destroy(ora_reg *r) {
  if (!r) return;
  free_levels(r->ref, r->nlevels);
  free(r);
}

static void scale_image_half(const real *src, const uint8_t *m, int w, int h, real **dst,
                             uint8_t **dm, int *dw, int *dh) {
  /* scaleImage with scale 0.5 (c_frame_registration.cc:230-250) */
  *dw = (w + 1) / 2;
  *dh = (h + 1) / 2;
  *dst = (real *)xcalloc((size_t)(*dw) * (*dh), sizeof(real));
  ora_pyr_down(src, w, h, *dst, *dw, *dh, ORA_BORDER_REFLECT101);
  *dm = NULL;
  if (m) {
    *dm = (uint8_t *)xcalloc((size_t)(*dw) * (*dh), 1);
    ora_pyr_down_u8(m, w, h, *dm, *dw, *dh);
    for (size_t i = 0; i < (size_t)(*dw) * (*dh); i++) (*dm)[i] = (*dm)[i] >= 250 ? 255 : 0;
  }
}

int ora_reg_setup_reference(ora_reg *r, const real *img, const uint8_t *mask, int w, int h) {
  free_levels(r->ref, r->nlevels);
  r->ref_w = w;
  r->ref_h = h;
  const real *eimg = img;
  const uint8_t *emask = mask;
  real *simg = NULL;
  uint8_t *smask = NULL;
  int ew = w, eh = h;
  if (r->o.scale > 0 && r->o.scale != 1) {
    scale_image_half(img, mask, w, h, &simg, &smask, &ew, &eh);
    eimg = simg;
    emask = smask;
  }
  /* level count (ecc2.cc:1009-1026) */
  int min_size = r->o.ecch_minimum_image_size > 4 ? r->o.ecch_minimum_image_size : 4;
  int tw = ew, th = eh, lv = 1;
  for (;;) {
    int maxlevel = r->o.ecch_max_level;
    if (maxlevel >= 0 && lv >= (maxlevel > 0 ? maxlevel : 0)) break;
    int nw, nh;
    next_size(tw, th, &nw, &nh);
    if (nw < min_size || nh < min_size) break;
    tw = nw;
    th = nh;
    ++lv;
    if (lv >= ORA_MAX_LEVELS) break;
  }
  r->nlevels = lv;
  build_levels(eimg, emask, ew, eh, lv, r->o.reference_smooth_sigma, 1, r->o.ecc_method, r->ref);
  memset(r->jvalid, 0, sizeof(r->jvalid));
  free(simg);
  free(smask);
  return 1;
}

int ora_reg_num_levels(const ora_reg *r) { return r->nlevels; }

int ora_reg_level_size(const ora_reg *r, int lvl, int *w, int *h) {
  if (lvl < 0 || lvl >= r->nlevels) return 0;
  *w = r->ref[lvl].w;
  *h = r->ref[lvl].h;
  return 1;
}

int ora_reg_get_ref_level(const ora_reg *r, int lvl, real *img, uint8_t *mask) {
  if (lvl < 0 || lvl >= r->nlevels) return 0;
  size_t n = (size_t)r->ref[lvl].w * r->ref[lvl].h;
  if (img) memcpy(img, r->ref[lvl].img, sizeof(real) * n);
  if (mask) {
    if (r->ref[lvl].mask) memcpy(mask, r->ref[lvl].mask, n);
    else memset(mask, 255, n);
  }
  return 1;
}

/* c_ecch::align (ecc2.cc:1133-1176): coarse to fine, transform pre-scaled by
 * the width ratio; a failed level is logged only and skips the rescale */
static void ecch_align(ora_reg *r, level_img *cur, ora_xform *t, int *iters, int *failed_levels,
                       int *level_iters) {
  const int L = r->nlevels;
  int lvl = L - 1;
  if (lvl > 0) ora_xform_scale(t, (double)r->ref[lvl].w / (double)r->ref[0].w);
  int total = 0;
  for (; lvl >= 0; --lvl) {
    ora_ecc_level_result res;
    double max_eps = r->o.eps * (double)(1 << lvl);
    const level_img *R = &r->ref[lvl], *Cu = &cur[lvl];
    int ok;
    switch (r->o.ecc_method) {
      case ORA_ECC_FORWARD_ADDITIVE:
        ok = ora_ecc_fa_align(R->img, R->mask, R->w, R->h, Cu->img, Cu->mask, Cu->w, Cu->h, t, r->o.interpolation,
                              r->o.max_iterations, max_eps, r->o.update_step_scale, &res);
        break;
      case ORA_ECC_INVERSE_COMPOSITIONAL:
      case ORA_ECC_INVERSE_COMPOSITIONAL_LM:
        if (!r->jvalid[lvl]) {
          r->jx[lvl] = *t;
          r->jvalid[lvl] = 1;
        }
        if (r->o.ecc_method == ORA_ECC_INVERSE_COMPOSITIONAL)
          ok = ora_ecc_ic_align(R->img, R->mask, R->w, R->h, Cu->img, Cu->mask, Cu->w, Cu->h, t, &r->jx[lvl],
                                r->o.max_iterations, max_eps, r->o.update_step_scale, &res);
        else
          ok = ora_ecc_iclm_align(R->img, R->mask, R->w, R->h, Cu->img, Cu->mask, Cu->w, Cu->h, t, &r->jx[lvl],
                                  r->o.max_iterations, max_eps, r->o.update_step_scale, &res);
        break;
      default:
        ok = ora_ecc_lm_align(R->img, R->mask, R->w, R->h, Cu->img, Cu->mask, Cu->w, Cu->h, t, r->o.max_iterations,
                              max_eps, r->o.update_step_scale, &res);
        break;
    }
    if (!ok) {
      (*failed_levels)++;
    } else if (lvl > 0) {
      ora_xform_scale(t, (double)r->ref[lvl - 1].w / (double)r->ref[lvl].w);
    }
    if (level_iters) level_iters[lvl] += res.num_iterations;
    total += res.num_iterations;
  }
  *iters = total;
}

int ora_reg_register(ora_reg *r, const real *img, const uint8_t *mask, int w, int h,
                     ora_reg_status *st) {
  memset(st, 0, sizeof(*st));
  const real *eimg = img;
  const uint8_t *emask = mask;
  real *simg = NULL;
  uint8_t *smask = NULL;
  int ew = w, eh = h;
  if (r->o.scale > 0 && r->o.scale != 1) {
    scale_image_half(img, mask, w, h, &simg, &smask, &ew, &eh);
    eimg = simg;
    emask = smask;
  }
  level_img cur[ORA_MAX_LEVELS];
  memset(cur, 0, sizeof(cur));
  build_levels(eimg, emask, ew, eh, r->nlevels, r->o.input_smooth_sigma, 0, r->o.ecc_method, cur);
  /* forward-additive always holds a current mask (all 255 when none given) */
  ora_xform t;
  ora_xform_reset(&t, r->o.motion_type);
  real *map = (real *)xcalloc(2 * (size_t)r->ref[0].w * r->ref[0].h, sizeof(real));
  int code = 0;
  const int translation_first = r->o.motion_type != ORA_MOTION_TRANSLATION &&
                                r->o.ecch_estimate_translation_first && r->o.ecch_max_level != 0;
  if (translation_first) {
    ora_xform tt;
    ora_xform_reset(&tt, ORA_MOTION_TRANSLATION);
    real T[2];
    ora_xform_translation(&t, T);
    ora_xform_set_translation(&tt, T);
    /* c_ecch::set_image_transform to the translation transform and back
     * (c_frame_registration.cc:823,844) drops the inverse-compositional J */
    memset(r->jvalid, 0, sizeof(r->jvalid));
    ecch_align(r, cur, &tt, &st->iterations_translation, &st->levels_failed, NULL);
    memset(r->jvalid, 0, sizeof(r->jvalid));
    ora_xform_create_map(&tt, r->ref[0].w, r->ref[0].h, map);
    st->rho_translation = ora_compute_correlation(cur[0].img, cur[0].mask, cur[0].w, cur[0].h,
                                                  r->ref[0].img, r->ref[0].mask, r->ref[0].w,
                                                  r->ref[0].h, map);
    if (st->rho_translation < 0.75 * r->o.min_rho) code = 3;
    ora_xform_translation(&tt, T);
    ora_xform_set_translation(&t, T);
  }
  if (!code) {
    ecch_align(r, cur, &t, &st->iterations, &st->levels_failed, st->level_iterations);
    ora_xform_create_map(&t, r->ref[0].w, r->ref[0].h, map);
    st->rho = ora_compute_correlation(cur[0].img, cur[0].mask, cur[0].w, cur[0].h, r->ref[0].img,
                                      r->ref[0].mask, r->ref[0].w, r->ref[0].h, map);
    if (st->rho < r->o.min_rho) code = 4;
    if (r->o.scale > 0 && r->o.scale != 1) ora_xform_scale(&t, 1. / r->o.scale);
  }
  st->code = code;
  st->ok = code == 0;
  st->nparams = t.nparams;
  for (int i = 0; i < t.nparams; i++) st->params[i] = t.p[i];
  ora_xform_matrix(&t, st->matrix);
  free(map);
  free_levels(cur, r->nlevels);
  free(simg);
  free(smask);
  return st->ok;
}

/* ------------------------------------------------------------------------ */
/* c_frame_registration::base_remap (c_frame_registration.cc:1265-1338):     */
/* dst = remap(src); mask = remap(src_mask or 255, CONSTANT 0) >= 255, then  */
/* erode 5x5 with BORDER_CONSTANT 255                                        */
/* ------------------------------------------------------------------------ */
void ora_base_remap(const real *src, int sw, int sh, int cn, const uint8_t *src_mask,
                    const real *map, int dw, int dh, real *dst, uint8_t *dst_mask, int interp,
                    int border, real cval) {
  if (dst) {
    if (cn == 1) {
      ora_remap(src, sw, sh, map, dw, dh, dst, interp, border, cval);
    } else {
      real *plane = (real *)xcalloc((size_t)sw * sh, sizeof(real));
      real *out = (real *)xcalloc((size_t)dw * dh, sizeof(real));
      for (int c = 0; c < cn; c++) {
        for (size_t i = 0; i < (size_t)sw * sh; i++) plane[i] = src[i * cn + c];
        for (size_t i = 0; i < (size_t)dw * dh; i++) out[i] = dst[i * cn + c];
        ora_remap(plane, sw, sh, map, dw, dh, out, interp, border, cval);
        for (size_t i = 0; i < (size_t)dw * dh; i++) dst[i * cn + c] = out[i];
      }
      free(plane);
      free(out);
    }
  }
  if (dst_mask) {
    uint8_t *all = NULL;
    if (!src_mask) {
      all = (uint8_t *)xcalloc((size_t)sw * sh, 1);
      memset(all, 255, (size_t)sw * sh);
      src_mask = all;
    }
    /* the mask is remapped with the frame's interpolation (:1319-1327) */
    if (interp == ORA_INTER_CUBIC)
      ora_remap_u8_cubic(src_mask, sw, sh, map, dw, dh, dst_mask, ORA_BORDER_CONSTANT, 0);
    else
      ora_remap_u8(src_mask, sw, sh, map, dw, dh, dst_mask, ORA_BORDER_CONSTANT, 0);
    for (size_t i = 0; i < (size_t)dw * dh; i++) dst_mask[i] = dst_mask[i] >= 255 ? 255 : 0;
    ora_erode_rect_u8(dst_mask, dw, dh, 5, 5, dst_mask, ORA_BORDER_CONSTANT, 255);
    free(all);
  }
}

/* ------------------------------------------------------------------------ */
/* _weighted_average_update (c_frame_accumulation.cc:20-129): incremental    */
/* weighted mean, A += (I - A) * w / (W + w), canonical fma((I-A), f, A)     */
/* ------------------------------------------------------------------------ */
void ora_weighted_average_update(const real *src, int w, int h, int cn, const uint8_t *mask_u8,
                                 const real *weights_f, real *acc, real *accw) {
#pragma omp parallel for schedule(static)
  for (int y = 0; y < h; y++) {
    for (int x = 0; x < w; x++) {
      size_t o = (size_t)y * w + x;
      real wn;
      if (weights_f) {
        wn = weights_f[o];
        if (!(wn > 0)) continue;
      } else if (mask_u8) {
        if (!mask_u8[o]) continue;
        wn = 1;
      } else {
        wn = 1;
      }
      const real Wn = accw[o] + wn;
      const real factor = weights_f ? wn / Wn : (real)1 / Wn;
      accw[o] = Wn;
      for (int c = 0; c < cn; c++) {
        real A = acc[o * cn + c];
        acc[o * cn + c] = RFMA(src[o * cn + c] - A, factor, A);
      }
    }
  }
}

/* Per-frame sharpness weight: mean of the squared 3x3 laplacian
 * [0 1 0; 1 -4 1; 0 1 0] with BORDER_REPLICATE.  Canonical per pixel:
 * L = fma(-4, c, (up + down) + (left + right)); sum of L*L in double. */
double ora_sharpness_mean_sq_laplacian(const real *src, int w, int h) {
  double *rs = (double *)xcalloc(h, sizeof(double));
#pragma omp parallel for schedule(static)
  for (int y = 0; y < h; y++) {
    const real *up = src + (size_t)(y > 0 ? y - 1 : 0) * w;
    const real *c = src + (size_t)y * w;
    const real *dn = src + (size_t)(y < h - 1 ? y + 1 : h - 1) * w;
    double s = 0;
    for (int x = 0; x < w; x++) {
      real l = c[x > 0 ? x - 1 : 0], r = c[x < w - 1 ? x + 1 : w - 1];
      real L = RFMA((real)-4, c[x], (up[x] + dn[x]) + (l + r));
      s += (double)L * (double)L;
    }
    rs[y] = s;
  }
  double s = 0;
  for (int y = 0; y < h; y++) s += rs[y];
  free(rs);
  return s / ((double)w * h);
}

/* ------------------------------------------------------------------------ */
/* c_bayer_average (c_frame_accumulation.cc:988-1334)                        */
/* ------------------------------------------------------------------------ */
int ora_bayer_channel(int colorid, int x, int y) {
  /* generate_bayer_pattern_mask (:1267-1334): 2x2 cells, B = 0, G = 1, R = 2 */
  static const int tab[4][4] = {
      {2, 1, 1, 0}, /* RGGB: R G / G B */
      {1, 2, 0, 1}, /* GRBG: G R / B G */
      {1, 0, 2, 1}, /* GBRG: G B / R G */
      {0, 1, 1, 2}, /* BGGR: B G / G R */
  };
  return tab[colorid - 8][((y & 1) << 1) | (x & 1)];
}

void ora_bayer_accumulate(const real *src, int w, int h, const real *rmap, int colorid,
                          const uint8_t *mask_u8, const real *weights_f, real *acc, real *cntr,
                          int32_t *counts) {
#pragma omp parallel for schedule(static)
  for (int y = 0; y < h; y++) {
    for (int x = 0; x < w; x++) {
      const size_t o = (size_t)y * w + x;
      if (mask_u8 && !mask_u8[o]) continue;
      if (!rmap) {
        /* :996-1039: the pixel's own channel; float arithmetic */
        const int cc = ora_bayer_channel(colorid, x, y);
        const real wt = weights_f ? weights_f[o] : (real)1;
        acc[o * 3 + cc] += weights_f ? src[o] * wt : src[o];
        cntr[o * 3 + cc] += wt;
        counts[o * 3 + cc] += wt != 0;
        continue;
      }
      /* :1044-1081: p = rmap(x, y); (int) truncation; side weights in the map's
       * precision, products in double, float accumulators */
      const real px = rmap[2 * o], py = rmap[2 * o + 1];
      const int sx = (int)px, sy = (int)py;
      if (!(sx >= 0 && sx < w - 1 && sy >= 0 && sy < h - 1)) continue;
      const double wt = weights_f ? (double)weights_f[o] : 1.0;
      const double ax = (double)((real)(sx + 1) - px), ay = (double)((real)(sy + 1) - py);
      const double bx = (double)(px - (real)sx), by = (double)(py - (real)sy);
      const double s[4] = {ax * ay * wt, bx * ay * wt, ax * by * wt, bx * by * wt};
      for (int k = 0; k < 4; k++) {
        const int xx = sx + (k & 1), yy = sy + (k >> 1);
        const int c = ora_bayer_channel(colorid, xx, yy);
        const double v = (double)src[(size_t)yy * w + xx];
        acc[o * 3 + c] = (real)((double)acc[o * 3 + c] + v * s[k]);
        cntr[o * 3 + c] = (real)((double)cntr[o * 3 + c] + s[k]);
        counts[o * 3 + c] += s[k] != 0.0;
      }
    }
  }
}

void ora_bayer_compute(const real *acc, const real *cntr, int w, int h, real *avg, uint8_t *mask) {
  for (size_t o = 0; o < (size_t)w * h; o++) {
    int any = 0;
    for (int c = 0; c < 3; c++) {
      const real n = cntr[o * 3 + c];
      avg[o * 3 + c] = n > 0 ? acc[o * 3 + c] / n : (real)0;
      any |= n > 0;
    }
    if (mask) mask[o] = any ? 255 : 0;
  }
}

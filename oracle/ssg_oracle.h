/*
 * ssg_oracle.h -- CPU oracle for the align+stack hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: only tests/, the
 * smoke() entry point and bench.py's cpu_baseline leg may load it.  The
 * product path (serstacker_amd/) never links, imports or calls it.
 *
 * It restates, in plain C, the reference algorithm (SerStacker core/proc/
 * image_registration and core/average) together with the OpenCV 4.5.x
 * primitive semantics the reference relies on (remap with 1/32-px fixed-point
 * positions, pyrDown, sepFilter2D, meanStdDev, erode, Cholesky invert).
 *
 * The same source is compiled twice:
 *   liboracle_f32.so : SSG_REAL=float  -- the FP32 "reference-faithful" path
 *   liboracle_f64.so : SSG_REAL=double -- the FP64 oracle for warp accuracy
 *
 * Canonical arithmetic.  Where OpenCV's result depends on the compiler's
 * contraction of a*b+c*d into fused multiply-adds, the oracle fixes one order
 * (explicit fma calls, -ffp-contract=off).  The HIP kernels use the same
 * order, so per-pixel image operations agree bit for bit.
 */
#ifndef SSG_ORACLE_H
#define SSG_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#ifndef SSG_REAL
#define SSG_REAL float
#endif
typedef SSG_REAL ora_real;

/* values follow OpenCV's BorderTypes / InterpolationFlags */
enum { ORA_BORDER_CONSTANT = 0, ORA_BORDER_REPLICATE = 1, ORA_BORDER_REFLECT = 2,
       ORA_BORDER_WRAP = 3, ORA_BORDER_REFLECT101 = 4, ORA_BORDER_TRANSPARENT = 5 };
enum { ORA_INTER_NEAREST = 0, ORA_INTER_LINEAR = 1, ORA_INTER_CUBIC = 2, ORA_INTER_AREA = 3 };
/* values follow the reference's IMAGE_MOTION_TYPE (image_transform.h:15-26) */
enum { ORA_MOTION_TRANSLATION = 0, ORA_MOTION_EUCLIDEAN = 1, ORA_MOTION_SCALED_EUCLIDEAN = 2,
       ORA_MOTION_AFFINE = 3 };
/* values follow the reference's ECC_ALIGN_METHOD (ecc2.h:59-64) */
enum { ORA_ECC_FORWARD_ADDITIVE = 0, ORA_ECC_INVERSE_COMPOSITIONAL = 1, ORA_ECC_LM = 2,
       ORA_ECC_INVERSE_COMPOSITIONAL_LM = 3 };

int  ora_real_size(void);
void ora_set_threads(int n);

/* ---- OpenCV primitive restatements ---- */
int  ora_border_interpolate(int p, int len, int border);
void ora_gaussian_kernel(int n, double sigma, ora_real *out);
void ora_sep_filter(const ora_real *src, int w, int h, ora_real *dst,
                    const ora_real *kx, int nx, const ora_real *ky, int ny,
                    int border, ora_real delta);
void ora_pyr_down(const ora_real *src, int sw, int sh, ora_real *dst, int dw, int dh, int border);
void ora_pyr_down_u8(const uint8_t *src, int sw, int sh, uint8_t *dst, int dw, int dh);
void ora_pyr_up(const ora_real *src, int sw, int sh, ora_real *dst, int dw, int dh);
void ora_resize_nearest_u8(const uint8_t *src, int sw, int sh, uint8_t *dst, int dw, int dh);
void ora_remap(const ora_real *src, int sw, int sh, const ora_real *map, int dw, int dh,
               ora_real *dst, int interp, int border, ora_real cval);
void ora_remap_u8(const uint8_t *src, int sw, int sh, const ora_real *map, int dw, int dh,
                  uint8_t *dst, int border, uint8_t cval);
void ora_remap_u8_cubic(const uint8_t *src, int sw, int sh, const ora_real *map, int dw, int dh,
                        uint8_t *dst, int border, uint8_t cval);
void ora_bicubic_itab(int fx, int fy, int *w);
void ora_erode_rect_u8(const uint8_t *src, int w, int h, int kw, int kh, uint8_t *dst,
                       int border, uint8_t cval);
void ora_mean_stddev(const ora_real *src, const uint8_t *mask, int n, double *mean, double *stdev);
int  ora_cholesky_invert(ora_real *A, int n);  /* in place; 0 on a non-positive pivot */
void ora_ecc_differentiate(const ora_real *src, int w, int h, ora_real *gx, ora_real *gy);

/* ---- image transforms (c_image_transform.cc) ---- */
typedef struct ora_xform {
  int32_t  motion;
  int32_t  nparams;
  ora_real p[6];        /* parameter vector in the reference's layout */
  ora_real C[2];        /* euclidean rotation center (reference default 0,0) */
  ora_real fixed_scale; /* euclidean _scale used when the scale is fixed */
} ora_xform;

void     ora_xform_reset(ora_xform *t, int motion);
void     ora_xform_create_map(const ora_xform *t, int w, int h, ora_real *map);
void     ora_xform_scale(ora_xform *t, double f);
ora_real ora_xform_eps(const ora_xform *t, const ora_real *dp, int w, int h);
void     ora_xform_translation(const ora_xform *t, ora_real T[2]);
void     ora_xform_set_translation(ora_xform *t, const ora_real T[2]);
void     ora_xform_sd_images(const ora_xform *t, const ora_real *gx, const ora_real *gy,
                             int w, int h, ora_real **J);
void     ora_xform_matrix(const ora_xform *t, double m[6]);
/* invert_and_compose(p, dp) of the inverse-compositional aligners */
void     ora_xform_invert_and_compose(const ora_xform *t, const ora_real *dp, ora_real *out);

/* ---- ECC (ecc2.cc) ---- */
typedef struct ora_ecc_level_result {
  int32_t num_iterations;
  int32_t failed;
  double  eps;
} ora_ecc_level_result;

int ora_ecc_fa_align(const ora_real *ref, const uint8_t *ref_mask, int w, int h,
                     const ora_real *cur, const uint8_t *cur_mask, int cw, int ch,
                     ora_xform *t, int interp, int max_iterations, double max_eps,
                     double update_step_scale, ora_ecc_level_result *res);

/* c_ecclm::align (ecc2.cc:1526-1650); ref_mask already eroded */
int ora_ecc_lm_align(const ora_real *ref, const uint8_t *ref_mask, int w, int h,
                     const ora_real *cur, const uint8_t *cur_mask, int cw, int ch,
                     ora_xform *t, int max_iterations, double max_eps,
                     double update_step_scale, ora_ecc_level_result *res);
/* c_ecc_inverse_compositional::align (ecc2.cc:1690-1792); J from jx */
int ora_ecc_ic_align(const ora_real *ref, const uint8_t *ref_mask, int w, int h,
                     const ora_real *cur, const uint8_t *cur_mask, int cw, int ch,
                     ora_xform *t, const ora_xform *jx, int max_iterations, double max_eps,
                     double update_step_scale, ora_ecc_level_result *res);
/* c_ecclm_inverse_compositional::align (ecc2.cc:1926-2086); J from jx */
int ora_ecc_iclm_align(const ora_real *ref, const uint8_t *ref_mask, int w, int h,
                       const ora_real *cur, const uint8_t *cur_mask, int cw, int ch,
                       ora_xform *t, const ora_xform *jx, int max_iterations, double max_eps,
                       double update_step_scale, ora_ecc_level_result *res);
double ora_compute_correlation(const ora_real *cur, const uint8_t *cur_mask, int cw, int ch,
                               const ora_real *ref, const uint8_t *ref_mask, int w, int h,
                               const ora_real *map);

/* ---- frame registration (c_frame_registration.cc) ---- */
typedef struct ora_reg_options {
  int32_t motion_type;
  int32_t interpolation;
  int32_t border_mode;
  double  border_value;
  double  scale;
  double  eps;
  double  min_rho;
  double  input_smooth_sigma;
  double  reference_smooth_sigma;
  double  update_step_scale;
  int32_t max_iterations;
  int32_t ecch_max_level;
  int32_t ecch_minimum_image_size;
  int32_t ecch_estimate_translation_first;
  int32_t ecc_method;
} ora_reg_options;

typedef struct ora_reg_status {
  int32_t ok;          /* 1 registered, 0 rejected */
  int32_t code;        /* 0 ok, 3 rho (translation pass), 4 rho (final) */
  double  rho_translation;
  double  rho;
  int32_t iterations_translation;
  int32_t iterations;
  int32_t levels_failed;
  int32_t nparams;
  double  params[6];   /* full-resolution parameter vector */
  double  matrix[6];   /* equivalent x' = m0 x + m1 y + m2, y' = m3 x + m4 y + m5 */
  int32_t level_iterations[16];
} ora_reg_status;

typedef struct ora_reg ora_reg;
ora_reg *ora_reg_create(const ora_reg_options *opts);
void     ora_reg_destroy(ora_reg *r);
int      ora_reg_setup_reference(ora_reg *r, const ora_real *img, const uint8_t *mask, int w, int h);
int      ora_reg_register(ora_reg *r, const ora_real *img, const uint8_t *mask, int w, int h,
                          ora_reg_status *st);
int      ora_reg_num_levels(const ora_reg *r);
int      ora_reg_level_size(const ora_reg *r, int lvl, int *w, int *h);
int      ora_reg_get_ref_level(const ora_reg *r, int lvl, ora_real *img, uint8_t *mask);

/* c_frame_registration::base_remap (c_frame_registration.cc:1265-1338) */
void ora_base_remap(const ora_real *src, int sw, int sh, int cn, const uint8_t *src_mask,
                    const ora_real *map, int dw, int dh, ora_real *dst, uint8_t *dst_mask,
                    int interp, int border, ora_real cval);

/* ---- accumulation (c_frame_accumulation.cc) ---- */
void ora_weighted_average_update(const ora_real *src, int w, int h, int cn,
                                 const uint8_t *mask_u8, const ora_real *weights_f,
                                 ora_real *acc, ora_real *accw);
double ora_sharpness_mean_sq_laplacian(const ora_real *src, int w, int h);

/* c_bayer_average (c_frame_accumulation.cc:988-1334).  Channels in the
 * reference's BGR order (BAYER_B = 0, BAYER_G = 1, BAYER_R = 2,
 * c_frame_accumulation.h:230-234); colorid = COLORID_BAYER_RGGB (8), GRBG (9),
 * GBRG (10), BGGR (11) (core/io/debayer.h:22-25). */
int ora_bayer_channel(int colorid, int x, int y);
/* one add(src, weights) with the remap set by set_remap (rmap may be NULL):
 * acc, cntr: w*h*3 (the reference's Mat3f accumulator and counter); counts:
 * w*h*3 taps with a non-zero weight (integer per-cell sample counts). */
void ora_bayer_accumulate(const ora_real *src, int w, int h, const ora_real *rmap, int colorid,
                          const uint8_t *mask_u8, const ora_real *weights_f, ora_real *acc,
                          ora_real *cntr, int32_t *counts);
/* compute(): avg = acc / cntr per channel (0 where cntr <= 0); mask = any
 * channel's cntr > 0 (255). */
void ora_bayer_compute(const ora_real *acc, const ora_real *cntr, int w, int h, ora_real *avg,
                       uint8_t *mask);

#ifdef __cplusplus
}
#endif
#endif

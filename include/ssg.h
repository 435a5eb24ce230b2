/*
 * ssg.h -- C ABI of the MI355X (gfx950) align+stack hot path.
 *
 * The entry points replace SerStacker's registration and frame-accumulation
 * C++ interfaces for the per-frame hot path.  Each declaration names the
 * reference interface it stands for (paths relative to the SerStacker tree).
 *
 * Conventions
 *  - Plain pointers and sizes only.  Images are described by ssg_image
 *    (OpenCV-like type codes, byte stride, and where the memory lives).
 *  - The caller owns every buffer it passes; the library never frees one.
 *    Opaque handles own their device state and are released by *_destroy.
 *  - Every call returns 0 on success or a negative SSG_E_* code; the
 *    message of the last failure on the calling thread is ssg_last_error().
 *  - Handles are single-owner and not thread-safe (as the reference classes).
 *    Each handle runs its work on its own HIP stream and calls return after
 *    the stream has finished, unless documented otherwise.
 */
#ifndef SSG_H
#define SSG_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SSG_ABI_VERSION 1

/* ---- status codes ---------------------------------------------------------------------------- */
enum {
  SSG_OK = 0,
  SSG_E_INVALID_ARG = -1,
  SSG_E_HIP = -2,
  SSG_E_NOT_READY = -3, /* e.g. no reference frame set */
  SSG_E_UNSUPPORTED = -4,
  SSG_E_NO_MEMORY = -5
};

/* per-frame registration outcome (ssg_registration_status.code) */
enum {
  SSG_REG_OK = 0,
  SSG_REG_NOT_RUN = 1,
  SSG_REG_LEVEL_FAILED = 2,     /* informational: a pyramid level failed its Cholesky, frame kept */
  SSG_REG_RHO_TRANSLATION = 3,  /* c_frame_registration.cc:842-846  rho < 0.75*min_rho */
  SSG_REG_RHO = 4               /* c_frame_registration.cc:865-868  rho < min_rho */
};

/* ---- enums (values = the reference's / OpenCV's) ------------------------------------------- */
/* IMAGE_MOTION_TYPE, core/proc/image_registration/image_transform.h:15-26 */
enum { SSG_MOTION_TRANSLATION = 0, SSG_MOTION_EUCLIDEAN = 1, SSG_MOTION_SCALED_EUCLIDEAN = 2,
       SSG_MOTION_AFFINE = 3 };
/* ECC_ALIGN_METHOD, ecc2.h:59-64 */
enum { SSG_ECC_ALIGN_FORWARD_ADDITIVE = 0, SSG_ECC_ALIGN_INVERSE_COMPOSITIONAL = 1,
       SSG_ECC_ALIGN_LM = 2, SSG_ECC_ALIGN_INVERSE_COMPOSITIONAL_LM = 3 };
/* ECC_INTERPOLATION_METHOD, ecc2.h:32-43 (cv::InterpolationFlags) */
enum { SSG_INTER_NEAREST = 0, SSG_INTER_LINEAR = 1, SSG_INTER_CUBIC = 2, SSG_INTER_AREA = 3 };
/* ECC_BORDER_MODE, ecc2.h:46-56 (cv::BorderTypes) */
enum { SSG_BORDER_CONSTANT = 0, SSG_BORDER_REPLICATE = 1, SSG_BORDER_REFLECT = 2, SSG_BORDER_WRAP = 3,
       SSG_BORDER_REFLECT101 = 4, SSG_BORDER_TRANSPARENT = 5 };
/* image element types (OpenCV type codes) */
enum { SSG_8UC1 = 0, SSG_16UC1 = 2, SSG_32SC1 = 4, SSG_32FC1 = 5, SSG_64FC1 = 6, SSG_32FC2 = 13, SSG_8UC3 = 16,
       SSG_16UC3 = 18, SSG_32SC3 = 20, SSG_32FC3 = 21, SSG_64FC3 = 22 };
/* COLORID, core/io/debayer.h:19-35 */
enum { SSG_COLORID_MONO = 0, SSG_COLORID_BAYER_RGGB = 8, SSG_COLORID_BAYER_GRBG = 9, SSG_COLORID_BAYER_GBRG = 10,
       SSG_COLORID_BAYER_BGGR = 11, SSG_COLORID_BAYER_CYYM = 16, SSG_COLORID_BAYER_YCMY = 17,
       SSG_COLORID_BAYER_YMCY = 18, SSG_COLORID_BAYER_MYYC = 19, SSG_COLORID_RGB = 100, SSG_COLORID_BGR = 101 };
/* DEBAYER_ALGORITHM, core/io/debayer.h:38-56 */
enum { SSG_DEBAYER_DEFAULT = -1, SSG_DEBAYER_DISABLE = 0, SSG_DEBAYER_NN = 1, SSG_DEBAYER_NN2 = 2,
       SSG_DEBAYER_VNG = 3, SSG_DEBAYER_EA = 4, SSG_DEBAYER_SP = 5, SSG_DEBAYER_MATRIX = 6,
       SSG_DEBAYER_PLANE_0 = 7, SSG_DEBAYER_PLANE_1 = 8, SSG_DEBAYER_PLANE_2 = 9, SSG_DEBAYER_PLANE_3 = 10,
       SSG_DEBAYER_AVGBP = 11 };
/* frame_accumulation_method, core/pipeline/c_image_stacking_pipeline/c_image_stacking_pipeline.h:26-34 */
enum { SSG_ACC_AVERAGE = 0, SSG_ACC_WEIGHTED_AVERAGE = 1, SSG_ACC_FOCUS_STACK = 2, SSG_ACC_FFT = 3,
       SSG_ACC_BAYER_AVERAGE = 4 };

/* ---- image view ---------------------------------------------------------------------------- */
typedef struct ssg_image {
  void *data;
  int32_t width;
  int32_t height;
  int64_t step;      /* bytes between rows */
  int32_t type;      /* SSG_8UC1, SSG_32FC1, ... */
  int32_t on_device; /* 1: data is device memory of the current HIP device; 0: host memory */
} ssg_image;

/* ---- options (field names, order and defaults mirror the reference structs) ------------------ */
/* c_ecc_registration_options, c_frame_registration.h:47-64 */
typedef struct ssg_ecc_registration_options {
  double scale;
  double eps;
  double min_rho;
  double input_smooth_sigma;
  double reference_smooth_sigma;
  double update_step_scale;
  int32_t se_radius;
  int32_t ecc_method;
  int32_t max_iterations;
  int32_t ecch_max_level;
  int32_t ecch_minimum_image_size;
  double normalization_noise;
  int32_t normalization_scale;
  int32_t ecch_estimate_translation_first;
  int32_t replace_planetary_disk_with_mask;
} ssg_ecc_registration_options;

/* c_eccflow_registration_options, c_frame_registration.h:88-100 */
typedef struct ssg_eccflow_registration_options {
  double update_multiplier;
  double input_smooth_sigma;
  double reference_smooth_sigma;
  double noise_level;
  double scale_factor;
  int32_t max_iterations;
  int32_t support_scale;
  int32_t min_image_size;
  int32_t max_pyramid_level;
  int32_t downscale_method; /* ECCFlowDownscaleMethod, ecc2.h:507-511 */
} ssg_eccflow_registration_options;

/* color_channel_type (core/proc/extract_channel.h:13-50): the channel ECC registers on.  Colour
 * (SSG_32FC3, BGR) frames take GRAY (cv::cvtColor BGR2GRAY, the reference default), BLUE / GREEN
 * / RED or a channel index 0..2; gray frames are used as they are. */
enum { SSG_COLOR_CHANNEL_0 = 0, SSG_COLOR_CHANNEL_1 = 1, SSG_COLOR_CHANNEL_2 = 2, SSG_COLOR_CHANNEL_DONT_CHANGE = 100,
       SSG_COLOR_CHANNEL_RED = 101, SSG_COLOR_CHANNEL_GREEN = 102, SSG_COLOR_CHANNEL_BLUE = 103,
       SSG_COLOR_CHANNEL_GRAY = 104 };

/* c_image_registration_options, c_frame_registration.h:119-136 */
typedef struct ssg_image_registration_options {
  int32_t motion_type;
  int32_t ecc_registration_channel;
  int32_t interpolation;
  int32_t border_mode;
  double border_value[4];
  ssg_ecc_registration_options ecc;
  ssg_eccflow_registration_options eccflow;
  int32_t enable_feature_registration;
  int32_t enable_ecc_registration;
  int32_t enable_eccflow_registration;
  int32_t accumulate_and_compensate_turbulent_flow;
} ssg_image_registration_options;

/* Fills *o with the reference defaults (c_frame_registration.h:47-136). */
void ssg_default_image_registration_options(ssg_image_registration_options *o);

/* c_image_registration_status (c_frame_registration.h:185-208) plus the
 * registered transform.  params[] holds the reference's parameter vector
 * (translation 2, euclidean 3, scaled euclidean 4, affine 6); matrix[] the
 * equivalent 2x3 map x' = m0 x + m1 y + m2, y' = m3 x + m4 y + m5 from the
 * reference frame grid to the current frame, at full resolution. */
typedef struct ssg_registration_status {
  int32_t ok;
  int32_t code;
  double rho_translation;
  double rho;
  double eps;
  int32_t iterations_translation;
  int32_t iterations;
  int32_t levels_failed;
  int32_t nparams;
  float params[6];
  double matrix[6];
  int32_t level_iterations[8];
} ssg_registration_status;

/* ---- library ------------------------------------------------------------------------------- */
int ssg_abi_version(void);
const char *ssg_last_error(void);
/* Number of HIP devices and the name of the current one, "<name> (<gcnArchName>, <n> CUs)";
 * fails if no GPU. */
int ssg_device_info(int *num_devices, char *name, int name_len);

/* ---- registration: c_frame_registration (c_frame_registration.h:210-354) ------------------- */
typedef struct ssg_registration ssg_registration;

/* c_frame_registration::c_frame_registration(options) (c_frame_registration.cc:263-) */
int ssg_registration_create(const ssg_image_registration_options *opts, ssg_registration **out);
void ssg_registration_destroy(ssg_registration *r);

/* c_frame_registration::setup_reference_frame (c_frame_registration.cc:565-719).
 * image: SSG_32FC1 (host or device); mask: SSG_8UC1 or NULL. */
int ssg_registration_setup_reference_frame(ssg_registration *r, const ssg_image *image,
                                           const ssg_image *mask);

/* c_frame_registration::register_frame (c_frame_registration.cc:721-964), the
 * plain per-frame form: the ECC loop of each pyramid level syncs with the host
 * after every iteration, as the reference's loop does (ecc2.cc:1295). */
int ssg_registration_register_frame(ssg_registration *r, const ssg_image *image, const ssg_image *mask,
                                    ssg_registration_status *status);

/* Batched form of register_frame for n frames of the reference size, all
 * resident in device memory: pyramid build, translation-first ECC, rho gates
 * and the main ECC run for every frame at once, with a per-frame convergence
 * mask (no counterpart in the reference, which registers frames one by one;
 * results per frame are those of register_frame).  masks may be NULL. */
int ssg_registration_register_batch(ssg_registration *r, int n, const ssg_image *frames,
                                    const ssg_image *masks, ssg_registration_status *status);

/* c_image_transform::create_remap (c_image_transform.h:54): writes the SSG_32FC2
 * remap field of a status's transform at the reference frame size. */
int ssg_registration_create_remap(ssg_registration *r, const ssg_registration_status *status,
                                  ssg_image *map);

/* c_frame_registration::current_remap (c_frame_registration.h:292): the remap of the last frame
 * passed to ssg_registration_register_frame, at the reference size (SSG_32FC2, host or device).
 * With enable_eccflow_registration it is the ECC transform's map refined by the dense ECC flow
 * (c_frame_registration.cc:883-917), the field c_frame_registration::remap warps with. */
int ssg_registration_current_remap(ssg_registration *r, ssg_image *map);

/* c_frame_registration::remap (c_frame_registration.cc:1405-1417): custom_remap of src (and
 * src_mask) with the current remap of the last frame passed to ssg_registration_register_frame;
 * interpolation / border_mode < 0 take the options' values (and border_value[0]). */
int ssg_registration_remap(ssg_registration *r, const ssg_image *src, const ssg_image *src_mask, ssg_image *dst,
                           ssg_image *dst_mask, int32_t interpolation, int32_t border_mode, double border_value);

/* Batched register_frame that also returns each frame's current remap (no reference
 * counterpart): remaps[i] (device SSG_32FC2, reference size, one step) receives the frame's
 * transform map, refined by the dense ECC flow for the accepted frames when
 * enable_eccflow_registration is set (then frames and masks are device images of one step). */
int ssg_registration_register_batch_remaps(ssg_registration *r, int n, const ssg_image *frames,
                                           const ssg_image *masks, ssg_registration_status *status,
                                           ssg_image *remaps);

/* c_frame_registration::custom_remap / base_remap (c_frame_registration.cc:1265-1417):
 * dst = remap(src, map) with the given interpolation/border; dst_mask (optional)
 * = erode5x5(remap(src_mask or 255, CONSTANT 0) >= 255).  The map is either
 * a SSG_32FC2 field (map != NULL) or the 2x3 matrix of a registration status
 * (map == NULL, matrix != NULL).  src: SSG_32FC1 or SSG_32FC3. */
int ssg_remap(const ssg_image *src, const ssg_image *src_mask, const ssg_image *map, const double *matrix,
              int32_t motion_type, const float *params, ssg_image *dst, ssg_image *dst_mask,
              int32_t interpolation, int32_t border_mode, double border_value);

/* ---- accumulation: c_frame_accumulation (core/average/c_frame_accumulation.h:14-37) --------- */
typedef struct ssg_accumulation ssg_accumulation;

/* create_frame_accumulation (c_image_stacking_pipeline.cc:450-466): method is
 * SSG_ACC_WEIGHTED_AVERAGE or SSG_ACC_AVERAGE (both c_weigthed_average),
 * SSG_ACC_BAYER_AVERAGE (c_bayer_average, c_frame_accumulation.h:222-262; channels = 1,
 * the raw input; the accumulator has the reference's 3 channels, BGR order), or
 * SSG_ACC_FOCUS_STACK (c_laplacian_pyramid_focus_stacking, c_frame_accumulation.h:141-184;
 * frames SSG_32FC1/SSG_32FC3 added without masks; compute() writes no counters and a mask
 * of 255, get_acc_counters() writes nothing, as the reference releases accw). */
int ssg_accumulation_create(int32_t method, int32_t width, int32_t height, int32_t channels,
                            ssg_accumulation **out);
void ssg_accumulation_destroy(ssg_accumulation *a);

/* c_bayer_average::set_bayer_pattern (c_frame_accumulation.cc:1136-1142): SSG_COLORID_BAYER_*. */
int ssg_accumulation_set_bayer_pattern(ssg_accumulation *a, int32_t colorid);
/* c_bayer_average::set_remap (c_frame_accumulation.cc:1149-1152): the SSG_32FC2 field of
 * source positions for each accumulator pixel, used by the following add() calls; NULL
 * clears it (clear() also clears it, :1154-1161). */
int ssg_accumulation_set_remap(ssg_accumulation *a, const ssg_image *map);
/* SSG_16UC1 raw frames are read as value / 2^bpp, the pipeline's conversion
 * (c_image_stacking_pipeline_base.cc:269-273); default 16. */
int ssg_accumulation_set_input_bpp(ssg_accumulation *a, int32_t bpp);

/* c_laplacian_pyramid_focus_stacking::options (c_frame_accumulation.h:155-161) and the weight
 * source of the focus configuration.  SSG_FOCUS_SELECT_MAX_ENERGY and SSG_FOCUS_WEIGHTED_AVERAGE
 * are the reference's fusing policies (c_frame_accumulation.cc:491-681).  SSG_FOCUS_LPG_WEIGHTS
 * (no reference counterpart) weights every laplacian level by W = gaussian(lpg(k, p), lpg_sigma)
 * reduced to that level with pyrDown, on every level including the top one, and divides by the
 * summed weights. */
enum { SSG_FOCUS_SELECT_MAX_ENERGY = 0, SSG_FOCUS_WEIGHTED_AVERAGE = 1, SSG_FOCUS_LPG_WEIGHTS = 2 };
typedef struct ssg_focus_stacking_options {
  int32_t fusing_policy;      /* SSG_FOCUS_*; reference default select_max_energy */
  int32_t inpaint_mask_holes; /* reference default 1; masks are not taken on this path */
  int32_t avgchannel;         /* reference default 1 */
  int32_t kradius;            /* reference default 0 */
  double ksigma;              /* reference default 0 */
  double lpg_k, lpg_p, lpg_sigma; /* SSG_FOCUS_LPG_WEIGHTS: 6, 2, 5 */
  int32_t max_levels;         /* 0: levels while the next level's smaller side is > 8
                                 (laplacian_pyramid.cc:62-69); else at most this many levels */
} ssg_focus_stacking_options;
void ssg_default_focus_stacking_options(ssg_focus_stacking_options *o);
/* Options of a SSG_ACC_FOCUS_STACK accumulator, before its first add(). */
int ssg_accumulation_set_focus_options(ssg_accumulation *a, const ssg_focus_stacking_options *o);

/* lpg() (core/proc/lpg.cc:277-344) with dscale = uscale = 0: the sharpness map
 * alpha*lap5x5^2 + beta*grad5x5^2 + 1e-9 of the channel average of src, raised to p
 * (alpha = 625 k/(k+1), beta = (100/36)^2/(k+1)).  src: SSG_32FC1/SSG_32FC3; dst: SSG_32FC1 of the
 * same size (host or device); sharpness (optional): the map's mean. */
int ssg_lpg(const ssg_image *src, ssg_image *dst, double k, double p, double *sharpness);

/* ecc_normalize(src, mask, dst, lvl) (ecc2.cc:385-397), the local-mean removal that registration
 * applies when ecc.normalization_scale > 0 and ecc.normalization_noise > 0
 * (c_frame_registration.cc:624-628, 807-812): dst = src - the pyrUp chain of
 * pyrDown^lvl(src, BORDER_REPLICATE) back to the source size, then 0 where mask (optional,
 * SSG_8UC1) is not 255.  src / dst: SSG_32FC1 of one size, host or device. */
int ssg_ecc_normalize(const ssg_image *src, const ssg_image *mask, ssg_image *dst, int32_t lvl);

/* compute_local_variance_map (core/proc/sharpness_measure/c_local_variance_sharpness_measure.cc:193-249),
 * the per-pixel weight of weighted-average stacking (c_image_stacking_pipeline.cc:2013-2020), for the
 * gray channel with uscale 0 (the pipeline's settings, c_image_stacking_pipeline.h:94-99: dscale 1,
 * kradius 1): G = morphological gradient of the pyrDown'ed frame, W = |G|_1,
 * Q = 20^3 sum(G^4) / W (summed in double; the reference's float atomic sum varies with its thread
 * schedule), map = G^3 20^3 + 0.05 Q resized INTER_LINEAR to the frame size.  image: SSG_32FC1 or
 * SSG_32FC3 (BGR); map: SSG_32FC1 of the image size or NULL; q: the quality Q or NULL.  A frame
 * without gradient gives Q = 0 and a zero map. */
int ssg_local_variance_map(const ssg_image *image, int32_t kradius, int32_t dscale, ssg_image *map, double *q);

/* c_frame_accumulation::add(src, weights) (c_frame_accumulation.cc:182-221, bayer :1169-1196):
 * weights may be NULL, a SSG_8UC1 mask (nonzero = 1) or a SSG_32FC1 weight map.  Bayer
 * sources are SSG_32FC1 or SSG_16UC1 raw frames of the accumulator size. */
int ssg_accumulation_add(ssg_accumulation *a, const ssg_image *src, const ssg_image *weights);

/* n successive add(frames[i]) calls without weights (no reference counterpart: a batch of the
 * reference's per-frame adds, c_image_stacking_pipeline.cc:1644-1776).  Focus stacks build the
 * laplacian pyramids of a batch together and fuse each level frame after frame per pixel in one
 * launch; the results equal those of the single adds. */
int ssg_accumulation_add_frames(ssg_accumulation *a, int n, const ssg_image *frames);

/* Focus stacking of several independent sequences in one batch (no reference counterpart: one
 * c_laplacian_pyramid_focus_stacking per sequence, c_frame_accumulation.cc:491-683, each fed
 * by its own pipeline).  accs[q] are focus accumulators with the SSG_FOCUS_LPG_WEIGHTS policy,
 * of any sizes and 1 or 3 channels; frames[q][0..nframes[q]) are f32 device frames of that
 * accumulator's size (one row step per sequence); shifts[q] (optional, may be NULL) holds per
 * frame the translation of its registration (tx, ty): the frame is remapped from (x + tx,
 * y + ty) with INTER_LINEAR and BORDER_REFLECT101 before it is added.  Every step of a chunk of
 * frames is one launch over all sequences; the sums equal those of one add() per aligned frame,
 * bit for bit.  At most 32 sequences per call; the accumulators' profiling counters are those of
 * accs[0] (SSG_PROF_FOCUS_*). */
int ssg_focus_stack_batch(int32_t nseq, ssg_accumulation *const *accs, const int32_t *nframes,
                          const ssg_image *const *frames, const float *const *shifts);

/* Fused remap + base_remap mask + weighted add of n registered frames (the
 * per-frame sequence custom_remap -> weights*mask/255 -> add of
 * c_image_stacking_pipeline.cc:1644-1776 for every frame at once).
 * statuses[i].ok == 0 skips frame i.  frame_weights: per-frame scalar weight
 * or NULL (1), remapped as a constant image; weight_maps: n per-frame weight images on
 * the frames' own grid (SSG_32FC1 device images, e.g. ssg_local_variance_map of the
 * unaligned frames, c_image_stacking_pipeline.cc:1466-1472) remapped with the frame's
 * interpolation and BORDER_CONSTANT 0 (:1653-1660), or NULL; the weight is then gated
 * by the mask (:1704-1712).  Frames: SSG_32FC1/SSG_32FC3 device images.
 * Bayer accumulators: the pipeline's set_remap(current_remap) + add(raw,
 * current_mask) (:1747-1748) for every frame, frames in order: raw SSG_32FC1 or
 * SSG_16UC1 device frames, the remap of each status's transform, the base_remap
 * mask (LINEAR only), frame_weights and weight_maps NULL. */
int ssg_accumulation_add_registered_batch(ssg_accumulation *a, int n, const ssg_image *frames,
                                          const ssg_registration_status *statuses, int32_t motion_type,
                                          const float *frame_weights, const ssg_image *weight_maps,
                                          int32_t interpolation, int32_t border_mode, double border_value);

/* The per-frame pipeline of a batch with every per-frame decision on the device (no reference
 * counterpart: c_image_stacking_pipeline.cc:1466-1790 runs it frame by frame): the weight of each
 * unaligned frame (the ssg_frame_sharpness measure), the registration of register_batch, and the
 * fused warp + mask + weighted accumulate (add_registered_batch with those weights) of the frames
 * that pass the rho gates -- with no host round trip between them: each pyramid level queues its
 * whole pass budget and the per-frame convergence mask ends every frame's iterations on the
 * device.  use_graph != 0 captures that fixed launch sequence into a hipGraph on the first call
 * and replays it while the handles, frame count and frame buffers stay the same (one graph
 * launch per batch).  frames: SSG_32FC1 device frames of the reference size and one step;
 * `a`: a 1-channel weighted-average accumulator of the reference size.  Statuses as
 * register_batch.  Not with masks, normalization_scale > 0 or the eccflow refine. */
int ssg_stack_batch(ssg_registration *r, ssg_accumulation *a, int32_t n, const ssg_image *frames,
                    int32_t interpolation, int32_t border_mode, double border_value, int32_t use_graph,
                    ssg_registration_status *status);

/* The same fused remap -> mask -> weight -> accumulate with a dense remap field per frame
 * instead of a transform: the frames warped by c_frame_registration::remap with the
 * eccflow-refined current remap (c_frame_registration.cc:1405-1417), then accumulated
 * (c_image_stacking_pipeline.cc:1644-1776).  rmaps: n device SSG_32FC2 fields of the
 * accumulator size and one step; every frame counts (weights <= 0 excepted). */
int ssg_accumulation_add_remapped_batch(ssg_accumulation *a, int n, const ssg_image *frames, const ssg_image *rmaps,
                                        const float *frame_weights, const ssg_image *weight_maps,
                                        int32_t interpolation, int32_t border_mode, double border_value);

/* c_frame_accumulation::compute(avg, mask, dscale, ddepth) (c_frame_accumulation.cc:223-248):
 * avg = sum / weight (FP64 division), mask = weight > 0 (SSG_8UC1, optional).
 * Bayer (:1198-1238): avg SSG_32FC3 = acc / counter per channel (float), mask = any
 * channel's counter > 0. */
int ssg_accumulation_compute(ssg_accumulation *a, ssg_image *avg, ssg_image *mask, double dscale);

/* c_frame_accumulation::get_acc_counters (c_frame_accumulation.cc:250-260): the weight plane
 * (SSG_32FC1); Bayer (:1240-1250): the counter times (1, 0.5, 1), SSG_32FC3. */
int ssg_accumulation_get_acc_counters(ssg_accumulation *a, ssg_image *accw);

/* Integer per-pixel sample counts (frames whose mask covered the pixel), SSG_32SC1; Bayer:
 * per pixel and channel the taps added with a non-zero weight, SSG_32SC3. */
int ssg_accumulation_get_counts(ssg_accumulation *a, ssg_image *counts);

/* c_frame_accumulation::accumulated_frames (c_frame_accumulation.h:31). */
int ssg_accumulation_accumulated_frames(ssg_accumulation *a, int32_t *n);

/* c_frame_accumulation::reinitialize(src, accw) (c_frame_accumulation.h:25; c_weigthed_average
 * :171-180): the accumulator restarts from a mean image `src` (f32, the accumulator's channels)
 * with weights `accw` (SSG_32FC1), as one accumulated frame.  The sum-form planes hold src*accw
 * exactly, so compute() returns src where accw > 0 (where accw == 0 it returns 0, the empty
 * pixel of the sum form); per-pixel counts restart at 1 where accw > 0.  Bayer and focus
 * accumulators return SSG_E_UNSUPPORTED, as the reference's return false (:486-490,1170-1173). */
int ssg_accumulation_reinitialize(ssg_accumulation *a, const ssg_image *src, const ssg_image *accw);

/* Accumulator checkpoint / resume (no reference counterpart: the reference keeps its state in
 * cv::Mat members).  state_size gives the bytes of a state blob; state_save writes it into a host
 * buffer; state_load restores it into an accumulator created with the same method, size,
 * channels (and Bayer pattern / focus options), after which add() continues the stack and
 * compute() returns what the saved accumulator would have: the sums are additive, so a stack
 * split into sessions (or shards) resumes bit for bit.  The blob is a versioned header followed
 * by the raw accumulator planes (host byte order). */
int ssg_accumulation_state_size(ssg_accumulation *a, uint64_t *bytes);
int ssg_accumulation_state_save(ssg_accumulation *a, void *buf, uint64_t bytes);
int ssg_accumulation_state_load(ssg_accumulation *a, const void *buf, uint64_t bytes);

/* c_frame_accumulation::clear (c_frame_accumulation.h:30). */
int ssg_accumulation_clear(ssg_accumulation *a);

/* Per-frame scalar weight of the unaligned frames, computed where the pipeline
 * computes its weights (c_image_stacking_pipeline.cc:1466-1467): the mean of the
 * squared 3x3 laplacian [0 1 0; 1 -4 1; 0 1 0] with BORDER_REPLICATE.  frames:
 * SSG_32FC1 device images of one size; weights: n floats (host). */
int ssg_frame_sharpness(int n, const ssg_image *frames, float *weights);

/* Allocation-free form of ssg_frame_sharpness, same arguments and the same results bit for bit.
 * No counterpart in the reference: it is for callers that cannot allocate device memory
 * (code under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).  It calls
 * no hipMalloc/hipFree and creates no stream: the pointer table, the per-block partial sums
 * and the weights live in a static device array of SSG_SHARPNESS_NOALLOC_SCRATCH_BYTES per GPU,
 * held until the process ends, and the work runs on the null stream of the current device.
 * Limits: n <= SSG_SHARPNESS_NOALLOC_MAX_FRAMES and 24576 + 8 * n * ceil(height / 16) bytes
 * <= SSG_SHARPNESS_NOALLOC_SCRATCH_BYTES, else SSG_E_INVALID_ARG before any GPU work.  Calls
 * from threads of one process are serialised; calls that overlap on one GPU from other
 * processes, or GPU work that writes the scratch array concurrently, race. */
#define SSG_SHARPNESS_NOALLOC_SCRATCH_BYTES (2 * 1024 * 1024)
#define SSG_SHARPNESS_NOALLOC_MAX_FRAMES 2048
int ssg_frame_sharpness_noalloc(int n, const ssg_image *frames, float *weights);

/* ---- FFT phase correlation (no counterpart in the reference, which has none on this path;
 * its nearest code is core/proc/fft.cc:17-160).  Optional 2x2 binning (the luminance of raw
 * Bayer frames), mean removal, separable Hann apodisation, zero padding to powers of two, normalised
 * cross-power spectrum weighted so the correlation peak is a gaussian of sigma 1 px, integer peak
 * and 3-point gaussian refinement per axis (FP64).
 * Estimates T with ref(x) = frame(x + T): the parameters of a translation transform. ------ */
typedef struct ssg_phasecorr ssg_phasecorr;
/* width/height: frame size; binning: 1, or 2 (2x2 cell means).  The padded binned size must be
 * at most 4096 per axis. */
int ssg_phasecorr_create(int32_t width, int32_t height, int32_t binning, ssg_phasecorr **out);
void ssg_phasecorr_destroy(ssg_phasecorr *pc);
/* Reference frame: SSG_32FC1 or SSG_16UC1, host or device. */
int ssg_phasecorr_set_reference(ssg_phasecorr *pc, const ssg_image *ref);
/* n device frames (SSG_32FC1 or SSG_16UC1, one size and step): out[3i] = Tx, out[3i+1] = Ty,
 * out[3i+2] = correlation peak (1 for an exact circular shift). */
int ssg_phasecorr_estimate(ssg_phasecorr *pc, int n, const ssg_image *frames, double *out);

/* ---- dense ECC flow: c_eccflow (core/proc/image_registration/ecc2.h:548-662, ecc2.cc:2103-2865).
 * Coarse-to-fine smooth optical flow: per pyramid level and iteration the current level is
 * warped by the flow, the residual products It*Ix, It*Iy are averaged over cells of
 * 2^support_scale px (INTER_AREA + 3x3 gaussian), a 2x2 system is solved per cell and the
 * update is upsampled (INTER_CUBIC) onto the flow. ----------------------------------------- */
/* ECCFlowDownscaleMethod, ecc2.h:507-511 */
enum { SSG_ECCFLOW_DOWNSCALE_RECURSIVE_RESIZE = 0, SSG_ECCFLOW_DOWNSCALE_FULL_RESIZE = 1,
       SSG_ECCFLOW_DOWNSCALE_PYRAMID = 2 };
/* c_eccflow_options, ecc2.h:515-527 (same fields, order and defaults) */
typedef struct ssg_eccflow_options {
  double input_smooth_sigma;     /* carried, unused (as in c_eccflow) */
  double reference_smooth_sigma; /* carried, unused (as in c_eccflow) */
  double update_multiplier;
  double scale_factor;           /* RecursiveResize / FullResize level step */
  double noise_level;            /* < 0: 1e-3 */
  int32_t max_iterations;
  int32_t support_scale;
  int32_t min_image_size;
  int32_t max_pyramid_level;     /* < 0: no limit */
  int32_t downscale;             /* SSG_ECCFLOW_DOWNSCALE_* */
} ssg_eccflow_options;
typedef struct ssg_eccflow ssg_eccflow;
void ssg_default_eccflow_options(ssg_eccflow_options *o);
int ssg_eccflow_create(const ssg_eccflow_options *o, ssg_eccflow **out);
void ssg_eccflow_destroy(ssg_eccflow *f);
/* c_eccflow::set_reference_image (ecc2.cc:2489-2672).  image: SSG_32FC1, mask: SSG_8UC1 or
 * NULL; host or device. */
int ssg_eccflow_set_reference_image(ssg_eccflow *f, const ssg_image *image, const ssg_image *mask);
/* Pyramid geometry of the reference (c_eccflow::current_pyramid, ecc2.cc:2215): the number of
 * levels, and a level's size. */
int ssg_eccflow_num_levels(const ssg_eccflow *f, int32_t *n);
int ssg_eccflow_level_size(const ssg_eccflow *f, int32_t level, int32_t *width, int32_t *height);
/* Reference pyramid entry planes (pyramid_entry, ecc2.h:555-560): which = 0 image, 1 Ix, 2 Iy
 * (level size), 3 a00, 4 a01, 5 a11, 6 idet (the D cells, cell size).  out: SSG_32FC1 of that
 * size, host or device. */
int ssg_eccflow_reference_plane(const ssg_eccflow *f, int32_t level, int32_t which, ssg_image *out);
/* c_eccflow::compute(input, rmap, mask) (ecc2.cc:2856-2865): input SSG_32FC1 at the reference
 * size; input_mask SSG_8UC1 or NULL; rmap_in the initial remap (SSG_32FC2 at the reference
 * size) or NULL for the zero flow; rmap_out receives identity + flow (SSG_32FC2, may be the
 * same memory as rmap_in).  Host or device images. */
int ssg_eccflow_compute(ssg_eccflow *f, const ssg_image *input, const ssg_image *input_mask,
                        const ssg_image *rmap_in, ssg_image *rmap_out);
/* Batched compute over n device frames of the reference size (no reference counterpart):
 * input_masks and rmaps_in are NULL or n entries; rmaps_out has n device SSG_32FC2 maps. */
int ssg_eccflow_compute_batch(ssg_eccflow *f, int n, const ssg_image *inputs, const ssg_image *input_masks,
                              const ssg_image *rmaps_in, ssg_image *rmaps_out);

/* ---- kernel timing (no reference counterpart; the reference's INSTRUMENT_REGION timers,
 * core/debug.h:183-188, are its closest analogue).  When enabled, every launch of a hot
 * kernel is bracketed by HIP events on the handle's stream; algorithmic bytes follow
 * the per-pixel minimum traffic of each kernel. ------------------------------------------ */
enum { SSG_PROF_ECC_TRANSLATION = 0, SSG_PROF_ECC_MAIN = 1, SSG_PROF_CORRELATION = 2, SSG_PROF_PYRAMID = 3,
       SSG_PROF_ACCUMULATE = 4, SSG_PROF_ECC_SOLVE = 5, SSG_PROF_FFT = 6, SSG_PROF_ECCFLOW = 7,
       SSG_PROF_FOCUS_PREP = 8, SSG_PROF_FOCUS_BLUR = 9, SSG_PROF_FOCUS_DOWN = 10, SSG_PROF_FOCUS_FUSE = 11,
       SSG_PROF_GRAPH = 12, SSG_PROF_COUNT = 13 };
typedef struct ssg_profile {
  int64_t launches[SSG_PROF_COUNT];
  double ms[SSG_PROF_COUNT];     /* summed kernel durations */
  double bytes[SSG_PROF_COUNT];  /* summed algorithmic bytes */
} ssg_profile;
int ssg_registration_set_profiling(ssg_registration *r, int enable);
int ssg_registration_get_profile(ssg_registration *r, ssg_profile *out, int reset);
int ssg_accumulation_set_profiling(ssg_accumulation *a, int enable);
int ssg_accumulation_get_profile(ssg_accumulation *a, ssg_profile *out, int reset);
int ssg_phasecorr_set_profiling(ssg_phasecorr *pc, int enable);
int ssg_phasecorr_get_profile(ssg_phasecorr *pc, ssg_profile *out, int reset);
int ssg_eccflow_set_profiling(ssg_eccflow *f, int enable);
int ssg_eccflow_get_profile(ssg_eccflow *f, ssg_profile *out, int reset);

/* ---- tuning knobs (no reference counterpart) -------------------------------------------------
 * Launch-shape choices of the kernels, process-wide, set by the project's autotuner
 * (serstacker_amd.autotune) per size class, or by hand.  ACC_ROWS, ECC_GRID, FOCUS_CHUNK and
 * PYR_STRIP change speed only (every value adds the same values in the same order); ACC_BLOCKS
 * regroups a batch's frames into chunks whose compensated partial sums are then merged, so the
 * stack agrees to the precision of those compensated sums rather than bit for bit (counts stay
 * exact).
 *   SSG_TUNE_ACC_ROWS     rows per thread of the fused warp+accumulate tile: 2 (64x8) or 4 (64x16)
 *   SSG_TUNE_ACC_BLOCKS   blocks the fused accumulate aims at per launch (256..65536): the frames
 *                         per chunk follow (more chunks = more partial planes to merge)
 *   SSG_TUNE_ECC_GRID     blocks per CU of the persistent ECC passes (1..16; capped by occupancy)
 *   SSG_TUNE_FOCUS_CHUNK  frames per sequence and chunk of ssg_focus_stack_batch (1..32)
 *   SSG_TUNE_PYR_STRIP    1 (default): the registration pyramid by column strips, its smoothed
 *                         level 0 and level 1 in one pass; 0: tiled smoothing, then tiled pyrDown */
enum { SSG_TUNE_ACC_ROWS = 0, SSG_TUNE_ACC_BLOCKS = 1, SSG_TUNE_ECC_GRID = 2, SSG_TUNE_FOCUS_CHUNK = 3,
       SSG_TUNE_PYR_STRIP = 4, SSG_TUNE_COUNT = 5 };
int ssg_tune_set(int32_t knob, int32_t value);
int ssg_tune_get(int32_t knob, int32_t *value);

/* ---- synthetic planetary-disk frames (benchmark and test inputs; no reference counterpart) -- */
typedef struct ssg_synth_params {
  int32_t width, height;
  double disk_radius;
  double texture_sigma;     /* gaussian filter sigma of the uniform texture noise */
  double texture_amplitude; /* 0.1 = 10% */
  double limb_darkening;    /* I = 1 - u*(1-mu), u = 0.6 */
  double max_shift;         /* translation U(-max_shift, max_shift) px */
  double max_rotation_deg;  /* rotation U(-r, r) degrees about the centre */
  double max_scale_dev;     /* scale 1 + U(-d, d) */
  double max_shear;         /* shear U(-s, s) */
  double blur_sigma_min, blur_sigma_max; /* per-frame gaussian blur sigma U(min, max); 0 = none */
  double peak_electrons;    /* Poisson photon noise at this peak (0 = no noise) */
  double read_noise;        /* gaussian read noise sigma, DN */
  double output_scale;      /* stored value = DN * output_scale (1/65536 for the f32 configs) */
  uint64_t seed;
} ssg_synth_params;

/* Renders frames [first, first+n) into device images (SSG_32FC1 or SSG_16UC1)
 * and returns each frame's true map (reference grid -> frame) in truth[6*i].
 * Frame index -1 renders the un-warped reference (identity map, mean blur). */
int ssg_synth_render(const ssg_synth_params *p, int first, int n, const ssg_image *frames, double *truth);

/* ---- input stage: SER sequences and raw-frame conversion ------------------------------------ */

/* c_ser_file::file_header (core/io/c_ser_file.h:38-53): the 178-byte SER header.
 * is_little_endian holds the reader's interpretation (c_ser_file.cc:279: the
 * stored field is 0 for little-endian frames). */
#pragma pack(push, 1)
typedef struct ssg_ser_header {
  char file_id[14];
  int32_t luid;
  int32_t color_id;         /* SSG_COLORID_* */
  int32_t is_little_endian;
  int32_t image_width;
  int32_t image_height;
  int32_t bits_per_plane;   /* 1..8 u8, 9..16 u16, 17..32 s32, -32 f32, -64 f64 */
  int32_t frames_count;
  char observer[40];
  char instrument[40];
  char telescope[40];
  uint64_t date_time;
  uint64_t date_time_utc;
} ssg_ser_header;
#pragma pack(pop)

typedef struct ssg_ser_reader ssg_ser_reader;

/* c_ser_reader::open (c_ser_file.cc:246-353): header checks, the trailing timestamp
 * array when the file holds one; the read position is 0. */
int ssg_ser_open(const char *path, ssg_ser_reader **out);
/* c_ser_reader::close */
void ssg_ser_close(ssg_ser_reader *r);
/* c_ser_file::header(); *frame_type = cvtype() as SSG_* (channels 3 for RGB/BGR) */
int ssg_ser_get_header(const ssg_ser_reader *r, ssg_ser_header *out, int32_t *frame_type);
/* c_ser_file::timestamps(): writes min(n, count) values, *count = timestamps in the file (0 or frames_count) */
int ssg_ser_timestamps(const ssg_ser_reader *r, uint64_t *ts, int32_t n, int32_t *count);
/* c_ser_reader::seek / curpos (c_ser_file.cc:355-386) */
int ssg_ser_seek(ssg_ser_reader *r, int32_t frame_index);
int ssg_ser_curpos(const ssg_ser_reader *r, int32_t *pos);
/* c_ser_reader::read (c_ser_file.cc:388-450): the frame at the read position into a host or
 * device image of the file's type and size (byte order as the reference handles it); advances
 * the position. */
int ssg_ser_read(ssg_ser_reader *r, ssg_image *image);
/* n successive read() calls into device images (no reference counterpart: the batch form of
 * the pipeline's frame loop, c_image_stacking_pipeline.cc:1358-1390).  File reads go through two
 * pinned buffers, each frame's host-to-device copy overlapping the next frame's read. */
int ssg_ser_read_batch(ssg_ser_reader *r, int32_t n, ssg_image *images);

/* debayer() (core/io/debayer.cc:1197-1355) on the GPU; src and dst device or host images.
 * src: one channel, SSG_8UC1, SSG_16UC1 or SSG_32FC1, even size for the Bayer methods.  dst has
 * the source depth: 3 channels for NN / EA / NN2 / MATRIX / SP (SP at half size), 1 channel for
 * DISABLE (a copy), PLANE_k and AVGBP (half size).  NN and EA on 8/16-bit sources are OpenCV's
 * cv::demosaicing COLOR_Bayer*2BGR (bilinear) and COLOR_Bayer*2BGR_EA (edge-aware), borders copied
 * from their inner neighbours; VNG on 8-bit sources (OpenCV's VNG) returns SSG_E_UNSUPPORTED; on
 * f32 sources NN / EA / VNG (and VNG on 16-bit) fall back to NN2 as the reference does.  DEFAULT
 * is NN2. */
int ssg_debayer(const ssg_image *src, int32_t colorid, int32_t algorithm, ssg_image *dst);

/* The conversion of c_image_stacking_pipeline_base::read_input_frame
 * (c_image_stacking_pipeline_base.cc:124-279) for n raw frames of one sequence: with a
 * darkbayer or flatbayer (SSG_32FC1 of the raw size, or NULL) the frame is converted to f32
 * (x / 2^bpp), the dark subtracted and the flat divided; Bayer frames are debayered with
 * `algorithm` (in the source type when no dark/flat is set, with the reference's integer
 * rounding); the result is converted to f32 (x / 2^bpp) when not yet.  raw: SSG_8UC1 /
 * SSG_16UC1 / SSG_32FC1 (or 3-channel RGB/BGR frames, converted only); out: device or host
 * SSG_32FC1 / SSG_32FC3 of the debayered size.  Batches of device frames run in one launch. */
int ssg_read_input_frames(int32_t n, const ssg_image *raw, int32_t colorid, int32_t bpp, int32_t algorithm,
                          const ssg_image *darkbayer, const ssg_image *flatbayer, ssg_image *out);

/* ---- master frame (c_image_stacking_pipeline_base.cc:282-404, c_image_stacking_pipeline.cc:1112-1312) */

/* The quality metric of select_master_frame's best-of scan (c_image_stacking_pipeline_base.cc:
 * 350-359): average_bayer_planes of Bayer frames (colorid a Bayer pattern), then
 * compute_local_variance_map's Q (c_local_variance_sharpness_measure.cc:193-249) with
 * kradius / dscale of the options' quality_estimation.  frames: raw SSG_8UC1 / SSG_16UC1 /
 * SSG_32FC1 (or SSG_32FC3 for colour), host or device; quality[i] receives frame i's Q. */
int ssg_master_frame_quality(int32_t n, const ssg_image *frames, int32_t colorid, int32_t kradius, int32_t dscale,
                             double *quality);
/* The scan window of master_frame_best_of_100_in_middle (c_image_stacking_pipeline_base.cc:312-326). */
int ssg_master_scan_window(int32_t size, int32_t max_frames_to_scan, int32_t *start_pos, int32_t *end_pos);
/* The selection rule (c_image_stacking_pipeline_base.cc:369-395): the first frame whose quality
 * exceeds every earlier one (and 0); *index = its position + start_pos. */
int ssg_select_master_frame(int32_t n, const double *quality, int32_t start_pos, int32_t *index);
/* create_reference_frame's stacking window around the master frame
 * (c_image_stacking_pipeline.cc:1219-1232). */
int ssg_master_stack_window(int32_t size, int32_t start_frame_index, int32_t master_frame_pos,
                            int32_t max_frames_to_stack, int32_t *startpos, int32_t *endpos);

/* ---- finish of a stack ------------------------------------------------------------------------ */

/* linear_interpolation_inpaint(image, mask) (core/proc/inpaint/linear_interpolation_inpaint.cc:
 * 335-361), in place: holes (mask == 0) are filled from the linear interpolations along their
 * row and their column, weighted by distance, repeated while holes get filled.  image: SSG_32FC1
 * or SSG_32FC3, host or device; mask: SSG_8UC1 of the image size (not modified). */
int ssg_inpaint_linear(ssg_image *image, const ssg_image *mask);

/* ---- stacking at a larger size: upscale of the frames before align, of the remaps after ------- */

/* frame_upscale_option (c_image_stacking_pipeline.h:43-49). */
enum { SSG_UPSCALE_NONE = 0, SSG_UPSCALE_X2 = 1, SSG_UPSCALE_X15 = 2, SSG_UPSCALE_X3 = 3 };

/* The upscaled size: x2 (pyrUp) 2w x 2h, x1.5 (w*3/2) x (h*3/2), x3 3w x 3h, none w x h. */
int ssg_upscale_size(int32_t option, int32_t w, int32_t h, int32_t *dw, int32_t *dh);
/* upscale_remap (c_image_stacking_pipeline.cc:1869-1905) of n remaps, the
 * frame_upscale_after_align step of the frame loop (:1633-1641): cv::resize INTER_LINEAR for
 * x1.5 and x3 (float data: INTER_LINEAR_EXACT is INTER_LINEAR), cv::pyrUp for x2, a copy for
 * none; the values (source positions) are not scaled.  src: SSG_32FC2 of one size; dst:
 * SSG_32FC2 of ssg_upscale_size, 8-byte aligned steps.  Device maps with one step per side
 * go in one launch; host maps are staged.  The upscaled remaps feed
 * ssg_accumulation_add_remapped_batch with an accumulator of the upscaled size. */
int ssg_upscale_remaps(int32_t option, int32_t n, const ssg_image *src, ssg_image *dst);
/* upscale_image(scale, src, srcmask, dst, dstmask) (c_image_stacking_pipeline.cc:1949-2000) of n
 * frames, the frame_upscale_before_align steps: the reference frame and its mask before
 * setup_reference_frame (:914-918), each frame with its mask and its weight map before
 * register_frame (:1474-1492).  x1.5 cv::resize INTER_LINEAR, x2 cv::pyrUp, x3 INTER_LINEAR_EXACT
 * (float data: INTER_LINEAR); the masks take OpenCV's 8-bit arithmetic of the same resize and are
 * then compare(dstmask, 255, CMP_GE) (0 / 255); none copies both.  src: SSG_32FC1 or SSG_32FC3
 * of one size and type, host or device, 4-byte aligned steps; dst: the same type at
 * ssg_upscale_size.  srcmask: NULL (no masks) or n SSG_8UC1 images of the source size, an entry
 * with data NULL meaning that frame has no mask; dstmask: n SSG_8UC1 of the upscaled size (the
 * entries of frames without a mask are not written).  Device images with one step per side go
 * in one launch per 1024; host images are staged. */
int ssg_upscale_images(int32_t option, int32_t n, const ssg_image *src, const ssg_image *srcmask, ssg_image *dst,
                       ssg_image *dstmask);

/* average_pyramid_inpaint(src, mask, dst, dstmask, max_levels)
 * (core/proc/inpaint/average_pyramid_inpaint.cc:96-126), the pipeline's fill of the accumulated
 * image (c_image_stacking_pipeline.cc:763-767, max_levels 100): a pyramid of 3x3 box averages
 * over the valid pixels (BORDER_REPLICATE; downstrike_even / upject_even keep the odd rows and
 * columns), each level falling back to its own valid pixels.  src / dst: SSG_32FC1 or
 * SSG_32FC3 of one size (dst may be src); mask: SSG_8UC1, 0 = hole, or NULL (then a copy);
 * dstmask: SSG_8UC1 or NULL.  Box sums are exact double sums rounded to float. */
int ssg_average_pyramid_inpaint(const ssg_image *src, const ssg_image *mask, ssg_image *dst, ssg_image *dstmask,
                                int32_t max_levels);
/* unsharp_mask(src, dst, sigma, alpha) (core/proc/unsharp_mask.cc:72-115) for f32 images: the
 * low-pass image of create_lpass_image (a gaussian with BORDER_REFLECT, through a pyrDown /
 * pyrUp chain for sigma > 2), then dst = src / (1 - alpha) - lpass * alpha / (1 - alpha).
 * src, dst: SSG_32FC1 or SSG_32FC3 of one size, host or device (dst may be src). */
int ssg_unsharp_mask(const ssg_image *src, ssg_image *dst, double sigma, double alpha);

#ifdef __cplusplus
}
#endif
#endif

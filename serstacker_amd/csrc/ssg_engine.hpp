// ssg_engine.hpp -- structures shared by the host engine and the kernels, and
// the launcher entry points of each kernel family.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>

#include "ssg_device.hpp"

namespace ssg {

// ---- per-frame ECC state (device) ----------------------------------------------------------
enum : int { XF_ACTIVE = 1 };

struct FrameXf {
  float p[6];     // parameter vector, reference layout of the pass's transform
  float C[2];     // euclidean centre
  MapCoef mc;     // map coefficients derived from p (see warp_xy)
  float eps;      // eps of the last iteration
  float cg;       // centring constant for the warped image (previous iteration's mean)
  int iters;      // iterations done in the current level
  int flags;      // XF_ACTIVE
};

constexpr int MAX_LEVELS = 8;

// ECC_ALIGN_METHOD values (ecc2.h:59-64)
enum : int { ECC_FA = 0, ECC_IC = 1, ECC_LM = 2, ECC_ICLM = 3 };

// Per-frame state of the Levenberg-Marquardt and inverse-compositional aligners
// (c_ecclm, c_ecc_inverse_compositional, c_ecclm_inverse_compositional).  A pass
// evaluates the map in FrameXf::mc: the current parameters (phase 0) or the
// trial parameters newp (phase 1).
struct FrameLm {
  double lambda;
  double err;       // error at the current parameters
  double rmsold;    // plain inverse compositional: the previous iteration's error
  double Hp[36];    // Hessian without damping
  double v[6];      // J^T rhs at the current parameters
  double dp[6];     // increment of the pending trial
  float newp[6];    // trial parameters
  float jc[5];      // inverse compositional: sa, ca, s, Cx, Cy of the J precompute
  int phase;
  int iteration;
  float dpe;        // eps of the last increment
  int pad;
};

// The J parameters of the inverse-compositional aligners when the reference
// keeps one J per level across frames (no translation-first pass)
struct JRec {
  int valid;
  float c[5];
};

struct FrameRes {
  int level_iters[2][MAX_LEVELS];  // [pass: 0 translation-first, 1 main][level]
  int level_failed[2][MAX_LEVELS];
  int rejected;                    // 0, SSG_REG_RHO_TRANSLATION or SSG_REG_RHO
  float eps;
  double rho[2];
};

// moment counts of the fused ECC pass
constexpr int NMOM_FULL = 41;   // affine moments (see ssg_ecc.hip)
constexpr int NMOM_TRANS = 14;  // translation moments
constexpr int NMOM_CORR = 10;   // correlation moments (6 sums, the extremes of f' and g')
constexpr int NMOM_LM_FULL = 25;  // c_ecclm pass: 18 Hessian + 6 J^T rhs moments + |rhs|^2
constexpr int NMOM_LM_TRANS = 6;  // gx^2, gx gy, gy^2, gx rhs, gy rhs, |rhs|^2
constexpr int NMOM_IC = 8;        // inverse-compositional pass: 6 J^T rhs moments, |rhs|^2, good pixels
constexpr int NMOM_REF = 18;      // reference-gradient Hessian moments of a level

constexpr int ECC_LW = 64;                    // LDS row width: one column per lane
constexpr int ECC_TW = ECC_LW - 4;            // output columns per tile (2-px gradient halo each side)
constexpr int ECC_RPT = 16;                   // rows per thread
constexpr int ECC_WAVES = 4;                  // waves (row groups) per block
constexpr int ECC_TH = ECC_RPT * ECC_WAVES;   // tile height
constexpr int ECC_LH = ECC_TH + 4;            // LDS rows
constexpr int ECC_THREADS = WAVE * ECC_WAVES; // 256
constexpr int CORR_TW = 64, CORR_TH = 64;     // correlation-pass tile (one column per lane, no halo)

struct LevelGeom {
  int w, h;          // level size
  int pitch;         // elements per row (current pyramid)
  long long offset;  // element offset of the level inside one frame's pyramid
};

// Arguments of one ECC iteration launch over a batch
struct EccPassArgs {
  // current pyramids: frame f, level image = cur + f * cur_fstride (+ level offset already applied)
  const float* cur;
  long long cur_fstride;
  int cur_pitch, cw, ch;
  const uint8_t* cmask;  // eroded current masks (same layout, u8) or null
  long long cmask_fstride;
  int cmask_pitch;
  // reference level
  const float* ref;
  int ref_pitch, w, h;
  const uint8_t* rmask;  // eroded reference mask or null
  int rmask_pitch;
  float cf;              // centring constant of the reference level
  int tiles_x, tiles_y, ntiles;
  // state
  FrameXf* xf;
  FrameRes* res;
  float* partials;        // [nframes][ntiles][NMOM]
  const int* active_in;
  const int* active_count_in;
  int* active_out;
  int* active_count_out;
  int nframes;
  int pass, level;
  int max_iterations;
  float max_eps;
  double update_step_scale;
  // Levenberg-Marquardt / inverse-compositional aligners
  int method;              // ECC_*
  FrameLm* lm;
  const float* rgx;        // reference gradients of the level (ref_pitch), zero off the reference mask
  const float* rgy;
  const double* ref_h;     // NMOM_REF moments of the reference gradients of the level
  double rma;              // reference-mask pixels of the level (the area without a mask)
};

struct CorrPassArgs {
  const float* cur;
  long long cur_fstride;
  int cur_pitch, cw, ch;
  const uint8_t* cmask;
  long long cmask_fstride;
  int cmask_pitch;
  const float* ref;
  int ref_pitch, w, h;
  const uint8_t* rmask;
  int rmask_pitch;
  float cf;
  int tiles_x, tiles_y, ntiles;
  const FrameXf* xf;
  FrameRes* res;
  float* partials;
  const int* frames;   // frames to check (alive ones)
  int nlist;
  int pass;            // 0: translation gate (0.75*min_rho), 1: final gate (min_rho)
  double min_rho;
};

struct LevelBeginArgs {
  FrameXf* xf_t;  // translation-first state
  FrameXf* xf_m;  // main state
  FrameRes* res;
  int nframes;
  int pass;       // 0 or 1
  int level, nlevels;
  double ratio;   // scale factor applied to the transform at this level start
  int main_motion;
  float cf;
  int* active_out;
  int* active_count_out;
  int method;              // ECC_*
  FrameLm* lm;
  JRec* jrec;              // per-level record when J persists across frames, else null
};

// ---- input stage (raw frames -> pipeline f32 images, debayer) ----------------------------------
// INPUT_PLANE0 + k (k = 0..3): Bayer plane k; INPUT_CV_*: OpenCV's bilinear / edge-aware demosaicing
enum : int { INPUT_COPY = 0, INPUT_NN2 = 1, INPUT_MATRIX = 2, INPUT_SP = 3, INPUT_AVGBP = 4, INPUT_PLANE0 = 5,
             INPUT_CV_BILINEAR = 9, INPUT_CV_EA = 10 };

struct InputArgs {
  const void* src;         // frame f at src + f * src_fstride (bytes), rows of src_pitch bytes
  long long src_fstride;
  int src_pitch;
  int w, h;                // source size
  int ow, oh;              // output size
  int pattern;             // Bayer layout: 0 RGGB, 1 GRBG, 2 GBRG, 3 BGGR
  int method;              // INPUT_*
  int convert;             // integer source converted to f32: multiply by scale
  float scale;             // 1 / 2^bpp
  const float* dark;       // darkbayer (floats, dark_pitch per row) or null
  int dark_pitch;
  const float* flat;       // flatbayer or null
  int flat_pitch;
  void* dst;               // frame f at dst + f * dst_fstride (bytes)
  long long dst_fstride;
  int dst_pitch;
};

// in_depth: 0 u8, 2 u16, 5 f32 (OpenCV depth codes); out_f32: f32 output (else the
// source depth); float_work: the pipeline's float samples (dark/flat or f32 sources)
hipError_t launch_input(const InputArgs& a, int in_depth, bool out_f32, bool float_work, int nframes, hipStream_t s);

// ---- finish of a stack: linear-interpolation inpainting, unsharp mask -------------------------
struct InpaintArgs {
  float* img;              // interleaved f32 image, pitch floats per row
  int pitch, w, h, cn;
  uint8_t* mask;           // 0 = hole; filled holes become 255
  int mpitch;
  float* ih;               // horizontal / vertical fills [h][w][cn]
  float* iv;
  float* hd;               // their distances [h][w], 0 where the row / column has no data
  float* vd;
  int* s_tmp;              // [h][w] scratch of the row scan
  unsigned* filled;        // holes filled by the pass
};
hipError_t launch_inpaint_pass(const InpaintArgs& a, hipStream_t s);
// frame_upscale_option (c_image_stacking_pipeline.h:43-49)
enum { UPSCALE_NONE = 0, UPSCALE_X2 = 1, UPSCALE_X15 = 2, UPSCALE_X3 = 3 };
void upscale_size(int option, int w, int h, int* dw, int* dh);
// n interleaved 2-channel f32 fields (pitches in floats) upscaled by the option (not NONE)
hipError_t launch_upscale_maps(int option, int n, const float* const* src, int spitch, int sw, int sh,
                               float* const* dst, int dpitch, hipStream_t s);
// n interleaved cn-channel f32 images (cn 1..3, pitches in floats) upscaled by the option (not NONE)
hipError_t launch_upscale_f32(int option, int cn, int n, const float* const* src, int spitch, int sw, int sh,
                              float* const* dst, int dpitch, hipStream_t s);
// n 8-bit masks (pitches in bytes) upscaled by the option (not NONE), then compared with 255
hipError_t launch_upscale_masks(int option, int n, const uint8_t* const* src, int spitch, int sw, int sh,
                                uint8_t* const* dst, int dpitch, hipStream_t s);
hipError_t launch_apr_init(const float* src, int spitch, const uint8_t* mask, int mpitch, int w, int h, int cn,
                           float* img, float* msk, unsigned* nz, hipStream_t s);
hipError_t launch_apr_down(const float* img, const float* msk, int w, int h, int cn, float* dimg, float* dmsk, int dw,
                           int dh, unsigned* nz, hipStream_t s);
hipError_t launch_apr_up(float* img, float* msk, int w, int h, int cn, const float* simg, int sw, hipStream_t s);
hipError_t launch_apr_mask_out(const float* msk, int w, int h, uint8_t* dst, int dpitch, hipStream_t s);
hipError_t launch_unsharp_blend(const float* src, int spitch, const float* lp, int lpitch, float* dst, int dpitch,
                                int wcn, int h, float alpha, float beta, hipStream_t s);

struct AccumArgs {
  const float* const* frames;  // n device pointers
  int src_pitch, sw, sh, cn;   // all frames share the geometry
  const MapCoef* maps;         // per frame (parametric motions)
  const float2* const* rmaps;  // per frame dense remap fields (MOTION_MAP), pitch rmap_pitch float2
  int rmap_pitch;
  const float* weights;        // per frame scalar weight (when wmaps is null)
  const float* const* wmaps;   // per frame weight image on the frame grid (f32) or null
  int wmap_pitch;              // elements per row of the weight images
  const int* valid;            // per frame 0/1
  int nframes, per_chunk;      // frames [c*per_chunk, min((c+1)*per_chunk, nframes)) go to chunk c
  int dw, dh;                  // accumulator size
  int interp, border;
  float cval;
  // chunk partial output planes [chunk][plane][dh*dw]
  float* part;
  unsigned* part_cnt;
  long long plane;             // elements per plane (dw*dh)
  int nchunks;
  int rows_per_thread;         // tile height / 4 waves: 2 or 4 (the tuning knob SSG_TUNE_ACC_ROWS)
};

// the tuning knobs (ssg_tune_set): the value in force for `knob`
int tune_value(int knob);
// chunks of the fused accumulate for n frames on a w x h accumulator (frames per chunk follow)
int accumulate_chunks(int n, int w, int h);

// c_bayer_average state: planes [c][h*w] for c = B, G, R (the reference's BGR order)
struct BayerState {
  float* acc;       // 3 planes: sum of value * weight (float, the reference's Mat3f)
  float* cntr;      // 3 planes: sum of weights
  unsigned* cnt;    // 3 planes: taps with a non-zero weight
  long long plane;  // elements per plane
  int w, h;
  unsigned pattern; // 2-bit channel of each 2x2 cell position ((y&1)<<1 | (x&1))
};

struct BayerBatchArgs {
  const void* const* frames;  // n device pointers, f32 or u16 raw
  int src_pitch;              // elements per row
  float scale;                // u16 -> f32 scale (2^-bpp)
  const MapCoef* maps;
  const int* valid;
  int nframes;
  BayerState st;
};

// ---- launchers --------------------------------------------------------------------------------
// pyramids (ssg_pyr.hip)
constexpr int SEP_MAX_TAPS = 49;
struct SepKernel {
  float kx[SEP_MAX_TAPS];
  float ky[SEP_MAX_TAPS];
  int nx, ny, tx, ty;  // taps and symmetry type (0 other, 1 symmetric, 2 antisymmetric)
  float delta;         // added to every output (cv::sepFilter2D's delta)
};
hipError_t launch_sepfilter(const float* const* src_ptrs, int src_pitch, float* dst, long long dst_fstride,
                            int dst_pitch, int w, int h, int nframes, const SepKernel& k, int border,
                            hipStream_t s);
hipError_t launch_pyrdown(const float* src, long long src_fstride, int src_pitch, int sw, int sh, float* dst,
                          long long dst_fstride, int dst_pitch, int dw, int dh, int nframes, int border,
                          hipStream_t s);
// level 0 smoothed (a symmetric gaussian of 3..9 taps, BORDER_REPLICATE) and level 1 (pyrDown,
// BORDER_REFLECT101) of a pyramid in one pass; level 1 starts off1 floats after level 0 in each
// frame's block of fstride floats
bool smooth_pyrdown_supported(const SepKernel& k, int src_pitch, const float* pyr, long long fstride, int pitch0,
                              int w, int h, long long off1, int pitch1, int dw, int dh, int nframes);
hipError_t launch_smooth_pyrdown(const float* const* src_ptrs, int src_pitch, float* pyr, long long fstride,
                                 int pitch0, int w, int h, long long off1, int pitch1, int dw, int dh, int nframes,
                                 const SepKernel& k, hipStream_t s);
// pyrDown (BORDER_REFLECT101) by column strips, one wave each: the values of launch_pyrdown. The
// frames come from src + f * src_fstride, or (src_ptrs, with src null) from a device pointer table;
// strip_pyrdown_supported takes src null for the table form
bool strip_pyrdown_supported(const float* src, long long src_fstride, int src_pitch, int w, int h, const float* dst,
                             long long dst_fstride, int dst_pitch, int dw, int dh, int nframes);
hipError_t launch_strip_pyrdown(const float* const* src_ptrs, const float* src, long long src_fstride, int src_pitch,
                                int w, int h, float* dst, long long dst_fstride, int dst_pitch, int dw, int dh,
                                int nframes, hipStream_t s);
hipError_t launch_copy_frames(const float* const* src_ptrs, int src_pitch, float* dst, long long dst_fstride,
                              int dst_pitch, int w, int h, int nframes, hipStream_t s);
hipError_t launch_mask_prepare(const uint8_t* const* src_ptrs, int src_pitch, uint8_t* dst, long long dst_fstride,
                               int dst_pitch, int sw, int sh, int dw, int dh, int nframes, hipStream_t s);
hipError_t launch_erode5(const uint8_t* src, long long src_fstride, int src_pitch, uint8_t* dst,
                         long long dst_fstride, int dst_pitch, int w, int h, int nframes, int border, uint8_t cval,
                         hipStream_t s);
hipError_t launch_count_zero_u8(const uint8_t* src, long long fstride, int pitch, int w, int h, int nframes,
                                unsigned* out, hipStream_t s);
hipError_t launch_sum_f32(const float* src, int pitch, int w, int h, double* out, hipStream_t s);
hipError_t launch_sharpness(const float* const* src_ptrs, int pitch, int w, int h, int nframes, double* partial,
                            hipStream_t s);
hipError_t launch_ecc_channel(const float* const* src, int pitch, int w, int h, int n, int channel, float* dst,
                              long long dst_fs, int dst_pitch, hipStream_t s);
hipError_t launch_zero(void* p, size_t bytes, hipStream_t s);
hipError_t launch_xf_to_acc(const FrameXf* xf, const FrameRes* res, const float* w, int n, int motion,
                            double inv_scale, MapCoef* maps, int* valid, hipStream_t s);
hipError_t launch_sharpness_finish(const double* partial, int nb, int nframes, double area, float* out, hipStream_t s);
// dst = src - up, 0 where masks[f] (if any) is not 255; src / masks: device arrays of per-frame pointers
hipError_t launch_norm_sub(const float* const* src, int spitch, const float* up, long long ufs,
                           const uint8_t* const* masks, int mpitch, float* dst, int dpitch, long long dfs, int w, int h,
                           int nframes, hipStream_t s);
hipError_t launch_u16_to_f32(const uint16_t* src, int src_pitch, float* dst, int dst_pitch, int w, int h,
                             float scale, hipStream_t s);

// ECC (ssg_ecc.hip)
hipError_t launch_level_begin(const LevelBeginArgs& a, hipStream_t s);
hipError_t launch_ecc_pass(int motion, const EccPassArgs& a, int grid, hipStream_t s);
// c_ecclm pass (gradients of the warped image, J at the evaluated parameters) and the
// inverse-compositional pass (reference gradients); then the per-frame state machines
hipError_t launch_lm_pass(int motion, const EccPassArgs& a, int grid, hipStream_t s);
hipError_t launch_ic_pass(int motion, const EccPassArgs& a, int grid, hipStream_t s);
hipError_t launch_lm_solve(int motion, const EccPassArgs& a, hipStream_t s);
// reference gradients of a level (ecc_differentiate, zeroed where mask == 0) and their
// Hessian moments (deterministic tile order); part: scratch of ntiles * NMOM_REF doubles
hipError_t launch_ic_reference(const float* ref, int pitch, int w, int h, const uint8_t* mask, int mpitch, float* gx,
                               float* gy, double* part, double* out, hipStream_t s);
hipError_t launch_ecc_solve(int motion, const EccPassArgs& a, hipStream_t s);
hipError_t launch_corr_pass(int motion, const CorrPassArgs& a, int grid, hipStream_t s);
int ecc_persistent_grid();

// accumulation / remap (ssg_accum.hip)
hipError_t launch_accumulate(int motion, const AccumArgs& a, hipStream_t s);
hipError_t launch_accum_merge(const float* part, const unsigned* part_cnt, int nchunks, long long plane, int cn,
                              float* acc, float* accw, unsigned* cnt, hipStream_t s);
hipError_t launch_acc_add(const float* src, int src_pitch, int w, int h, int cn, const uint8_t* mask, int mask_pitch,
                          const float* wmap, int wmap_pitch, float* acc, float* accw, unsigned* cnt,
                          long long plane, hipStream_t s);
hipError_t launch_acc_reinit(const float* src, int src_pitch, const float* accw, int accw_pitch, int w, int h, int cn,
                             long long plane, float* acc, unsigned* cnt, hipStream_t s);
hipError_t launch_acc_compute(const float* acc, const float* accw, int w, int h, int cn, long long plane,
                              float* avg, int avg_pitch, uint8_t* mask, int mask_pitch, double dscale,
                              hipStream_t s);
hipError_t launch_remap(int motion, const float* src, int src_pitch, int sw, int sh, int cn, const MapCoef& mc,
                        const float* map, int map_pitch, float* dst, int dst_pitch, int dw, int dh, int interp,
                        int border, float cval, hipStream_t s);
hipError_t launch_remap_mask(int motion, const uint8_t* src_mask, int src_pitch, int sw, int sh, const MapCoef& mc,
                             const float* map, int map_pitch, uint8_t* dst, int dst_pitch, int dw, int dh,
                             hipStream_t s, int interp);
hipError_t set_cubic_itab(const short* tab, hipStream_t s);
hipError_t launch_create_map(int motion, const MapCoef& mc, float* map, int map_pitch, int w, int h, hipStream_t s);

// bayer average (ssg_bayer.hip)
unsigned bayer_pattern_bits(int colorid);  // 0 for a colorid that is not a Bayer pattern
hipError_t launch_bayer_add(const BayerState& st, const void* src, int src_pitch, bool u16, float scale,
                            const float* rmap, int rmap_pitch, const uint8_t* mask, int mask_pitch,
                            const float* wts, int wts_pitch, hipStream_t s);
hipError_t launch_bayer_batch(int motion, bool u16, const BayerBatchArgs& a, hipStream_t s);
hipError_t launch_bayer_compute(const BayerState& st, float* avg, int avg_pitch, uint8_t* mask, int mask_pitch,
                                hipStream_t s);
hipError_t launch_bayer_counters(const BayerState& st, float* out, int out_pitch, hipStream_t s);

// focus stacking (ssg_focus.hip)
hipError_t launch_split(const float* src, int pitch, int w, int h, int cn, float* planes, hipStream_t s);
hipError_t launch_merge(const float* planes, int w, int h, int cn, float* dst, int pitch, hipStream_t s);
hipError_t launch_gray(const float* planes, size_t n, int cn, float* g, hipStream_t s);
hipError_t launch_lpg(const float* g, int w, int h, float alpha, float beta, float eps, float p, float* m,
                      hipStream_t s);
hipError_t launch_pyrup(const float* src, long long src_fstride, int sw, int sh, float* dst, long long dst_fstride,
                        int dw, int dh, int nplanes, hipStream_t s);
hipError_t launch_ew(int op, const float* a, const float* b, float* d, size_t n, hipStream_t s);  // 0 a-b, 1 a+b, 2 |a|
hipError_t launch_div(const float* a, const float* w, int wc, float* d, size_t n, int cn, hipStream_t s);
hipError_t launch_wacc(float* acc, float* wsum, const float* W, const float* L, size_t n, int cn, hipStream_t s);
hipError_t launch_select(float* acc, const float* pyr, const float* w0, const float* w1, size_t n, hipStream_t s);
hipError_t launch_energy(const float* lap, float* w, size_t n, int cn, int avgc, hipStream_t s);
hipError_t launch_mulw(float* lap, const float* w, int wc, size_t n, int cn, hipStream_t s);
hipError_t launch_blend(float* a, const float* b, float alpha, float beta, size_t n, hipStream_t s);
hipError_t launch_plane_stdev(const float* planes, size_t n, int nplanes, double* out, hipStream_t s);
hipError_t launch_select_frames(float* acc, const float* L, size_t ne, size_t fstride, int m, bool first,
                                hipStream_t s);
hipError_t launch_wavg_frames(float* acc, float* wsum, const float* L, size_t n, int cn, int avgc, int wc, int m,
                              bool first, hipStream_t s);
hipError_t launch_wacc_frames(float* acc, float* wsum, const float* W, const float* L, size_t n, int cn, int m,
                              hipStream_t s);
hipError_t launch_top_frames(float* acc, const float* L, size_t n, int cn, int m, bool first, hipStream_t s);

// local-variance sharpness map (ssg_lvm.hip)
hipError_t launch_bgr2gray(const float* src, int pitch, int w, int h, float* g, hipStream_t s);
hipError_t launch_morph_grad(const float* m, int w, int h, int r, float* g, hipStream_t s);
hipError_t launch_int_to_f32(const void* src, int pitch, int w, int h, bool u16, float* dst, hipStream_t s);
hipError_t launch_round_half_up(float* v, size_t n, hipStream_t s);
int lvm_sum_blocks(size_t n);
hipError_t launch_lvm_sums(const float* g, size_t n, double* partial, hipStream_t s);
hipError_t launch_lvm_map(const float* g, size_t n, float map_scale, float add, float* m, hipStream_t s);
hipError_t launch_resize_linear(const float* src, int sw, int sh, float* dst, int dst_pitch, int dw, int dh,
                                hipStream_t s);

// FFT phase correlation (ssg_fft.hip)
hipError_t launch_fft2(float2* buf, float2* tmp, int M, int N, int nframes, const float2* tw_m, const float2* tw_n,
                       bool inverse, hipStream_t s);
constexpr int PC_PARTS_MAX = 64;  // blocks per frame of the phase-correlation reductions
hipError_t launch_pc_load(const void* const* frames, bool u16, int pitch, int bw, int bh, int binning, float* mean,
                          double* part, const float* wx, const float* wy, float2* out, int M, int N, int nframes,
                          hipStream_t s);
hipError_t launch_pc_cross(float2* spec, const float2* ref, int M, int N, int nframes, hipStream_t s);
hipError_t launch_pc_peak(const float2* r, int M, int N, int binning, int nframes, float* pv, int* pi, double* out,
                          hipStream_t s);

// dense ECC flow (ssg_eccflow.hip)
struct AreaGeom {  // one cv::resize INTER_AREA geometry, sw x sh -> dw x dh
  int sw, sh, dw, dh;
  int fast, isx, isy;  // whole factors on both axes (resizeAreaFast_)
  // otherwise the computeResizeAreaTab entries of output column dx are
  // [xoff[dx], xoff[dx+1]) of (xsi, xal); rows likewise
  const int* xoff;
  const int* xsi;
  const float* xal;
  const int* yoff;
  const int* ysi;
  const float* yal;
};
struct FlowLevelArgs {
  const float* cur;      // current level, frame-strided (cfs), pitch w
  long long cfs;
  const uint8_t* cmask;  // current mask level (frame-strided, mfs) or null
  long long mfs;
  const float* ref;      // reference level, pitch w
  const float* Ix;
  const float* Iy;
  const uint8_t* rmask;  // reference mask level or null
  int w, h;
};
enum { CUBIC_ADD = 0, CUBIC_PLAIN = 1, CUBIC_INIT = 2 };
struct FlowCubicArgs {
  const float2* src;  // source flow, frame-strided (sfs), pitch sw
  long long sfs;
  int sw, sh;
  const float2* const* maps;  // CUBIC_INIT: the caller's remaps (pitch map_pitch)
  int map_pitch;
  float2* dst;  // output, frame-strided (dfs), pitch dw
  long long dfs;
  int dw, dh;
  const int* xs;  // per output column: floor of the source position; per output row likewise
  const float4* xa;
  const int* ys;
  const float4* ya;
  double rx, ry;
};
// One fused eccflow iteration over a level whose cells are k x k pixels (k = 4..32, a power of
// two dividing the level size): optionally adds the pending cell update (cubic-upsampled) to uv,
// then warps, forms the residual products and sums them per cell in the INTER_AREA order.
struct FlowIterArgs {
  FlowLevelArgs L;
  float2* uv;  // level flow, frame-strided (ufs), pitch L.w
  long long ufs;
  const float2* cuv;  // pending cell update (pitch cw), frame-strided (cfs)
  long long cfs;
  int cw, ch, k;
  const int* xs;  // cells -> level INTER_CUBIC tables
  const float4* xa;
  const int* ys;
  const float4* ya;
  float* cells;  // cell sums, float2 interleaved (pitch 2 cw floats), frame-strided (cellfs floats)
  long long cellfs;
  float2* const* maps;  // final pass at level 0: identity + uv into the caller's maps
  int map_pitch;
};
hipError_t launch_flow_iter(bool add, bool resid, bool to_map, const FlowIterArgs& a, int nframes, hipStream_t s);
hipError_t launch_area(int cn, const float* src, long long sfs, int spitch, float* dst, long long dfs, int dpitch,
                       const AreaGeom& g, int nframes, hipStream_t s);
hipError_t launch_flow_products(const float* Ix, const float* Iy, int n, float2* P2, float* P1, hipStream_t s);
hipError_t launch_flow_D(const float* A2, const float* A1, int cw, int ch, float um, float reg, float4* D,
                         hipStream_t s);
hipError_t launch_flow_residual(const FlowLevelArgs& a, const float2* uv, long long ufs, float2* itxy, long long ifs,
                                int nframes, hipStream_t s);
hipError_t launch_flow_solve(const float* cells, long long cfs, const float4* D, int cw, int ch, float2* cuv,
                             long long ufs, int nframes, hipStream_t s);
hipError_t launch_flow_cubic(int mode, const FlowCubicArgs& a, int nframes, hipStream_t s);
hipError_t launch_flow_scale(float2* uv, long long ufs, int w, int h, double rx, double ry, int nframes, hipStream_t s);
hipError_t launch_flow_to_remap(const float2* uv, long long ufs, int w, int h, float2* const* maps, int pitch,
                                int nframes, hipStream_t s);
hipError_t launch_nearest_u8(const uint8_t* src, long long sfs, int sw, int sh, uint8_t* dst, long long dfs, int dw,
                             int dh, int nframes, hipStream_t s);

// synthetic frames (ssg_synth.hip)
struct SynthFrame {
  double A[6];  // scene coordinate = A * (x, y, 1): maps frame pixels to the scene
  float blur_sigma;
  uint32_t key;
};

void set_cubic_table(const float tab[32][4], hipStream_t s);

// ---- multi-sequence focus stacking (ssg_focus_multi.hip) -------------------------------------
// One launch per step serves every sequence of a batch: the blocks of a launch are split over
// jobs (a sequence, a level, a kind) by their prefix block counts.
constexpr int FM_MAXL = 16;
struct FmSeq {
  int w, h, cn, L, m;               // level-0 size, channels, levels, frames in this chunk
  int lw[FM_MAXL], lh[FM_MAXL];
  long long loff[FM_MAXL];          // single-plane element offset of each level
  float* G;     // chunk pyramids: level i at G + m*cn*loff[i], planes [frame][channel] of lw*lh
  float* W;     // chunk weight pyramids: level i at W + m*loff[i], planes [frame]
  float* gray;  // [frame][pixel] channel average of the aligned level-0 frame
  float* M;     // [frame][pixel] lpg maps
  float* acc;   // fused sums: level i at acc + cn*loff[i], planes [channel]
  float* wsum;  // weight sums: level i at wsum + loff[i]
  const float* const* frames;  // m device frames, interleaved cn channels
  const float* shift;          // 2 floats per frame (source offset x, y) or null
  int frame_pitch;             // floats per frame row
  float alpha, beta, eps, p;   // lpg coefficients and power
  SepKernel K;                 // the weight smoothing (symmetric gaussian)
};
enum { FM_PREP = 0, FM_LPG = 1, FM_BLUR = 2, FM_DOWN_G = 3, FM_DOWN_W = 4, FM_FUSE = 5, FM_TOP = 6 };
struct FmJob {
  int seq, level, kind, end;  // end: the launch's block index one past this job's last block
};
constexpr int FM_MAXJOBS = 64;
struct FmJobs {
  int n;
  FmJob j[FM_MAXJOBS];
};
// blocks of one job (kind, level) of sequence q
long long fm_job_blocks(const FmSeq& q, int kind, int level);
// step `kind` (FM_PREP, FM_LPG, FM_BLUR, FM_DOWN_*, FM_FUSE/FM_TOP) of every job in J; seqs: device table;
// blur_radius: the smoothing radius shared by every job of an FM_BLUR launch (0: they differ)
hipError_t launch_fm_step(int kind, const FmJobs& J, const FmSeq* seqs, int blur_radius, hipStream_t s);

}  // namespace ssg

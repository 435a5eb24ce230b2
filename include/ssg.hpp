/*
 * ssg.hpp -- header-only C++ host interface over the C ABI (ssg.h), with the
 * class and method names of SerStacker's registration and accumulation
 * interfaces, so pipeline code binds to it one call for one call:
 *
 *   ssg::c_frame_registration   core/proc/image_registration/c_frame_registration.h:210-354
 *   ssg::c_eccflow              core/proc/image_registration/ecc2.h:548-662
 *   ssg::c_frame_accumulation   core/average/c_frame_accumulation.h:14-37
 *   ssg::c_weigthed_average     core/average/c_frame_accumulation.h:39-63
 *   ssg::c_bayer_average        core/average/c_frame_accumulation.h:222-262
 *   ssg::c_laplacian_pyramid_focus_stacking  core/average/c_frame_accumulation.h:141-184
 *   ssg::phase_correlation      (no reference counterpart: the FFT pre-alignment)
 *   ssg::c_ser_reader           core/io/c_ser_file.h:136-215
 *   ssg::debayer, ssg::read_input_frames   core/io/debayer.cc:1197-1355,
 *                               c_image_stacking_pipeline_base.cc:124-279
 *   ssg::master_frame_quality and the master-frame windows
 *                               c_image_stacking_pipeline_base.cc:282-404,
 *                               c_image_stacking_pipeline.cc:1219-1232
 *   ssg::linear_interpolation_inpaint, ssg::average_pyramid_inpaint, ssg::unsharp_mask
 *                               core/proc/inpaint/, core/proc/unsharp_mask.cc
 *   ssg::upscale_remap          c_image_stacking_pipeline.cc:1869-1905
 *
 * Images are ssg::image views (pointer, size, byte stride, type, host/device)
 * instead of cv::InputArray; the caller owns the memory.  Methods return bool
 * as the reference's do; the reason of a failure is ssg::last_error().
 * Plain C++17, no HIP or OpenCV headers.
 */
#ifndef SSG_HPP
#define SSG_HPP

#include <cstdint>
#include <string>
#include <utility>
#include <vector>

#include "ssg.h"

namespace ssg {

inline std::string last_error() { return ssg_last_error(); }

// A non-owning image view.
struct image : ssg_image {
  image() : ssg_image{nullptr, 0, 0, 0, SSG_32FC1, 0} {}
  image(void *p, int w, int h, int64_t step_bytes, int type, bool device)
      : ssg_image{p, w, h, step_bytes, type, device ? 1 : 0} {}
  // dense host/device views
  static image f32(float *p, int w, int h, bool device, int channels = 1) {
    return image(p, w, h, (int64_t)w * channels * 4, channels == 3 ? SSG_32FC3 : SSG_32FC1, device);
  }
  static image u8(uint8_t *p, int w, int h, bool device) { return image(p, w, h, w, SSG_8UC1, device); }
  static image map2f(float *p, int w, int h, bool device) { return image(p, w, h, (int64_t)w * 8, SSG_32FC2, device); }
  bool empty() const { return data == nullptr; }
};

using registration_options = ssg_image_registration_options;
using registration_status = ssg_registration_status;

inline registration_options default_registration_options() {
  registration_options o;
  ssg_default_image_registration_options(&o);
  return o;
}

// c_frame_registration (c_frame_registration.h:210-354), ECC path.
class c_frame_registration {
 public:
  c_frame_registration() : c_frame_registration(default_registration_options()) {}
  explicit c_frame_registration(const registration_options &options) : opts_(options) {}
  ~c_frame_registration() { release(); }
  c_frame_registration(const c_frame_registration &) = delete;
  c_frame_registration &operator=(const c_frame_registration &) = delete;

  registration_options &options() { return opts_; }
  const registration_options &options() const { return opts_; }
  const registration_status &status() const { return status_; }

  // c_frame_registration::setup_reference_frame (c_frame_registration.cc:565-719)
  bool setup_reference_frame(const image &img, const image &mask = image()) {
    if (!ensure()) return false;
    return ssg_registration_setup_reference_frame(h_, &img, mask.empty() ? nullptr : &mask) == SSG_OK;
  }

  // c_frame_registration::register_frame (c_frame_registration.cc:721-964); with dst
  // given, also remaps the frame (and its mask) onto the reference grid (:922-932).
  bool register_frame(const image &src, const image &srcmask = image(), image *dst = nullptr,
                      image *dstmask = nullptr) {
    if (!h_) return fail_state();
    if (ssg_registration_register_frame(h_, &src, srcmask.empty() ? nullptr : &srcmask, &status_) != SSG_OK)
      return false;
    if (!status_.ok) return false;
    if (dst) return remap(src, *dst, srcmask, dstmask);
    return true;
  }

  // Batched register_frame over device-resident frames (no reference counterpart).
  bool register_batch(const std::vector<image> &frames, std::vector<registration_status> &out,
                      const std::vector<image> *masks = nullptr) {
    if (!h_) return fail_state();
    out.resize(frames.size());
    std::vector<ssg_image> f(frames.begin(), frames.end()), m;
    if (masks) m.assign(masks->begin(), masks->end());
    return ssg_registration_register_batch(h_, (int)f.size(), f.data(), masks ? m.data() : nullptr, out.data()) ==
           SSG_OK;
  }

  // c_frame_registration::current_remap (c_frame_registration.h:292): the remap
  // field of the last registered frame at the reference size (SSG_32FC2), refined
  // by the dense ECC flow when enable_eccflow_registration is set.
  bool current_remap(image &map) const {
    if (!h_) return fail_state();
    return ssg_registration_current_remap(h_, &map) == SSG_OK;
  }

  // Batched register_frame that also writes each frame's remap (device SSG_32FC2 maps).
  bool register_batch_remaps(const std::vector<image> &frames, std::vector<registration_status> &out,
                             std::vector<image> &remaps, const std::vector<image> *masks = nullptr) {
    if (!h_ || remaps.size() != frames.size()) return fail_state();
    out.resize(frames.size());
    std::vector<ssg_image> f(frames.begin(), frames.end()), r(remaps.begin(), remaps.end()), m;
    if (masks) m.assign(masks->begin(), masks->end());
    return ssg_registration_register_batch_remaps(h_, (int)f.size(), f.data(), masks ? m.data() : nullptr,
                                                  out.data(), r.data()) == SSG_OK;
  }

  // c_frame_registration::remap (c_frame_registration.cc:1405-1417): warps with the
  // current remap; unknown interpolation/border (-1) take the options' values.
  bool remap(const image &src, image &dst, const image &src_mask = image(), image *dst_mask = nullptr,
             int interpolation = -1, int border_mode = -1, double border_value = 0) const {
    if (!h_) return fail_state();
    return ssg_registration_remap(h_, &src, src_mask.empty() ? nullptr : &src_mask, &dst, dst_mask, interpolation,
                                  border_mode, border_value) == SSG_OK;
  }

  // c_frame_registration::custom_remap (c_frame_registration.cc:1302-1417)
  bool custom_remap(const image &rmap, const image &src, image &dst, const image &src_mask = image(),
                    image *dst_mask = nullptr, int interpolation = -1, int border_mode = -1,
                    double border_value = 0) const {
    return ssg_remap(&src, src_mask.empty() ? nullptr : &src_mask, &rmap, nullptr, opts_.motion_type, nullptr,
                     &dst, dst_mask, interpolation < 0 ? opts_.interpolation : interpolation,
                     border_mode < 0 ? opts_.border_mode : border_mode,
                     border_mode < 0 ? opts_.border_value[0] : border_value) == SSG_OK;
  }

  // weights + registration + fused warp/accumulate of a device batch with every per-frame decision
  // on the device, replayed as a captured hipGraph when use_graph (ssg_stack_batch)
  template <class Acc>
  bool stack_batch(Acc &acc, const std::vector<image> &frames, std::vector<registration_status> &out,
                   bool use_graph = true, int interpolation = SSG_INTER_LINEAR,
                   int border_mode = SSG_BORDER_REFLECT101, double border_value = 0) {
    if (!h_) return fail_state();
    out.resize(frames.size());
    std::vector<ssg_image> f(frames.begin(), frames.end());
    return ssg_stack_batch(h_, acc.handle(), (int)f.size(), f.data(), interpolation, border_mode, border_value,
                           use_graph ? 1 : 0, out.data()) == SSG_OK;
  }

  ssg_registration *handle() const { return h_; }

 private:
  bool ensure() {
    if (h_) return true;
    return ssg_registration_create(&opts_, &h_) == SSG_OK;
  }
  static bool fail_state() { return false; }
  void release() {
    if (h_) ssg_registration_destroy(h_);
    h_ = nullptr;
  }

  registration_options opts_;
  registration_status status_{};
  ssg_registration *h_ = nullptr;
};

// c_frame_accumulation (c_frame_accumulation.h:14-37)
class c_frame_accumulation {
 public:
  virtual ~c_frame_accumulation() = default;
  virtual bool add(const image &src, const image &mask = image()) = 0;
  virtual bool compute(image &avg, image *mask = nullptr, double dscale = 1.0) const = 0;
  virtual bool get_acc_counters(image &accw) const = 0;
  // c_frame_accumulation::reinitialize (c_frame_accumulation.h:25)
  virtual bool reinitialize(const image &src, const image &accw) = 0;
  virtual void clear() = 0;
  virtual std::pair<int, int> accumulator_size() const = 0;
  virtual int accumulated_frames() const = 0;
  // checkpoint / resume of the accumulator (ssg_accumulation_state_save / _load)
  virtual bool save_state(std::vector<uint8_t> &out) const = 0;
  virtual bool load_state(const std::vector<uint8_t> &in) = 0;

 protected:
  static bool save_state_of(ssg_accumulation *a, std::vector<uint8_t> &out) {
    uint64_t n = 0;
    if (!a || ssg_accumulation_state_size(a, &n) != SSG_OK) return false;
    out.resize(n);
    return ssg_accumulation_state_save(a, out.data(), n) == SSG_OK;
  }
  static bool load_state_of(ssg_accumulation *a, const std::vector<uint8_t> &in) {
    return a && ssg_accumulation_state_load(a, in.data(), in.size()) == SSG_OK;
  }
};

// c_weigthed_average (c_frame_accumulation.h:39-63, c_frame_accumulation.cc:20-260)
class c_weigthed_average : public c_frame_accumulation {
 public:
  c_weigthed_average(int width, int height, int channels = 1) : w_(width), h_(height), cn_(channels) {
    ssg_accumulation_create(SSG_ACC_WEIGHTED_AVERAGE, width, height, channels, &a_);
  }
  ~c_weigthed_average() override {
    if (a_) ssg_accumulation_destroy(a_);
  }
  c_weigthed_average(const c_weigthed_average &) = delete;
  c_weigthed_average &operator=(const c_weigthed_average &) = delete;

  bool valid() const { return a_ != nullptr; }

  // add(src, weights): weights empty, SSG_8UC1 mask (nonzero = 1) or SSG_32FC1 weights
  bool add(const image &src, const image &weights = image()) override {
    return a_ && ssg_accumulation_add(a_, &src, weights.empty() ? nullptr : &weights) == SSG_OK;
  }

  // The fused per-frame sequence remap -> mask -> weights*mask/255 -> add of the
  // stacking pipeline (c_image_stacking_pipeline.cc:1644-1776) for registered frames.
  bool add_registered_batch(const std::vector<image> &frames, const std::vector<registration_status> &st,
                            int motion_type, const std::vector<float> *frame_weights = nullptr,
                            int interpolation = SSG_INTER_LINEAR, int border_mode = SSG_BORDER_REFLECT101,
                            double border_value = 0, const std::vector<image> *weight_maps = nullptr) {
    if (!a_ || st.size() != frames.size()) return false;
    std::vector<ssg_image> f(frames.begin(), frames.end()), wm;
    if (weight_maps) wm.assign(weight_maps->begin(), weight_maps->end());
    return ssg_accumulation_add_registered_batch(a_, (int)f.size(), f.data(), st.data(), motion_type,
                                                 frame_weights ? frame_weights->data() : nullptr,
                                                 weight_maps ? wm.data() : nullptr, interpolation, border_mode,
                                                 border_value) == SSG_OK;
  }

  // The same fused sequence with a dense remap field per frame (eccflow-refined remaps).
  bool add_remapped_batch(const std::vector<image> &frames, const std::vector<image> &rmaps,
                          const std::vector<float> *frame_weights = nullptr, int interpolation = SSG_INTER_LINEAR,
                          int border_mode = SSG_BORDER_REFLECT101, double border_value = 0,
                          const std::vector<image> *weight_maps = nullptr) {
    if (!a_ || rmaps.size() != frames.size()) return false;
    std::vector<ssg_image> f(frames.begin(), frames.end()), m(rmaps.begin(), rmaps.end()), wm;
    if (weight_maps) wm.assign(weight_maps->begin(), weight_maps->end());
    return ssg_accumulation_add_remapped_batch(a_, (int)f.size(), f.data(), m.data(),
                                               frame_weights ? frame_weights->data() : nullptr,
                                               weight_maps ? wm.data() : nullptr, interpolation, border_mode,
                                               border_value) == SSG_OK;
  }

  bool compute(image &avg, image *mask = nullptr, double dscale = 1.0) const override {
    return a_ && ssg_accumulation_compute(a_, &avg, mask, dscale) == SSG_OK;
  }
  bool get_acc_counters(image &accw) const override {
    return a_ && ssg_accumulation_get_acc_counters(a_, &accw) == SSG_OK;
  }
  bool get_counts(image &counts) const { return a_ && ssg_accumulation_get_counts(a_, &counts) == SSG_OK; }
  void clear() override {
    if (a_) ssg_accumulation_clear(a_);
  }
  std::pair<int, int> accumulator_size() const override { return {w_, h_}; }
  int accumulated_frames() const override {
    int32_t n = 0;
    if (a_) ssg_accumulation_accumulated_frames(a_, &n);
    return n;
  }
  // c_weigthed_average::reinitialize (c_frame_accumulation.cc:171-180)
  bool reinitialize(const image &src, const image &accw) override {
    return a_ && ssg_accumulation_reinitialize(a_, &src, &accw) == SSG_OK;
  }
  bool save_state(std::vector<uint8_t> &out) const override { return save_state_of(a_, out); }
  bool load_state(const std::vector<uint8_t> &in) override { return load_state_of(a_, in); }
  ssg_accumulation *handle() const { return a_; }

 private:
  int w_, h_, cn_;
  ssg_accumulation *a_ = nullptr;
};

// Accumulators other than the weighted average share one handle type.
class accumulation_handle : public c_frame_accumulation {
 public:
  accumulation_handle(int32_t method, int width, int height, int channels) : w_(width), h_(height) {
    ssg_accumulation_create(method, width, height, channels, &a_);
  }
  ~accumulation_handle() override {
    if (a_) ssg_accumulation_destroy(a_);
  }
  accumulation_handle(const accumulation_handle &) = delete;
  accumulation_handle &operator=(const accumulation_handle &) = delete;
  bool valid() const { return a_ != nullptr; }
  bool add(const image &src, const image &weights = image()) override {
    return a_ && ssg_accumulation_add(a_, &src, weights.empty() ? nullptr : &weights) == SSG_OK;
  }
  bool compute(image &avg, image *mask = nullptr, double dscale = 1.0) const override {
    return a_ && ssg_accumulation_compute(a_, &avg, mask, dscale) == SSG_OK;
  }
  // add() of each frame in order, batched (focus stacks build the batch's pyramids together)
  bool add_frames(const std::vector<image> &frames) {
    std::vector<ssg_image> f(frames.begin(), frames.end());
    return a_ && ssg_accumulation_add_frames(a_, (int)f.size(), f.data()) == SSG_OK;
  }
  bool get_acc_counters(image &accw) const override {
    return a_ && ssg_accumulation_get_acc_counters(a_, &accw) == SSG_OK;
  }
  void clear() override {
    if (a_) ssg_accumulation_clear(a_);
  }
  std::pair<int, int> accumulator_size() const override { return {w_, h_}; }
  int accumulated_frames() const override {
    int32_t n = 0;
    if (a_) ssg_accumulation_accumulated_frames(a_, &n);
    return n;
  }
  // Bayer and focus accumulators refuse it, as the reference's return false
  bool reinitialize(const image &src, const image &accw) override {
    return a_ && ssg_accumulation_reinitialize(a_, &src, &accw) == SSG_OK;
  }
  bool save_state(std::vector<uint8_t> &out) const override { return save_state_of(a_, out); }
  bool load_state(const std::vector<uint8_t> &in) override { return load_state_of(a_, in); }
  ssg_accumulation *handle() const { return a_; }

 protected:
  int w_, h_;
  ssg_accumulation *a_ = nullptr;
};

// c_bayer_average (c_frame_accumulation.h:222-262, c_frame_accumulation.cc:988-1334): raw
// Bayer frames (SSG_32FC1 or SSG_16UC1) into a 3-channel BGR accumulator.
class c_bayer_average : public accumulation_handle {
 public:
  c_bayer_average(int width, int height) : accumulation_handle(SSG_ACC_BAYER_AVERAGE, width, height, 1) {}
  // set_bayer_pattern (c_frame_accumulation.cc:1136-1142): SSG_COLORID_BAYER_*
  bool set_bayer_pattern(int colorid) { return a_ && ssg_accumulation_set_bayer_pattern(a_, colorid) == SSG_OK; }
  // set_remap (c_frame_accumulation.cc:1149-1152); an empty image clears it
  bool set_remap(const image &rmap) {
    return a_ && ssg_accumulation_set_remap(a_, rmap.empty() ? nullptr : &rmap) == SSG_OK;
  }
  bool set_input_bpp(int bpp) { return a_ && ssg_accumulation_set_input_bpp(a_, bpp) == SSG_OK; }
  // integer per-cell sample counts (SSG_32SC3)
  bool get_counts(image &counts) const { return a_ && ssg_accumulation_get_counts(a_, &counts) == SSG_OK; }
  // set_remap(current_remap) + add(raw, mask) for n registered device frames, in order
  bool add_registered_batch(const std::vector<image> &frames, const std::vector<registration_status> &st,
                            int motion_type) {
    if (!a_ || st.size() != frames.size()) return false;
    std::vector<ssg_image> f(frames.begin(), frames.end());
    return ssg_accumulation_add_registered_batch(a_, (int)f.size(), f.data(), st.data(), motion_type, nullptr,
                                                 nullptr, SSG_INTER_LINEAR, SSG_BORDER_REFLECT101, 0) == SSG_OK;
  }
};

using focus_stacking_options = ssg_focus_stacking_options;
inline focus_stacking_options default_focus_stacking_options() {
  focus_stacking_options o;
  ssg_default_focus_stacking_options(&o);
  return o;
}

// c_laplacian_pyramid_focus_stacking (c_frame_accumulation.h:141-184, c_frame_accumulation.cc:449-700)
class c_laplacian_pyramid_focus_stacking : public accumulation_handle {
 public:
  c_laplacian_pyramid_focus_stacking(int width, int height, int channels,
                                     const focus_stacking_options &o = default_focus_stacking_options())
      : accumulation_handle(SSG_ACC_FOCUS_STACK, width, height, channels) {
    if (a_ && ssg_accumulation_set_focus_options(a_, &o) != SSG_OK) {
      ssg_accumulation_destroy(a_);
      a_ = nullptr;
    }
  }
};

using eccflow_options = ssg_eccflow_options;
inline eccflow_options default_eccflow_options() {
  eccflow_options o;
  ssg_default_eccflow_options(&o);
  return o;
}

// c_eccflow (ecc2.h:548-662, ecc2.cc:2103-2865): dense coarse-to-fine ECC flow
class c_eccflow {
 public:
  explicit c_eccflow(const eccflow_options &o = default_eccflow_options()) { ssg_eccflow_create(&o, &h_); }
  ~c_eccflow() {
    if (h_) ssg_eccflow_destroy(h_);
  }
  c_eccflow(const c_eccflow &) = delete;
  c_eccflow &operator=(const c_eccflow &) = delete;
  bool valid() const { return h_ != nullptr; }
  // set_reference_image (ecc2.cc:2489-2672)
  bool set_reference_image(const image &img, const image &mask = image()) {
    return h_ && ssg_eccflow_set_reference_image(h_, &img, mask.empty() ? nullptr : &mask) == SSG_OK;
  }
  // compute(input, rmap, mask) (ecc2.cc:2856-2865): rmap is the initial remap (empty: the zero
  // flow) and receives identity + flow
  bool compute(const image &input, image &rmap, const image &mask = image(), bool rmap_is_initial = true) {
    return h_ && ssg_eccflow_compute(h_, &input, mask.empty() ? nullptr : &mask, rmap_is_initial ? &rmap : nullptr,
                                     &rmap) == SSG_OK;
  }
  // batched compute over device frames; initial remaps optional (then the zero flow)
  bool compute_batch(const std::vector<image> &inputs, std::vector<image> &rmaps, bool rmaps_are_initial = true,
                     const std::vector<image> *masks = nullptr) {
    if (!h_ || rmaps.size() != inputs.size()) return false;
    std::vector<ssg_image> in(inputs.begin(), inputs.end()), r(rmaps.begin(), rmaps.end()), m;
    if (masks) m.assign(masks->begin(), masks->end());
    return ssg_eccflow_compute_batch(h_, (int)in.size(), in.data(), masks ? m.data() : nullptr,
                                     rmaps_are_initial ? r.data() : nullptr, r.data()) == SSG_OK;
  }
  int num_levels() const {
    int32_t n = 0;
    if (h_) ssg_eccflow_num_levels(h_, &n);
    return n;
  }
  ssg_eccflow *handle() const { return h_; }

 private:
  ssg_eccflow *h_ = nullptr;
};

// FFT phase-correlation translation estimate (the pre-alignment of the turbulence and Bayer
// configurations; the reference has none on this path).
class phase_correlation {
 public:
  phase_correlation(int width, int height, int binning = 1) { ssg_phasecorr_create(width, height, binning, &h_); }
  ~phase_correlation() {
    if (h_) ssg_phasecorr_destroy(h_);
  }
  phase_correlation(const phase_correlation &) = delete;
  phase_correlation &operator=(const phase_correlation &) = delete;
  bool valid() const { return h_ != nullptr; }
  bool set_reference(const image &ref) { return h_ && ssg_phasecorr_set_reference(h_, &ref) == SSG_OK; }
  // per frame (Tx, Ty, peak) with ref(x) = frame(x + T)
  bool estimate(const std::vector<image> &frames, std::vector<double> &out) {
    if (!h_) return false;
    out.resize(3 * frames.size());
    std::vector<ssg_image> f(frames.begin(), frames.end());
    return ssg_phasecorr_estimate(h_, (int)f.size(), f.data(), out.data()) == SSG_OK;
  }

 private:
  ssg_phasecorr *h_ = nullptr;
};

// lpg() sharpness map (core/proc/lpg.cc:277-344) and its mean
inline bool lpg(const image &src, image &dst, double k = 6, double p = 2, double *sharpness = nullptr) {
  return ssg_lpg(&src, &dst, k, p, sharpness) == SSG_OK;
}

// compute_local_variance_map (c_local_variance_sharpness_measure.cc:193-249)
inline bool local_variance_map(const image &src, image &map, int kradius = 1, int dscale = 1, double *q = nullptr) {
  return ssg_local_variance_map(&src, kradius, dscale, &map, q) == SSG_OK;
}

// Per-frame scalar weights of the C1 stacking configuration (mean squared laplacian).
inline bool frame_sharpness(const std::vector<image> &frames, std::vector<float> &weights) {
  weights.resize(frames.size());
  std::vector<ssg_image> f(frames.begin(), frames.end());
  return ssg_frame_sharpness((int)f.size(), f.data(), weights.data()) == SSG_OK;
}

// c_ser_reader (core/io/c_ser_file.h:136-215): header, seek, per-frame reads to host or device
// memory, batched reads to device memory.
class c_ser_reader {
 public:
  c_ser_reader() = default;
  explicit c_ser_reader(const std::string &path) { open(path); }
  ~c_ser_reader() { close(); }
  c_ser_reader(const c_ser_reader &) = delete;
  c_ser_reader &operator=(const c_ser_reader &) = delete;
  bool open(const std::string &path) {
    close();
    return ssg_ser_open(path.c_str(), &h_) == SSG_OK;
  }
  void close() {
    if (h_) ssg_ser_close(h_);
    h_ = nullptr;
  }
  bool is_open() const { return h_ != nullptr; }
  // the header as stored (is_little_endian as the reference interprets it) and the frame type
  bool header(ssg_ser_header &hdr, int *frame_type = nullptr) const {
    int32_t t = 0;
    const bool ok = ssg_ser_get_header(h_, &hdr, &t) == SSG_OK;
    if (frame_type) *frame_type = t;
    return ok;
  }
  int num_frames() const {
    ssg_ser_header hdr{};
    return header(hdr) ? hdr.frames_count : 0;
  }
  bool seek(int index) { return ssg_ser_seek(h_, index) == SSG_OK; }
  int curpos() const {
    int32_t p = -1;
    return ssg_ser_curpos(h_, &p) == SSG_OK ? p : -1;
  }
  bool read(image &frame) { return ssg_ser_read(h_, &frame) == SSG_OK; }
  bool read(std::vector<image> &frames) {
    std::vector<ssg_image> v(frames.begin(), frames.end());
    return ssg_ser_read_batch(h_, (int)v.size(), v.data()) == SSG_OK;
  }
  bool timestamps(std::vector<uint64_t> &ts) const {
    ts.resize((size_t)num_frames());
    int32_t count = 0;
    if (ssg_ser_timestamps(h_, ts.data(), (int32_t)ts.size(), &count) != SSG_OK) return false;
    ts.resize((size_t)count);
    return true;
  }

 private:
  ssg_ser_reader *h_ = nullptr;
};

// debayer(src, dst, colorid, algo) (core/io/debayer.cc:1197-1355)
inline bool debayer(const image &src, image &dst, int colorid, int algorithm = SSG_DEBAYER_DEFAULT) {
  return ssg_debayer(&src, colorid, algorithm, &dst) == SSG_OK;
}

// read_input_frame's calibration, debayer and conversion of raw frames
// (c_image_stacking_pipeline_base.cc:124-279), a batch at a time
inline bool read_input_frames(const std::vector<image> &raw, std::vector<image> &out, int colorid, int bpp,
                              int algorithm = SSG_DEBAYER_DEFAULT, const image *darkbayer = nullptr,
                              const image *flatbayer = nullptr) {
  std::vector<ssg_image> r(raw.begin(), raw.end()), o(out.begin(), out.end());
  return ssg_read_input_frames((int)r.size(), r.data(), colorid, bpp, algorithm, darkbayer, flatbayer, o.data()) ==
         SSG_OK;
}

// select_master_frame's per-frame metric (c_image_stacking_pipeline_base.cc:350-359)
inline bool master_frame_quality(const std::vector<image> &frames, int colorid, int kradius, int dscale,
                                 std::vector<double> &quality) {
  quality.resize(frames.size());
  std::vector<ssg_image> f(frames.begin(), frames.end());
  return ssg_master_frame_quality((int)f.size(), f.data(), colorid, kradius, dscale, quality.data()) == SSG_OK;
}
// master_frame_best_of_100_in_middle's scan window (c_image_stacking_pipeline_base.cc:312-326)
inline std::pair<int, int> master_scan_window(int size, int max_frames_to_scan) {
  int32_t a = 0, b = 0;
  ssg_master_scan_window(size, max_frames_to_scan, &a, &b);
  return {a, b};
}
// the selection rule (c_image_stacking_pipeline_base.cc:369-395): a global frame index
inline int select_master_frame(const std::vector<double> &quality, int start_pos) {
  int32_t i = start_pos;
  ssg_select_master_frame((int)quality.size(), quality.data(), start_pos, &i);
  return i;
}
// create_reference_frame's stacking window (c_image_stacking_pipeline.cc:1219-1232)
inline std::pair<int, int> master_stack_window(int size, int start_frame_index, int master_frame_pos,
                                               int max_frames_to_stack) {
  int32_t a = 0, b = 0;
  ssg_master_stack_window(size, start_frame_index, master_frame_pos, max_frames_to_stack, &a, &b);
  return {a, b};
}

// linear_interpolation_inpaint(image, mask, image) (linear_interpolation_inpaint.cc:335-361), in place
inline bool linear_interpolation_inpaint(image &img, const image &mask) {
  return ssg_inpaint_linear(&img, &mask) == SSG_OK;
}
// average_pyramid_inpaint(src, mask, dst, dstmask, max_levels) (average_pyramid_inpaint.cc:96-126)
inline bool average_pyramid_inpaint(const image &src, const image *mask, image &dst, image *dstmask,
                                    int max_levels = 100) {
  return ssg_average_pyramid_inpaint(&src, mask, &dst, dstmask, max_levels) == SSG_OK;
}
// unsharp_mask(src, dst, sigma, alpha) (core/proc/unsharp_mask.cc:72-115)
inline bool unsharp_mask(const image &src, image &dst, double sigma, double alpha) {
  return ssg_unsharp_mask(&src, &dst, sigma, alpha) == SSG_OK;
}

// upscale_remap(option, srcmap, dstmap) (c_image_stacking_pipeline.cc:1869-1905) of a batch of remaps
inline std::pair<int, int> upscale_size(int option, int w, int h) {
  int32_t a = w, b = h;
  ssg_upscale_size(option, w, h, &a, &b);
  return {a, b};
}
inline bool upscale_remap(int option, const std::vector<image> &src, std::vector<image> &dst) {
  std::vector<ssg_image> s(src.begin(), src.end()), d(dst.begin(), dst.end());
  return ssg_upscale_remaps(option, (int)s.size(), s.data(), d.data()) == SSG_OK;
}

}  // namespace ssg

#endif  // SSG_HPP

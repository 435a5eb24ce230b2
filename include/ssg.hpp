/*
 * ssg.hpp -- header-only C++ host interface over the C ABI (ssg.h), with the
 * class and method names of SerStacker's registration and accumulation
 * interfaces, so pipeline code binds to it one call for one call:
 *
 *   ssg::c_frame_registration   core/proc/image_registration/c_frame_registration.h:210-354
 *   ssg::c_frame_accumulation   core/average/c_frame_accumulation.h:14-37
 *   ssg::c_weigthed_average     core/average/c_frame_accumulation.h:39-63
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
  virtual void clear() = 0;
  virtual std::pair<int, int> accumulator_size() const = 0;
  virtual int accumulated_frames() const = 0;
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

 private:
  int w_, h_, cn_;
  ssg_accumulation *a_ = nullptr;
};

// Per-frame scalar weights of the C1 stacking configuration (mean squared laplacian).
inline bool frame_sharpness(const std::vector<image> &frames, std::vector<float> &weights) {
  weights.resize(frames.size());
  std::vector<ssg_image> f(frames.begin(), frames.end());
  return ssg_frame_sharpness((int)f.size(), f.data(), weights.data()) == SSG_OK;
}

}  // namespace ssg

#endif  // SSG_HPP

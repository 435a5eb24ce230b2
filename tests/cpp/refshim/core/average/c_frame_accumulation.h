// The members of the reference's accumulation interface that the shim in INTEGRATION.md uses,
// with the reference's signatures (a compile-time fixture for tests/test_abi.py):
//   c_frame_accumulation                core/average/c_frame_accumulation.h:14-37
//   c_weigthed_average (final methods)  c_frame_accumulation.h:39-63
//   c_laplacian_pyramid_focus_stacking  c_frame_accumulation.h:141-184
//   c_bayer_average (final methods)     c_frame_accumulation.h:222-262
//   COLORID                             core/io/debayer.h
#pragma once
#include <opencv2/opencv.hpp>

enum COLORID { COLORID_MONO = 0, COLORID_BAYER_RGGB = 8, COLORID_BAYER_GRBG = 9, COLORID_BAYER_GBRG = 10,
               COLORID_BAYER_BGGR = 11 };

class c_frame_accumulation {
 public:
  typedef std::shared_ptr<c_frame_accumulation> ptr;
  virtual ~c_frame_accumulation() = default;
  virtual bool add(cv::InputArray src, cv::InputArray mask = cv::noArray()) = 0;
  virtual bool compute(cv::OutputArray avg, cv::OutputArray mask = cv::noArray(), double dscale = 1.0,
                       int ddepth = -1) const = 0;
  virtual bool get_acc_counters(cv::Mat& accw) const = 0;
  virtual bool reinitialize(cv::InputArray src, cv::InputArray accw) = 0;
  virtual void clear() = 0;
  virtual cv::Size accumulator_size() const = 0;
  int accumulated_frames() const { return _accumulated_frames; }

 protected:
  int _accumulated_frames = 0;
};

class c_weigthed_average : public c_frame_accumulation {
 public:
  c_weigthed_average();
  bool add(cv::InputArray src, cv::InputArray weights = cv::noArray()) final;
  bool compute(cv::OutputArray avg, cv::OutputArray mask = cv::noArray(), double dscale = 1.0,
               int ddepth = -1) const final;
  bool get_acc_counters(cv::Mat& accw) const final;
  bool reinitialize(cv::InputArray src, cv::InputArray accw) final;
  void clear() final;
  cv::Size accumulator_size() const final;
};

class c_laplacian_pyramid_focus_stacking : public c_frame_accumulation {
 public:
  enum fusing_policy { select_max_energy, weighted_average };
  struct options {
    enum fusing_policy fusing_policy = select_max_energy;
    bool inpaint_mask_holes = true;
    bool avgchannel = true;
    int kradius = 0;
    double ksigma = 0;
  };
  c_laplacian_pyramid_focus_stacking(const options& opts);
};

class c_bayer_average : public c_frame_accumulation {
 public:
  void set_bayer_pattern(COLORID colorid);
  void set_remap(const cv::Mat2f& rmap);
  bool add(cv::InputArray src, cv::InputArray weights = cv::noArray()) final;
  bool compute(cv::OutputArray avg, cv::OutputArray mask = cv::noArray(), double dscale = 1.0,
               int ddepth = -1) const final;
  bool get_acc_counters(cv::Mat& accw) const final;
  bool reinitialize(cv::InputArray src, cv::InputArray accw) final;
  void clear() final;
  cv::Size accumulator_size() const final;
};

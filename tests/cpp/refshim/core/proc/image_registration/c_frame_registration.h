// The members of the reference's registration interface that the shim in INTEGRATION.md uses,
// with the reference's signatures (a compile-time fixture for tests/test_abi.py; the reference
// headers themselves pull OpenCV, features2d, ctrlbind and libconfig through their includes).
//   enums          core/proc/image_registration/image_transform.h:15-26, ecc2.h:32-64,507-511,
//                  core/proc/extract_channel.h:13-50
//   option structs c_frame_registration.h:47-64, 88-100, 119-136
//   the class      c_frame_registration.h:210-354 (virtual ops, current_remap(), protected state)
#pragma once
#include <opencv2/opencv.hpp>

enum IMAGE_MOTION_TYPE { IMAGE_MOTION_TRANSLATION, IMAGE_MOTION_EUCLIDEAN, IMAGE_MOTION_SCALED_EUCLIDEAN,
                         IMAGE_MOTION_AFFINE };
enum ECC_ALIGN_METHOD { ECC_ALIGN_FORWARD_ADDITIVE, ECC_ALIGN_INVERSE_COMPOSITIONAL, ECC_ALIGN_LM,
                        ECC_ALIGN_INVERSE_COMPOSITIONAL_LM };
enum ECC_INTERPOLATION_METHOD { ECC_INTER_LINEAR = 1, ECC_INTER_UNKNOWN = -1 };
enum ECC_BORDER_MODE { ECC_BORDER_CONSTANT = 0, ECC_BORDER_REFLECT101 = 4, ECC_BORDER_UNKNOWN = -1 };
enum ECCFlowDownscaleMethod { ECCFlowDownscaleRecursiveResize, ECCFlowDownscaleFullResize, ECCFlowDownscalePyramid };
enum color_channel_type { color_channel_gray = 104 };

struct c_feature_registration_options {
  double scale = 0.5;
};

struct c_ecc_registration_options {
  double scale = 0.5;
  double eps = 0.2;
  double min_rho = 0.8;
  double input_smooth_sigma = 1.0;
  double reference_smooth_sigma = 1.0;
  double update_step_scale = 1.5;
  int se_radius = 5;
  ECC_ALIGN_METHOD ecc_method = ECC_ALIGN_LM;
  int max_iterations = 50;
  int ecch_max_level = 0;
  int ecch_minimum_image_size = 16;
  double normalization_noise = 0.01;
  int normalization_scale = 0;
  bool ecch_estimate_translation_first = true;
  bool replace_planetary_disk_with_mask = false;
};

struct c_eccflow_registration_options {
  double update_multiplier = 1.5;
  double input_smooth_sigma = 0;
  double reference_smooth_sigma = 0;
  double noise_level = -1;
  double scale_factor = 0.75;
  int max_iterations = 3;
  int support_scale = 4;
  int min_image_size = -1;
  int max_pyramid_level = -1;
  ECCFlowDownscaleMethod downscale_method = ECCFlowDownscaleRecursiveResize;
};

struct c_image_registration_options {
  IMAGE_MOTION_TYPE motion_type = IMAGE_MOTION_AFFINE;
  color_channel_type ecc_registration_channel = color_channel_gray;
  enum ECC_INTERPOLATION_METHOD interpolation = ECC_INTER_LINEAR;
  enum ECC_BORDER_MODE border_mode = ECC_BORDER_REFLECT101;
  cv::Scalar border_value = cv::Scalar(0, 0, 0);
  struct c_feature_registration_options feature_registration;
  struct c_ecc_registration_options ecc;
  struct c_eccflow_registration_options eccflow;
  bool enable_feature_registration = true;
  bool enable_ecc_registration = false;
  bool enable_eccflow_registration = false;
  bool accumulate_and_compensate_turbulent_flow = false;
};

class c_frame_registration {
 public:
  typedef std::shared_ptr<c_frame_registration> sptr;
  c_frame_registration();
  c_frame_registration(const c_image_registration_options& options);
  virtual ~c_frame_registration() = default;
  c_image_registration_options& options();
  virtual bool setup_reference_frame(cv::InputArray image, cv::InputArray msk = cv::noArray());
  virtual bool create_image_transfrom();
  virtual bool register_frame(cv::InputArray src, cv::InputArray srcmask, cv::OutputArray dst = cv::noArray(),
                              cv::OutputArray dstmask = cv::noArray());
  virtual bool remap(cv::InputArray src, cv::OutputArray dst, cv::InputArray src_mask = cv::noArray(),
                     cv::OutputArray dst_mask = cv::noArray(),
                     enum ECC_INTERPOLATION_METHOD interpolation_method = ECC_INTER_UNKNOWN,
                     enum ECC_BORDER_MODE border_mode = ECC_BORDER_UNKNOWN,
                     const cv::Scalar& border_value = cv::Scalar()) const;
  virtual bool custom_remap(const cv::Mat2f& rmap, cv::InputArray src, cv::OutputArray dst,
                            cv::InputArray src_mask = cv::noArray(), cv::OutputArray dst_mask = cv::noArray(),
                            enum ECC_INTERPOLATION_METHOD interpolation_method = ECC_INTER_UNKNOWN,
                            enum ECC_BORDER_MODE border_mode = ECC_BORDER_UNKNOWN,
                            const cv::Scalar& border_value = cv::Scalar()) const;
  const cv::Mat2f& current_remap() const;

 protected:
  c_image_registration_options _options;
  cv::Size _reference_frame_size;
  cv::Size _current_frame_size;
  cv::Mat2f _current_remap;
};

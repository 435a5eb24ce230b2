// host_api_demo.cpp -- plain C++ caller of the library through include/ssg.hpp,
// written the way the stacking pipeline calls c_frame_registration and
// c_weigthed_average (c_image_stacking_pipeline.cc:1539-1776): set the
// reference, register every frame (per frame, then batched), warp + stack,
// compute; then the same frames refined by c_eccflow and stacked with their
// dense remap fields.  Prints one JSON line; the GPU test compares it with the Python
// binding's results on the same seeded frames.
//
//   host_api_demo <width> <height> <frames>
//
// Exit status: 0 ok, 77 no GPU, 1 failure.
#include <hip/hip_runtime_api.h>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "ssg.hpp"

#define CHECK(cond)                                                                   \
  do {                                                                                \
    if (!(cond)) {                                                                    \
      std::fprintf(stderr, "FAILED %s (line %d): %s\n", #cond, __LINE__,              \
                   ssg::last_error().c_str());                                        \
      return 1;                                                                       \
    }                                                                                 \
  } while (0)

int main(int argc, char** argv) {
  const int W = argc > 1 ? std::atoi(argv[1]) : 320;
  const int H = argc > 2 ? std::atoi(argv[2]) : 240;
  const int N = argc > 3 ? std::atoi(argv[3]) : 6;
  int ndev = 0;
  char name[256];
  if (ssg_device_info(&ndev, name, sizeof(name)) != SSG_OK || ndev <= 0) {
    std::fprintf(stderr, "no GPU: %s\n", ssg::last_error().c_str());
    return 77;
  }

  // seeded synthetic frames in device memory (the same generator as the Python side)
  ssg_synth_params p{};
  p.width = W;
  p.height = H;
  p.disk_radius = 0.38 * (W < H ? W : H);
  p.texture_sigma = 3.0;
  p.texture_amplitude = 0.1;
  p.limb_darkening = 0.6;
  p.max_shift = 6.0;
  p.max_rotation_deg = 0.5;
  p.max_scale_dev = 0.003;
  p.max_shear = 0.002;
  p.blur_sigma_min = 0.8;
  p.blur_sigma_max = 2.0;
  p.peak_electrons = 2000;
  p.read_noise = 4;
  p.output_scale = 1.0 / 65536;
  p.seed = 4242;

  const size_t plane = (size_t)W * H;
  float* dev = nullptr;
  CHECK(hipMalloc((void**)&dev, sizeof(float) * plane * (N + 1)) == hipSuccess);
  std::vector<ssg::image> frames;
  for (int i = 0; i < N; ++i) frames.push_back(ssg::image::f32(dev + plane * (i + 1), W, H, true));
  ssg::image ref = ssg::image::f32(dev, W, H, true);
  std::vector<double> truth(6 * N);
  CHECK(ssg_synth_render(&p, -1, 1, &ref, nullptr) == SSG_OK);
  CHECK(ssg_synth_render(&p, 0, N, frames.data(), truth.data()) == SSG_OK);

  ssg::registration_options o = ssg::default_registration_options();
  o.motion_type = SSG_MOTION_AFFINE;
  o.enable_feature_registration = 0;  // the reference defaults start from feature registration
  o.enable_ecc_registration = 1;
  o.ecc.ecc_method = SSG_ECC_ALIGN_FORWARD_ADDITIVE;
  o.ecc.scale = 1.0;
  o.ecc.eps = 0.02;
  o.ecc.ecch_max_level = 3;
  o.ecc.max_iterations = 50;
  ssg::c_frame_registration reg(o);
  CHECK(reg.setup_reference_frame(ref));

  // per-frame form, then the batched form: identical results
  std::vector<ssg::registration_status> plain(N), batch;
  for (int i = 0; i < N; ++i) {
    CHECK(reg.register_frame(frames[i]));
    plain[i] = reg.status();
  }
  CHECK(reg.register_batch(frames, batch));
  int same = 1;
  for (int i = 0; i < N; ++i)
    for (int k = 0; k < 6; ++k) same &= plain[i].params[k] == batch[i].params[k];

  // current_remap of the last frame vs its matrix
  std::vector<float> map(plane * 2);
  ssg::image mview = ssg::image::map2f(map.data(), W, H, false);
  CHECK(reg.current_remap(mview));
  const double* m = plain[N - 1].matrix;
  double map_err = 0;
  for (int y = 0; y < H; y += 17)
    for (int x = 0; x < W; x += 13) {
      const double ex = m[0] * x + m[1] * y + m[2], ey = m[3] * x + m[4] * y + m[5];
      map_err = std::fmax(map_err, std::fabs(map[2 * ((size_t)y * W + x)] - ex));
      map_err = std::fmax(map_err, std::fabs(map[2 * ((size_t)y * W + x) + 1] - ey));
    }

  // weights, fused warp + stack, compute
  std::vector<float> wts;
  CHECK(ssg::frame_sharpness(frames, wts));
  ssg::c_weigthed_average acc(W, H, 1);
  CHECK(acc.valid());
  CHECK(acc.add_registered_batch(frames, batch, SSG_MOTION_AFFINE, &wts));
  std::vector<float> avg(plane);
  std::vector<uint8_t> mask(plane);
  ssg::image aview = ssg::image::f32(avg.data(), W, H, false), kview = ssg::image::u8(mask.data(), W, H, false);
  CHECK(acc.compute(aview, &kview));
  double sum = 0;
  long covered = 0;
  for (size_t i = 0; i < plane; ++i) {
    sum += avg[i];
    covered += mask[i] != 0;
  }

  // dense refinement: the transform fields of the batch, refined by c_eccflow, then the fused
  // warp + stack with the per-frame fields
  float* maps = nullptr;
  CHECK(hipMalloc((void**)&maps, sizeof(float) * 2 * plane * N) == hipSuccess);
  std::vector<ssg::image> rmaps;
  for (int i = 0; i < N; ++i) rmaps.push_back(ssg::image::map2f(maps + 2 * plane * i, W, H, true));
  std::vector<ssg::registration_status> st2;
  CHECK(reg.register_batch_remaps(frames, st2, rmaps));
  ssg::eccflow_options fo = ssg::default_eccflow_options();
  fo.downscale = SSG_ECCFLOW_DOWNSCALE_PYRAMID;
  fo.support_scale = 3;
  fo.max_iterations = 2;
  fo.min_image_size = -1;
  fo.max_pyramid_level = 3;
  ssg::c_eccflow flow(fo);
  CHECK(flow.valid());
  CHECK(flow.set_reference_image(ref));
  CHECK(flow.compute_batch(frames, rmaps));
  ssg::c_weigthed_average acc2(W, H, 1);
  CHECK(acc2.add_remapped_batch(frames, rmaps, &wts));
  CHECK(acc2.compute(aview, &kview));
  double fsum = 0;
  long fcovered = 0;
  for (size_t i = 0; i < plane; ++i) {
    fsum += avg[i];
    fcovered += mask[i] != 0;
  }

  // stacking at x1.5 after align: the refined remaps upscaled, the frames warped onto the larger grid
  const auto usz = ssg::upscale_size(SSG_UPSCALE_X15, W, H);
  const size_t uplane = (size_t)usz.first * usz.second;
  float* umaps = nullptr;
  CHECK(hipMalloc((void**)&umaps, sizeof(float) * 2 * uplane * N) == hipSuccess);
  std::vector<ssg::image> urmaps;
  for (int i = 0; i < N; ++i) urmaps.push_back(ssg::image::map2f(umaps + 2 * uplane * i, usz.first, usz.second, true));
  CHECK(ssg::upscale_remap(SSG_UPSCALE_X15, rmaps, urmaps));
  ssg::c_weigthed_average acc3(usz.first, usz.second, 1);
  CHECK(acc3.add_remapped_batch(frames, urmaps, &wts));
  std::vector<float> uavg(uplane);
  std::vector<uint8_t> umask(uplane);
  ssg::image uav = ssg::image::f32(uavg.data(), usz.first, usz.second, false);
  ssg::image ukv = ssg::image::u8(umask.data(), usz.first, usz.second, false);
  CHECK(acc3.compute(uav, &ukv));
  double usum = 0;
  long ucovered = 0;
  for (size_t i = 0; i < uplane; ++i) {
    usum += uavg[i];
    ucovered += umask[i] != 0;
  }
  hipFree(umaps);
  hipFree(maps);

  // the finish: average_pyramid_inpaint of the x1.5 stack with its mask
  std::vector<float> fin(uplane);
  std::vector<uint8_t> finm(uplane);
  ssg::image fv = ssg::image::f32(fin.data(), usz.first, usz.second, false);
  ssg::image fmv = ssg::image::u8(finm.data(), usz.first, usz.second, false);
  CHECK(ssg::average_pyramid_inpaint(uav, &ukv, fv, &fmv, 100));
  double apr_sum = 0;
  for (size_t i = 0; i < uplane; ++i) apr_sum += fin[i];

  // master-frame scan over the frames (mono): quality and selection
  std::vector<double> quality;
  CHECK(ssg::master_frame_quality(frames, SSG_COLORID_MONO, 1, 1, quality));
  const auto win = ssg::master_scan_window(N, 3);
  const int master = ssg::select_master_frame(quality, 0);

  std::printf("{\"up_avg_sum\": %.17g, \"up_covered\": %ld, \"apr_sum\": %.17g, \"master\": %d, "
              "\"scan\": [%d, %d], \"quality\": [",
              usum, ucovered, apr_sum, master, win.first, win.second);
  for (int i = 0; i < N; ++i) std::printf("%s%.17g", i ? ", " : "", quality[i]);
  std::printf("], ");
  std::printf("\"device\": \"%s\", \"frames\": %d, \"plain_equals_batch\": %d, \"accumulated\": %d, "
              "\"map_err\": %.3g, \"avg_sum\": %.17g, \"covered\": %ld, \"flow_levels\": %d, "
              "\"flow_avg_sum\": %.17g, \"flow_covered\": %ld, \"params0\": [%.9g, %.9g, %.9g, %.9g, %.9g, "
              "%.9g], \"iterations\": [",
              name, N, same, acc.accumulated_frames(), map_err, sum, covered, flow.num_levels(), fsum, fcovered,
              batch[0].params[0], batch[0].params[1], batch[0].params[2], batch[0].params[3], batch[0].params[4],
              batch[0].params[5]);
  for (int i = 0; i < N; ++i) std::printf("%s%d", i ? ", " : "", batch[i].iterations);
  std::printf("]}\n");
  hipFree(dev);
  return 0;
}

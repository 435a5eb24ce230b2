// A minimal stand-in for the few OpenCV types the reference-side shim in INTEGRATION.md touches
// (cv::Mat and its typed forms, InputArray/OutputArray, Size, Scalar, the type constants).  It is
// a compile-time fixture for tests/test_abi.py: only declarations with OpenCV's signatures, no
// behaviour beyond what a syntax check needs.  It is not used to build or run reference code.
#pragma once
#include <cstddef>
#include <cstdint>
#include <memory>

typedef unsigned char uchar;

#define CV_8U 0
#define CV_16U 2
#define CV_32S 4
#define CV_32F 5
#define CV_MAKETYPE(depth, cn) ((depth) + (((cn)-1) << 3))
#define CV_8UC1 CV_MAKETYPE(CV_8U, 1)
#define CV_16UC1 CV_MAKETYPE(CV_16U, 1)
#define CV_32SC1 CV_MAKETYPE(CV_32S, 1)
#define CV_32SC3 CV_MAKETYPE(CV_32S, 3)
#define CV_32FC1 CV_MAKETYPE(CV_32F, 1)
#define CV_32FC2 CV_MAKETYPE(CV_32F, 2)
#define CV_32FC3 CV_MAKETYPE(CV_32F, 3)

namespace cv {
struct Size {
  int width = 0, height = 0;
  Size() = default;
  Size(int w, int h) : width(w), height(h) {}
  bool operator!=(const Size& o) const { return width != o.width || height != o.height; }
};
struct Scalar {
  double val[4] = {0, 0, 0, 0};
  Scalar() = default;
  Scalar(double v0, double v1 = 0, double v2 = 0, double v3 = 0) : val{v0, v1, v2, v3} {}
  double operator[](int i) const { return val[i]; }
};
class Mat {
 public:
  uchar* data = nullptr;
  int rows = 0, cols = 0;
  size_t step = 0;
  Mat() = default;
  Mat(int r, int c, int t);
  int type() const;
  int depth() const;
  int channels() const;
  bool empty() const;
  Size size() const;
  void create(int r, int c, int t);
  void create(Size s, int t);
  void release();
  void copyTo(const class _OutputArray& dst) const;
  void convertTo(const class _OutputArray& dst, int rtype, double alpha = 1, double beta = 0) const;
};
template <class T>
class Mat_ : public Mat {
 public:
  void create(Size s);
  void create(int r, int c);
};
template <class T, int n>
struct Vec {
  T v[n];
};
typedef Vec<float, 2> Vec2f;
typedef Mat_<float> Mat1f;
typedef Mat_<Vec2f> Mat2f;
class _InputArray {
 public:
  _InputArray();
  _InputArray(const Mat& m);
  Mat getMat() const;
  bool empty() const;
  Size size() const;
  int channels() const;
};
class _OutputArray : public _InputArray {
 public:
  _OutputArray();
  _OutputArray(Mat& m);
  bool needed() const;
  bool fixedType() const;
  int depth() const;
  void create(Size s, int t) const;
  void create(int r, int c, int t) const;
};
typedef const _InputArray& InputArray;
typedef const _OutputArray& OutputArray;
const _OutputArray& noArray();
}  // namespace cv

// Small value types standing in for the OpenCV / Eigen types of the reference's StereoMapper
// interface (motion_stereo/stereo_mapper.h:26-63): the GPU host has neither library.
//   pm::Mat       ~ cv::Mat        (rows, cols, type, data, step, at<T>(i, j), clone, setTo)
//   pm::DeviceMat ~ cv::cuda::GpuMat (rows, cols, type, data, step, download, setTo)
//   pm::Matrix3d / pm::Vector3d ~ Eigen::Matrix3d / Eigen::Vector3d
#ifndef PM_TYPES_H
#define PM_TYPES_H

#include <cstddef>
#include <cstdint>
#include <cstring>
#include <memory>
#include <vector>

namespace pm {

// element type codes (the cv::Mat depth codes the reference uses)
enum { CV_8U = 0, CV_32S = 4, CV_32F = 5, CV_64F = 6, PM_64U = 7, PM_16F = 8, PM_16BF = 9 };

inline size_t elem_size(int type) {
  switch (type) {
    case CV_8U: return 1;
    case CV_32S: case CV_32F: return 4;
    case CV_64F: case PM_64U: return 8;
    case PM_16F: case PM_16BF: return 2;
    default: return 0;
  }
}

class Mat {
 public:
  int rows = 0, cols = 0, type = CV_8U;
  size_t step = 0;          // bytes per row
  uint8_t* data = nullptr;

  Mat() = default;
  Mat(int r, int c, int t) { create(r, c, t); }
  // wraps external memory without copying (like cv::Mat(rows, cols, type, data))
  Mat(int r, int c, int t, void* ext) : rows(r), cols(c), type(t), step(c * elem_size(t)),
                                         data(static_cast<uint8_t*>(ext)) {}
  // (re)allocates only if the shape or type differs, like cv::Mat::create
  void create(int r, int c, int t) {
    if (data && rows == r && cols == c && type == t) return;
    rows = r; cols = c; type = t; step = c * elem_size(t);
    auto v = std::make_shared<std::vector<uint8_t>>(step * r);
    data = v->data();
    store_ = std::shared_ptr<uint8_t>(v, data);
    alloc_tag_ = nullptr;
  }
  // create with storage from alloc (freed by release when the last Mat sharing it goes);
  // tag identifies the allocator (the mapper uses it for page-locked result matrices)
  void create_with(int r, int c, int t, uint8_t* (*alloc)(size_t), void (*release)(uint8_t*),
                   const void* tag) {
    if (data && rows == r && cols == c && type == t && alloc_tag_ == tag) return;
    const size_t st = c * elem_size(t);
    uint8_t* p = alloc(st * r);   // may throw: the Mat is unchanged then
    rows = r; cols = c; type = t; step = st;
    data = p;
    store_ = std::shared_ptr<uint8_t>(data, release);
    alloc_tag_ = tag;
  }
  const void* alloc_tag() const { return alloc_tag_; }
  Mat clone() const {
    Mat m(rows, cols, type);
    if (data) std::memcpy(m.data, data, step * rows);
    return m;
  }
  template <typename T> T& at(int i, int j) { return reinterpret_cast<T*>(data + i * step)[j]; }
  template <typename T> const T& at(int i, int j) const {
    return reinterpret_cast<const T*>(data + i * step)[j];
  }
  template <typename T> void setTo(T v) {
    for (int i = 0; i < rows; ++i)
      for (int j = 0; j < cols; ++j) at<T>(i, j) = v;
  }
  size_t total() const { return (size_t)rows * cols; }
  bool empty() const { return data == nullptr || rows == 0 || cols == 0; }

 private:
  std::shared_ptr<uint8_t> store_;
  const void* alloc_tag_ = nullptr;
};

// Device buffer owned by its creator (StereoMapper).  Dense rows: step = cols * elem size.
class DeviceMat {
 public:
  int rows = 0, cols = 0, type = CV_32F;
  size_t step = 0;
  uint8_t* data = nullptr;  // device pointer
  size_t bytes() const { return step * rows; }
  // copies to host (synchronous), like cv::cuda::GpuMat::download
  void download(Mat& m) const;
  // fills every element (float or int32 buffers), like cv::cuda::GpuMat::setTo
  void setTo(float v) const;
};

struct Vector3d {
  double v[3] = {0, 0, 0};
  Vector3d() = default;
  Vector3d(double a, double b, double c) : v{a, b, c} {}
  double& operator()(int i) { return v[i]; }
  double operator()(int i) const { return v[i]; }
  Vector3d operator-(const Vector3d& o) const {
    return Vector3d(v[0] - o.v[0], v[1] - o.v[1], v[2] - o.v[2]);
  }
};

struct Matrix3d {
  double m[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
  double& operator()(int i, int j) { return m[i * 3 + j]; }
  double operator()(int i, int j) const { return m[i * 3 + j]; }
  Matrix3d transpose() const {
    Matrix3d t;
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) t.m[i * 3 + j] = m[j * 3 + i];
    return t;
  }
  Matrix3d operator*(const Matrix3d& o) const {
    Matrix3d r;
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) {
        double s = 0;
        for (int k = 0; k < 3; ++k) s += m[i * 3 + k] * o.m[k * 3 + j];
        r.m[i * 3 + j] = s;
      }
    return r;
  }
  Vector3d operator*(const Vector3d& x) const {
    Vector3d r;
    for (int i = 0; i < 3; ++i) r.v[i] = m[i * 3] * x.v[0] + m[i * 3 + 1] * x.v[1] + m[i * 3 + 2] * x.v[2];
    return r;
  }
};

}  // namespace pm

#endif  // PM_TYPES_H

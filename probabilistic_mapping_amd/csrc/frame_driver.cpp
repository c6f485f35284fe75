// Keyframe / source selection driver (evaluate.cc's processImg without ROS) and its C ABI.
// The behaviour is described in include/pm/frame_driver.h.
#include "pm/frame_driver.h"

#include <cstring>
#include <iterator>
#include <new>
#include <string>

#include "common.hpp"

namespace pm {

Matrix3d quaternion_to_rotation(double qx, double qy, double qz, double qw) {
  // the rotation of a unit quaternion, with the doubled products formed first
  const double tx = 2.0 * qx, ty = 2.0 * qy, tz = 2.0 * qz;
  const double twx = tx * qw, twy = ty * qw, twz = tz * qw;
  const double txx = tx * qx, txy = ty * qx, txz = tz * qx;
  const double tyy = ty * qy, tyz = tz * qy, tzz = tz * qz;
  Matrix3d R;
  R(0, 0) = 1.0 - (tyy + tzz); R(0, 1) = txy - twz;         R(0, 2) = txz + twy;
  R(1, 0) = txy + twz;         R(1, 1) = 1.0 - (txx + tzz); R(1, 2) = tyz - twx;
  R(2, 0) = txz - twy;         R(2, 1) = tyz + twx;         R(2, 2) = 1.0 - (txx + tyy);
  return R;
}

namespace {

Mat mat33(const Matrix3d& R) {
  Mat m(3, 3, CV_64F);
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) m.at<double>(i, j) = R(i, j);
  return m;
}

Mat mat31(double x, double y, double z) {
  Mat m(3, 1, CV_64F);
  m.at<double>(0, 0) = x; m.at<double>(1, 0) = y; m.at<double>(2, 0) = z;
  return m;
}

}  // namespace

FrameDriver::FrameDriver(const DriverConfig& cfg, bool with_mapper)
    : cfg_(cfg), with_mapper_(with_mapper), last_size_(cfg.agg_num * cfg.interval) {
  if (cfg.agg_num < 0 || cfg.agg_num > PM_MAX_SOURCES || cfg.interval < 1 || cfg.planes < 2)
    throw Error(PM_ERR_INVALID_ARGUMENT,
                "FrameDriver: agg_num must be in 0..16, interval >= 1, planes >= 2");
}

FrameDriver::~FrameDriver() = default;

void FrameDriver::grab_image(int64_t stamp_ns, const Mat& image) {
  if (image.type != CV_8U || image.empty())
    throw Error(PM_ERR_INVALID_ARGUMENT, "FrameDriver: images are 8-bit grayscale");
  images_.push_front(Image{stamp_ns, next_image_++, image.clone()});
  // positions only grow, so the image that just moved past the window is never read again
  const size_t window = (size_t)cfg_.agg_num * (size_t)cfg_.interval;
  if (images_.size() > window + 1) std::next(images_.begin(), (long)window + 1)->image = Mat();
}

int FrameDriver::images_with_pixels() const {
  int n = 0;
  for (const Image& im : images_) n += im.image.empty() ? 0 : 1;
  return n;
}

void FrameDriver::grab_pose(int64_t stamp_ns, double tx, double ty, double tz, double qx,
                            double qy, double qz, double qw) {
  poses_.push_front(Pose{stamp_ns, next_pose_++, tx, ty, tz, qx, qy, qz, qw});
}

bool FrameDriver::process(DriverJob& job, Mat* depth) {
  // evaluate.cc:157-172: enough images buffered, and the newest pose not older than the
  // newest image
  if ((int)images_.size() <= last_size_ || poses_.empty()) return false;
  auto img = images_.begin();
  auto pose = poses_.begin();
  if (pose->t < img->t) return false;
  // :180-201: the pose with exactly the image's timestamp
  int pose_pos = 0;
  while (pose != poses_.end() && pose->t != img->t) {
    ++pose;
    ++pose_pos;
  }
  if (pose == poses_.end()) return false;   // "can not find pose for the current Img"
  // the sources' poses lie agg_num * interval places further down the pose list; the
  // reference reads past its end when they do not exist, this driver waits instead
  if (pose_pos + cfg_.agg_num * cfg_.interval >= (int)poses_.size()) {
    ++skipped_;
    return false;
  }
  job = DriverJob{};
  job.stamp_ns = img->t;
  job.ref_image = img->index;
  job.ref_pose = pose->index;
  const Mat ref = img->image;   // shares the pixels; the lists may drop the entry below
  const int ref_index = img->index;
  const Matrix3d R_ref = quaternion_to_rotation(pose->qx, pose->qy, pose->qz, pose->qw);
  const Vector3d t_ref(pose->tx, pose->ty, pose->tz);
  std::vector<Mat> src;
  std::vector<int> src_index;
  std::vector<Matrix3d> R_src;
  std::vector<Vector3d> t_src;
  // :209-233: every interval-th older image, and the pose as many places further on
  auto si = img;
  auto sp = pose;
  for (int i = 0; i < cfg_.agg_num; ++i) {
    si = std::next(si, cfg_.interval);
    sp = std::next(sp, cfg_.interval);
    job.src_images.push_back(si->index);
    job.src_poses.push_back(sp->index);
    src.push_back(si->image);
    src_index.push_back(si->index);
    R_src.push_back(quaternion_to_rotation(sp->qx, sp->qy, sp->qz, sp->qw));
    t_src.push_back(Vector3d(sp->tx, sp->ty, sp->tz));
  }
  // :234-237
  ++last_size_;
  images_.pop_back();
  poses_.pop_back();
  if (!with_mapper_) return true;

  // :243-258: the mapper (created at the first step, with the reference image's size, and
  // initialised as main() does, :337-348)
  if (!mapper_) {
    mapper_.reset(new StereoMapper(ref.cols, ref.rows, cfg_.planes));
    Mat K(3, 3, CV_64F);
    K.setTo(0.0);
    K.at<double>(0, 0) = cfg_.fx;
    K.at<double>(1, 1) = cfg_.fy;
    K.at<double>(0, 2) = cfg_.cx;
    K.at<double>(1, 2) = cfg_.cy;
    K.at<double>(2, 2) = 1.0;
    mapper_->initIntrinsic(K, (float)(0.02 * cfg_.fx), cfg_.dep_sample);
  }
  if (ref.rows != mapper_->height() || ref.cols != mapper_->width())
    throw Error(PM_ERR_INVALID_ARGUMENT, "FrameDriver: every image must have the first one's size");
  const Mat R_l = mat33(R_ref), T_l = mat31(t_ref(0), t_ref(1), t_ref(2));
  if (cfg_.frame_ring) {
    // the reference and its sources are among the newest agg_num * interval + 1 images
    // grabbed, so a ring of that many slots never evicts a frame a step still needs; a frame
    // is uploaded at its first use
    if (!ring_)
      ring_.reset(new FrameRing(ref.cols, ref.rows, cfg_.agg_num * cfg_.interval + 1));
    if (!ring_->holds(ref_index)) ring_->put(ref_index, ref, nullptr);
    for (int i = 0; i < cfg_.agg_num; ++i)
      if (!ring_->holds(src_index[i])) ring_->put(src_index[i], src[i], nullptr);
    mapper_->initReference(*ring_, ref_index);
    for (int i = 0; i < cfg_.agg_num; ++i)
      mapper_->update(*ring_, src_index[i], R_l, T_l, mat33(R_src[i]),
                      mat31(t_src[i](0), t_src[i](1), t_src[i](2)));
  } else {
    mapper_->initReference(ref);
    for (int i = 0; i < cfg_.agg_num; ++i)
      mapper_->update(src[i], R_l, T_l, mat33(R_src[i]),
                      mat31(t_src[i](0), t_src[i](1), t_src[i](2)));
  }
  Mat result;
  mapper_->outputFusionCPU(result, R_ref, t_ref, (float)cfg_.fx, (float)cfg_.fy, (float)cfg_.cx,
                           (float)cfg_.cy);
  if (depth) *depth = result;
  return true;
}

}  // namespace pm

// ---- C wrappers ---------------------------------------------------------------------------

struct pm_driver {
  pm::FrameDriver* impl;
  int agg_num;
};

namespace {

template <typename F>
pm_status driver_guard(F&& f) {
  pm::clear_error();
  try {
    f();
    return PM_OK;
  } catch (const pm::Error& e) {
    return pm::fail(e.status, e.what());
  } catch (const std::bad_alloc&) {
    return pm::fail(PM_ERR_OUT_OF_MEMORY, "host allocation failed");
  } catch (const std::exception& e) {
    return pm::fail(PM_ERR_DEVICE, e.what());
  }
}

}  // namespace

extern "C" pm_status pm_driver_create(const pm_driver_config* cfg, int32_t with_mapper,
                                      pm_driver** out) {
  if (!out || !cfg) return pm::fail(PM_ERR_INVALID_ARGUMENT, "pm_driver_create: null argument");
  *out = nullptr;
  return driver_guard([&] {
    pm::DriverConfig c;
    c.agg_num = cfg->agg_num;
    c.interval = cfg->interval;
    c.planes = cfg->planes;
    c.fx = cfg->fx; c.fy = cfg->fy; c.cx = cfg->cx; c.cy = cfg->cy;
    c.dep_sample = cfg->dep_sample;
    c.frame_ring = cfg->frame_ring != 0;
    auto* d = new pm_driver{nullptr, c.agg_num};
    try {
      d->impl = new pm::FrameDriver(c, with_mapper != 0);
    } catch (...) {
      delete d;
      throw;
    }
    *out = d;
  });
}

extern "C" void pm_driver_destroy(pm_driver* d) {
  if (!d) return;
  delete d->impl;
  delete d;
}

extern "C" pm_status pm_driver_grab_image(pm_driver* d, int64_t stamp_ns, const uint8_t* image,
                                          int32_t width, int32_t height) {
  if (!d || !d->impl || !image || width <= 0 || height <= 0)
    return pm::fail(PM_ERR_INVALID_ARGUMENT, "pm_driver_grab_image: bad argument");
  return driver_guard([&] {
    d->impl->grab_image(stamp_ns, pm::Mat(height, width, pm::CV_8U, const_cast<uint8_t*>(image)));
  });
}

extern "C" pm_status pm_driver_grab_pose(pm_driver* d, int64_t stamp_ns, const double pose[7]) {
  if (!d || !d->impl || !pose)
    return pm::fail(PM_ERR_INVALID_ARGUMENT, "pm_driver_grab_pose: bad argument");
  return driver_guard([&] {
    d->impl->grab_pose(stamp_ns, pose[0], pose[1], pose[2], pose[3], pose[4], pose[5], pose[6]);
  });
}

extern "C" pm_status pm_driver_process(pm_driver* d, int32_t* ran, int64_t* stamp_ns,
                                       int32_t* ref_image, int32_t* src_images, float* depth,
                                       size_t depth_bytes) {
  if (!d || !d->impl || !ran)
    return pm::fail(PM_ERR_INVALID_ARGUMENT, "pm_driver_process: bad argument");
  *ran = 0;
  return driver_guard([&] {
    pm::DriverJob job;
    pm::Mat result;
    if (!d->impl->process(job, depth ? &result : nullptr)) return;
    *ran = 1;
    if (stamp_ns) *stamp_ns = job.stamp_ns;
    if (ref_image) *ref_image = job.ref_image;
    if (src_images)
      for (int i = 0; i < d->agg_num; ++i) src_images[i] = job.src_images[i];
    if (depth && !result.empty()) {
      const size_t need = result.total() * sizeof(float);
      if (depth_bytes < need)
        throw pm::Error(PM_ERR_INVALID_ARGUMENT, "pm_driver_process: depth buffer too small");
      std::memcpy(depth, result.data, need);
    }
  });
}

extern "C" pm_status pm_driver_skipped(const pm_driver* d, int32_t* skipped) {
  if (!d || !d->impl || !skipped)
    return pm::fail(PM_ERR_INVALID_ARGUMENT, "pm_driver_skipped: bad argument");
  *skipped = d->impl->skipped();
  return PM_OK;
}

extern "C" pm_status pm_driver_buffered(const pm_driver* d, int32_t* images, int32_t* poses,
                                        int32_t* images_with_pixels) {
  if (!d || !d->impl)
    return pm::fail(PM_ERR_INVALID_ARGUMENT, "pm_driver_buffered: bad argument");
  if (images) *images = d->impl->buffered_images();
  if (poses) *poses = d->impl->buffered_poses();
  if (images_with_pixels) *images_with_pixels = d->impl->images_with_pixels();
  return PM_OK;
}

extern "C" pm_status pm_driver_ring_puts(const pm_driver* d, int64_t* puts) {
  if (!d || !d->impl || !puts)
    return pm::fail(PM_ERR_INVALID_ARGUMENT, "pm_driver_ring_puts: bad argument");
  const pm::FrameRing* r = d->impl->ring();
  *puts = r ? r->puts() : 0;
  return PM_OK;
}

extern "C" pm_status pm_quaternion_to_rotation(const double q[4], double R[9]) {
  if (!q || !R) return pm::fail(PM_ERR_INVALID_ARGUMENT, "pm_quaternion_to_rotation: null");
  const pm::Matrix3d m = pm::quaternion_to_rotation(q[0], q[1], q[2], q[3]);
  for (int k = 0; k < 9; ++k) R[k] = m.m[k];
  return PM_OK;
}

// The keyframe / source selection driver of the reference's ROS node, without ROS.
//
// Reference: evaluate.cc:29-56 (aggNum, interval, intrinsics), :75-110 (GrabImage / GrabPose
// push the newest message to the front of a list), :137-276 (processImg), :311-360 (main:
// K, bf, dep_sample, lastSize = aggNum * interval).  One processImg step:
//   - runs only once the image list is longer than lastSize, and only if the newest pose is
//     not older than the newest image;
//   - takes the newest image as the reference and the pose with exactly its timestamp
//     (quaternion -> rotation matrix, Eigen's formula);
//   - takes as measurement frames the images aggNum times `interval` further down the image
//     list, with the poses the same number of places further down the pose list from the
//     reference's pose (by position, not by timestamp, as the reference does);
//   - then lastSize grows by one and the oldest image and pose are dropped;
//   - and runs StereoMapper::initReference, update per source and outputFusionCPU, whose
//     result is the node's 32FC1 depth message.
// Differences: a step whose source pose would lie past the end of the pose list (the
// reference dereferences past the end there) is not run and counts as skipped; timestamps
// are integer nanoseconds (ros::Time compares seconds and nanoseconds exactly).  The lists
// grow by about one message per step, as the reference's do, but an image more than
// agg_num * interval places from the front can never be a reference or a source again (its
// position only grows), so the driver frees its pixels and keeps only its list entry.
#ifndef PM_FRAME_DRIVER_H
#define PM_FRAME_DRIVER_H

#include <cstdint>
#include <list>
#include <memory>
#include <vector>

#include "pm/stereo_mapper.h"

namespace pm {

struct DriverConfig {
  int agg_num = 3;                        // evaluate.cc:24
  int interval = 6;                       // evaluate.cc:25
  // intrinsics (evaluate.cc:37-40, the ICL-NUIM set) and StereoMapper::initIntrinsic's
  // arguments (evaluate.cc:337-348)
  double fx = 481.2017, fy = -480.0002, cx = 319.5, cy = 239.5;
  float dep_sample = 1.0f / (0.15f * 160.0f);
  int planes = 64;                        // DEP_CNT
  // frames reach the mapper through an upload-once FrameRing (each used frame is uploaded and
  // converted once instead of at every use); the depths are identical either way
  bool frame_ring = true;
};

// One processImg step's choice: the reference frame and its sources, as indices into the
// order in which images (and poses) were grabbed.
struct DriverJob {
  int64_t stamp_ns = 0;                   // the reference image's timestamp
  int ref_image = -1, ref_pose = -1;
  std::vector<int> src_images, src_poses;
};

// Eigen::Quaterniond(w, x, y, z).toRotationMatrix() of a unit quaternion, row-major.
Matrix3d quaternion_to_rotation(double qx, double qy, double qz, double qw);

class FrameDriver {
 public:
  // with_mapper = false plans the steps only (no GPU work, no device memory)
  FrameDriver(const DriverConfig& cfg, bool with_mapper = true);
  ~FrameDriver();
  FrameDriver(const FrameDriver&) = delete;
  FrameDriver& operator=(const FrameDriver&) = delete;

  // GrabImage (8-bit grayscale, copied) and GrabPose (evaluate.cc:75-110)
  void grab_image(int64_t stamp_ns, const Mat& image);
  void grab_pose(int64_t stamp_ns, double tx, double ty, double tz, double qx, double qy,
                 double qz, double qw);

  // One processImg step.  Returns true if a step ran: job describes it and, with a mapper,
  // depth holds the 32FC1 result (CV_32F, rows x cols of the reference image).
  bool process(DriverJob& job, Mat* depth);

  int skipped() const { return skipped_; }   // steps refused for lack of source poses
  // list lengths, and how many buffered images still hold their pixels
  int buffered_images() const { return (int)images_.size(); }
  int buffered_poses() const { return (int)poses_.size(); }
  int images_with_pixels() const;
  StereoMapper* mapper() { return mapper_.get(); }
  const FrameRing* ring() const { return ring_.get(); }

 private:
  struct Image {
    int64_t t;
    int index;
    Mat image;
  };
  struct Pose {
    int64_t t;
    int index;
    double tx, ty, tz, qx, qy, qz, qw;
  };
  DriverConfig cfg_;
  bool with_mapper_;
  std::unique_ptr<StereoMapper> mapper_;
  std::unique_ptr<FrameRing> ring_;
  std::list<Image> images_;   // newest first
  std::list<Pose> poses_;     // newest first
  int last_size_;
  int next_image_ = 0, next_pose_ = 0;
  int skipped_ = 0;
};

}  // namespace pm

#endif  // PM_FRAME_DRIVER_H

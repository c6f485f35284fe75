// StereoMapper: the per-keyframe depth pipeline behind the reference's interface
// (motion_stereo/stereo_mapper.h:26-63).  Same class name, method set, argument order and
// semantics; cv::Mat / cv::cuda::GpuMat / Eigen arguments become the look-alikes of
// pm/types.h.  Differences a caller can observe:
//   * shapes are runtime (the default constructor uses WIDTH x HEIGHT x DEP_CNT);
//   * update() stages the measurement frame; the cost of all staged frames is computed in one
//     pass when an output method runs (the running mean of the reference is order-free);
//   * initReference() and update() copy the host image into pinned staging and return
//     without synchronising (the caller may reuse its buffer at once); output methods
//     synchronise before they hand back a host result, which, when the mapper creates it, lives
//     in page-locked memory;
//   * raw_cost may be stored as fp16 / bf16 (set_cost_type);
//   * no GUI calls: outputFusion's imshow (stereo_mapper.cpp:232-235) is dropped;
//   * errors raise pm::Error (a std::runtime_error with a pm_status) instead of exiting.
#ifndef PM_STEREO_MAPPER_H
#define PM_STEREO_MAPPER_H

#include <stdexcept>
#include <string>
#include <vector>

#include "pm/frame_ring.h"
#include "pm/parameters.h"
#include "pm/types.h"
#include "pm_capi.h"

namespace pm {

class Error : public std::runtime_error {
 public:
  Error(pm_status s, const std::string& msg) : std::runtime_error(msg), status(s) {}
  pm_status status;
};

class StereoMapper {
 public:
  StereoMapper();
  StereoMapper(int width, int height, int dep_cnt);
  ~StereoMapper();
  StereoMapper(const StereoMapper&) = delete;
  StereoMapper& operator=(const StereoMapper&) = delete;

  void initIntrinsic(const Mat& _K1, float bf, float dep_sample);
  void initReference(const Mat& _img_l);
  void update(const Mat& _img_r, const Mat& R_l, const Mat& T_l, const Mat& R_r,
              const Mat& T_r);
  void output(Mat& result);
  void outputFusion(Mat& result, Matrix3d R, Vector3d t, float fx, float fy, float cx,
                    float cy);
  // CASE 0 / 1 / 2 ablation outputs (stereo_mapper.cpp:625-719); CASE >= 3 runs the
  // reference's CPU fusion rules (:719-800) on the GPU, with their own state.
  void outputFusionCPU(Mat& result, Matrix3d R, Vector3d t, float fx, float fy, float cx,
                       float cy);

  // extensions
  // initReference / update reading an ingested frame of a FrameRing instead of a host image
  // (no upload, no conversion; the homography is set up as in update).  The frame must stay
  // in the ring until the next output call returns.
  void initReference(const FrameRing& ring, int64_t frame);
  void update(const FrameRing& ring, int64_t frame, const Mat& R_l, const Mat& T_l,
              const Mat& R_r, const Mat& T_r);
  void set_cost_type(pm_cost_type t);
  pm_cost_type cost_type() const { return cost_type_; }
  int width() const { return W_; }
  int height() const { return H_; }
  int planes() const { return D_; }
  int num_sources() const { return (int)homs_.size(); }
  void synchronize() const;
  // runs the pending cost, SGM and filter stages without downloading anything
  void compute_depth();
  // CASE 0 / 1 winner-take-all maps (host result [H][W])
  void wta(int case_mode, Mat& result);

  // members of the reference (stereo_mapper.h:44-62); device buffers
  DeviceMat img_l;                         // uint8 reference image
  DeviceMat raw_cost, sgm_cost;            // [H][W][D]
  DeviceMat dep;                           // filtered depth [H][W]
  DeviceMat propogate_table;               // int32 [H][W], filled table of the last fusion
  DeviceMat alpha0, beta0, mu0, sigma0;
  DeviceMat alpha1, beta1, mu1, sigma1;
  DeviceMat fuseDepth;                     // fused depth [H][W]
  // state and table of outputFusionCPU's CASE >= 3 step (the reference's CPU rules keep their
  // own filter state, stereo_mapper.cpp:6-16); allocated at its first use
  DeviceMat cpu_alpha[2], cpu_beta[2], cpu_mu[2], cpu_sigma[2];
  DeviceMat cpu_table;                     // int32 [H][W], (y << 16) | x

  int pre_index, cur_index;
  bool firstDepth;
  Mat nK1;
  Mat R, T;                                // last homography setup (H1, H2)
  Matrix3d R_b_2_w;
  Vector3d t_b_2_w;

  float DEP_SAMPLE;

 private:
  void alloc();
  void release();
  void ensure_cost();
  void run_sgm_filter(bool median = true);
  pm_homography stage_homography(const Mat& R_l, const Mat& T_l, const Mat& R_r,
                                 const Mat& T_r);
  void fuse_cpu_rules(Mat& result, Matrix3d R, Vector3d t, float fx, float fy, float cx,
                      float cy);
  void stage_upload(const Mat& img, uint8_t* dst, size_t slot, const char* who);
  void download_result(const DeviceMat& src, Mat& result);

  int W_, H_, D_;
  pm_cost_type cost_type_ = PM_COST_F32;
  pm_cost_type raw_cost_type_ = PM_COST_F32;
  void* stream_ = nullptr;
  std::vector<DeviceMat> quads_;           // per staged source, 8-byte texels [H][W]
  struct RingFrame {
    const FrameRing* ring;                 // null: the source's texture is quads_[slot]
    int64_t frame;
  };
  std::vector<RingFrame> staged_;          // per staged source
  std::vector<pm_homography> homs_;
  DeviceMat zbuf_, med_tmp_, wta_tmp_;   // zbuf_: two [H][W] z-buffers, zbuf_ half zpar_ empty
  int zpar_ = 0;
  DeviceMat sgm_ws_;                       // per-path L volumes (pm_sgm_filter_paths)
  DeviceMat cpu_work_, cpu_result_;        // outputFusionCPU's CASE >= 3 step
  bool cost_ready_ = false;
  bool depth_ready_ = false;
  bool raw_ready_ = false;     // SGM + filter done (into med_tmp_ if raw_in_med_, else dep)
  bool raw_in_med_ = false;    // the median knob when the filter ran: dep = median(med_tmp_)
  bool have_reference_ = false;
  // pinned host staging of host images (slot 0: the reference, slot 1 + s: source s), each
  // with the event recorded after its last copy; initReference and update copy the caller's
  // image into a slot and return without waiting for the device
  std::vector<uint8_t*> pinned_;
  std::vector<void*> pinned_ev_;
  Mat out_staging_;                        // page-locked staging of results for callers' Mats
};

}  // namespace pm

#endif  // PM_STEREO_MAPPER_H

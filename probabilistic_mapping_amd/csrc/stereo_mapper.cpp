// StereoMapper host shell: owns the device buffers and sequences the stage kernels.
// Reference: motion_stereo/stereo_mapper.cpp:27-244 (constructor, initIntrinsic,
// initReference, update, output, outputFusion) and :625-719 (outputFusionCPU CASE 0-2).
#include "pm/stereo_mapper.h"

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <string>

#include "common.hpp"

namespace pm {

namespace {

void check(pm_status s) {
  if (s != PM_OK) throw Error(s, pm_last_error());
}

void hip_check(hipError_t e, const char* what) {
  if (e != hipSuccess)
    throw Error(e == hipErrorOutOfMemory ? PM_ERR_OUT_OF_MEMORY : PM_ERR_DEVICE,
                std::string(what) + ": " + hipGetErrorString(e));
}

DeviceMat dalloc(int rows, int cols, int type) {
  DeviceMat m;
  m.rows = rows; m.cols = cols; m.type = type; m.step = (size_t)cols * elem_size(type);
  void* p = nullptr;
  hip_check(hipMalloc(&p, m.bytes()), "hipMalloc");
  m.data = static_cast<uint8_t*>(p);
  return m;
}

void dfree(DeviceMat& m) {
  if (m.data) (void)hipFree(m.data);
  m = DeviceMat();
}

int storage_type(pm_cost_type t) {
  return t == PM_COST_F16 ? PM_16F : (t == PM_COST_BF16 ? PM_16BF : CV_32F);
}

// The checks the reference's cv::Mat inputs get implicitly: size and 8-bit grayscale type.
void check_image(const Mat& img, int rows, int cols, const char* who) {
  if (img.empty() || img.rows != rows || img.cols != cols)
    throw Error(PM_ERR_INVALID_ARGUMENT, std::string(who) + ": image size must be " +
                                             std::to_string(cols) + "x" +
                                             std::to_string(rows));
  if (img.type != CV_8U)
    throw Error(PM_ERR_INVALID_ARGUMENT, std::string(who) + ": image must be 8-bit grayscale");
}

// Page-locked host storage for result matrices: device-to-host copies into it are plain DMA.
uint8_t* pinned_alloc(size_t n) {
  void* p = nullptr;
  hip_check(hipHostMalloc(&p, n ? n : 1, hipHostMallocDefault), "hipHostMalloc");
  return static_cast<uint8_t*>(p);
}
void pinned_release(uint8_t* p) {
  if (p) (void)hipHostFree(p);
}
const char kPinnedTag = 0;   // &kPinnedTag marks a Mat whose storage is pinned_alloc's

void mat_to_array(const Mat& m, int rows, int cols, double* out, const char* what) {
  if (m.rows != rows || m.cols != cols || m.type != CV_64F)
    throw Error(PM_ERR_INVALID_ARGUMENT, std::string(what) + " must be a " +
                                             std::to_string(rows) + "x" +
                                             std::to_string(cols) + " CV_64F matrix");
  for (int i = 0; i < rows; ++i)
    for (int j = 0; j < cols; ++j) out[i * cols + j] = m.at<double>(i, j);
}

}  // namespace

void DeviceMat::download(Mat& m) const {
  if (m.rows != rows || m.cols != cols || m.type != type) m.create(rows, cols, type);
  hip_check(hipMemcpy(m.data, data, bytes(), hipMemcpyDeviceToHost), "download");
}

void DeviceMat::setTo(float v) const {
  const size_t n = (size_t)rows * cols;
  if (type == CV_32F) {
    check(pm_fill_f32(reinterpret_cast<float*>(data), n, v, nullptr));
    hip_check(hipDeviceSynchronize(), "setTo");
  } else if (v == 0.0f) {
    hip_check(hipMemset(data, 0, bytes()), "setTo");
  } else {
    throw Error(PM_ERR_INVALID_ARGUMENT, "DeviceMat::setTo: non-zero fill of a non-float buffer");
  }
}

StereoMapper::StereoMapper() : StereoMapper(WIDTH, HEIGHT, DEP_CNT) {}

StereoMapper::StereoMapper(int width, int height, int dep_cnt)
    : pre_index(0), cur_index(1), firstDepth(true), DEP_SAMPLE(0.0f), W_(width), H_(height),
      D_(dep_cnt) {
  if (W_ <= 0 || H_ <= 0 || W_ > 65535 || H_ > 65535)
    throw Error(PM_ERR_INVALID_ARGUMENT, "StereoMapper: bad image size");
  if (D_ < 1 || D_ > PM_MAX_PLANES)
    throw Error(PM_ERR_INVALID_ARGUMENT, "StereoMapper: dep_cnt must be in 1..256");
  // the stages address a volume through 32-bit buffer offsets (pm_sgm and friends)
  if ((uint64_t)W_ * (uint64_t)H_ * (uint64_t)D_ * 4u >= (1ull << 32))
    throw Error(PM_ERR_INVALID_ARGUMENT,
                "StereoMapper: width * height * dep_cnt * 4 bytes must stay below 4 GiB");
  hipStream_t st = nullptr;
  hip_check(hipStreamCreate(&st), "hipStreamCreate");
  stream_ = st;
  try {
    alloc();
  } catch (...) {
    release();
    (void)hipStreamDestroy(st);
    throw;
  }
}

StereoMapper::~StereoMapper() {
  if (stream_) (void)hipStreamSynchronize((hipStream_t)stream_);
  release();
  for (uint8_t* p : pinned_)
    if (p) (void)hipHostFree(p);
  for (void* e : pinned_ev_)
    if (e) (void)hipEventDestroy((hipEvent_t)e);
  if (stream_) (void)hipStreamDestroy((hipStream_t)stream_);
}

// Copies a host image into pinned staging slot `slot` and enqueues its upload to `dst` (W x H
// bytes) on the mapper's stream.  The slot's previous upload is waited for first, so the
// caller's buffer is free when this returns and no stream synchronisation is needed.
void StereoMapper::stage_upload(const Mat& img, uint8_t* dst, size_t slot, const char* who) {
  check_image(img, H_, W_, who);
  const size_t bytes = (size_t)W_ * H_;
  if (pinned_.size() <= slot) {
    pinned_.resize(slot + 1, nullptr);
    pinned_ev_.resize(slot + 1, nullptr);
  }
  if (!pinned_[slot]) {
    void* p = nullptr;
    hip_check(hipHostMalloc(&p, bytes, hipHostMallocDefault), "hipHostMalloc");
    pinned_[slot] = static_cast<uint8_t*>(p);
    hipEvent_t e = nullptr;
    hip_check(hipEventCreateWithFlags(&e, hipEventDisableTiming), "hipEventCreate");
    pinned_ev_[slot] = e;
  } else {
    hip_check(hipEventSynchronize((hipEvent_t)pinned_ev_[slot]), "hipEventSynchronize");
  }
  uint8_t* p = pinned_[slot];
  if (img.step == (size_t)W_) {
    std::memcpy(p, img.data, bytes);
  } else {
    for (int y = 0; y < H_; ++y) std::memcpy(p + (size_t)y * W_, img.data + y * img.step, W_);
  }
  hipStream_t st = (hipStream_t)stream_;
  hip_check(hipMemcpyAsync(dst, p, bytes, hipMemcpyHostToDevice, st), "upload");
  hip_check(hipEventRecord((hipEvent_t)pinned_ev_[slot], st), "hipEventRecord");
}

void StereoMapper::alloc() {
  const long N = (long)W_ * H_;
  img_l = dalloc(H_, W_, CV_8U);
  raw_cost_type_ = cost_type_;
  raw_cost = dalloc(1, (int)(N * D_), storage_type(cost_type_));
  sgm_cost = dalloc(1, (int)(N * D_), CV_32F);
  dep = dalloc(H_, W_, CV_32F);
  fuseDepth = dalloc(H_, W_, CV_32F);
  propogate_table = dalloc(H_, W_, CV_32S);
  DeviceMat* st[8] = {&alpha0, &beta0, &mu0, &sigma0, &alpha1, &beta1, &mu1, &sigma1};
  for (DeviceMat* m : st) *m = dalloc(1, (int)N, CV_32F);
  zbuf_ = dalloc(2, (int)N, PM_64U);   // two z-buffers, used alternately by the fusion steps
  med_tmp_ = dalloc(H_, W_, CV_32F);
  wta_tmp_ = dalloc(H_, W_, CV_32F);
  hip_check(hipMemset(propogate_table.data, 0, propogate_table.bytes()), "hipMemset");
  hip_check(hipMemset(zbuf_.data, 0xFF, zbuf_.bytes()), "hipMemset");   // empty entries
  zpar_ = 0;
  for (DeviceMat* m : st) hip_check(hipMemset(m->data, 0, m->bytes()), "hipMemset");
}

void StereoMapper::release() {
  DeviceMat* all[] = {&img_l, &raw_cost, &sgm_cost, &dep, &fuseDepth, &propogate_table,
                      &alpha0, &beta0, &mu0, &sigma0, &alpha1, &beta1, &mu1, &sigma1,
                      &zbuf_, &med_tmp_, &wta_tmp_, &sgm_ws_, &cpu_table, &cpu_work_,
                      &cpu_result_};
  for (DeviceMat* m : all) dfree(*m);
  for (int h = 0; h < 2; ++h) {
    dfree(cpu_alpha[h]); dfree(cpu_beta[h]); dfree(cpu_mu[h]); dfree(cpu_sigma[h]);
  }
  for (DeviceMat& q : quads_) dfree(q);
  quads_.clear();
}

void StereoMapper::set_cost_type(pm_cost_type t) {
  if (t != PM_COST_F32 && t != PM_COST_F16 && t != PM_COST_BF16)
    throw Error(PM_ERR_INVALID_ARGUMENT, "set_cost_type: unknown type");
  cost_type_ = t;
}

// A device result handed back to the caller, as cv::cuda::GpuMat::download does (result is
// (re)created only when its shape or type differs).  A result the mapper created lives in
// page-locked memory and receives the copy directly; a caller's matrix of the right shape
// receives it through a page-locked staging buffer.
void StereoMapper::download_result(const DeviceMat& src, Mat& result) {
  hipStream_t st = (hipStream_t)stream_;
  const bool fits = result.data && result.rows == src.rows && result.cols == src.cols &&
                    result.type == src.type;
  if (!fits)
    result.create_with(src.rows, src.cols, src.type, pinned_alloc, pinned_release, &kPinnedTag);
  if (result.alloc_tag() == &kPinnedTag && result.step == src.step) {
    hip_check(hipMemcpyAsync(result.data, src.data, src.bytes(), hipMemcpyDeviceToHost, st),
              "download");
    synchronize();
    return;
  }
  if (out_staging_.rows != src.rows || out_staging_.cols != src.cols ||
      out_staging_.type != src.type)
    out_staging_.create_with(src.rows, src.cols, src.type, pinned_alloc, pinned_release,
                             &kPinnedTag);
  hip_check(hipMemcpyAsync(out_staging_.data, src.data, src.bytes(), hipMemcpyDeviceToHost, st),
            "download");
  synchronize();
  for (int y = 0; y < src.rows; ++y)
    std::memcpy(result.data + y * result.step, out_staging_.data + y * src.step, src.step);
}

void StereoMapper::synchronize() const {
  hip_check(hipStreamSynchronize((hipStream_t)stream_), "hipStreamSynchronize");
}

// stereo_mapper.cpp:51-61 (the depth listing it prints is not reproduced)
void StereoMapper::initIntrinsic(const Mat& _K1, float bf, float dep_sample) {
  (void)bf;   // unused by the reference as well
  double k[9];
  mat_to_array(_K1, 3, 3, k, "initIntrinsic: K");
  nK1 = _K1.clone();
  DEP_SAMPLE = dep_sample;
}

// stereo_mapper.cpp:63-71: new reference frame; the cost volume restarts at -1
void StereoMapper::initReference(const Mat& _img_l) {
  stage_upload(_img_l, img_l.data, 0, "initReference");
  if (raw_cost_type_ != cost_type_) {
    synchronize();
    dfree(raw_cost);
    raw_cost = dalloc(1, (int)((long)W_ * H_ * D_), storage_type(cost_type_));
    raw_cost_type_ = cost_type_;
  }
  homs_.clear();
  staged_.clear();
  have_reference_ = true;
  cost_ready_ = false;
  depth_ready_ = raw_ready_ = false;
}

// initReference from an ingested frame: a device copy into img_l, ordered after the frame's
// upload by its event
void StereoMapper::initReference(const FrameRing& ring, int64_t frame) {
  if (ring.width() != W_ || ring.height() != H_)
    throw Error(PM_ERR_INVALID_ARGUMENT, "initReference: the ring's frames must be " +
                                             std::to_string(W_) + "x" + std::to_string(H_));
  hipStream_t st = (hipStream_t)stream_;
  hip_check(hipStreamWaitEvent(st, (hipEvent_t)ring.ready_event(frame), 0), "hipStreamWaitEvent");
  hip_check(hipMemcpyAsync(img_l.data, ring.image(frame), img_l.bytes(), hipMemcpyDeviceToDevice,
                           st),
            "copy reference");
  if (raw_cost_type_ != cost_type_) {
    synchronize();
    dfree(raw_cost);
    raw_cost = dalloc(1, (int)((long)W_ * H_ * D_), storage_type(cost_type_));
    raw_cost_type_ = cost_type_;
  }
  homs_.clear();
  staged_.clear();
  have_reference_ = true;
  cost_ready_ = false;
  depth_ready_ = raw_ready_ = false;
}

// the homography of update (stereo_mapper.cpp:79-90), also kept in R, T as the reference does
pm_homography StereoMapper::stage_homography(const Mat& R_l, const Mat& T_l, const Mat& R_r,
                                             const Mat& T_r) {
  if (!have_reference_)
    throw Error(PM_ERR_INVALID_STATE, "update: initReference has not been called");
  if (nK1.empty()) throw Error(PM_ERR_INVALID_STATE, "update: initIntrinsic has not been called");
  if ((int)homs_.size() >= PM_MAX_SOURCES)
    throw Error(PM_ERR_INVALID_STATE, "update: at most 16 measurement frames per keyframe");
  double K[9], rl[9], tl[3], rr[9], tr[3], Hm[9], t[3];
  mat_to_array(nK1, 3, 3, K, "K");
  mat_to_array(R_l, 3, 3, rl, "update: R_l");
  mat_to_array(T_l, 3, 1, tl, "update: T_l");
  mat_to_array(R_r, 3, 3, rr, "update: R_r");
  mat_to_array(T_r, 3, 1, tr, "update: T_r");
  if (!homography_fp64(K, rl, tl, rr, tr, Hm, t))
    throw Error(PM_ERR_INVALID_ARGUMENT, "update: non-finite pose or singular K");
  pm_homography h;
  for (int k = 0; k < 9; ++k) h.h[k] = (float)Hm[k];
  for (int k = 0; k < 3; ++k) h.t[k] = (float)t[k];
  R.create(3, 3, CV_64F);
  T.create(3, 1, CV_64F);
  for (int k = 0; k < 9; ++k) R.at<double>(k / 3, k % 3) = Hm[k];
  for (int k = 0; k < 3; ++k) T.at<double>(k, 0) = t[k];
  return h;
}

// stereo_mapper.cpp:73-91: stage a measurement frame (image + homography)
void StereoMapper::update(const Mat& _img_r, const Mat& R_l, const Mat& T_l, const Mat& R_r,
                          const Mat& T_r) {
  const pm_homography h = stage_homography(R_l, T_l, R_r, T_r);
  const size_t slot = homs_.size();
  if (quads_.size() <= slot) quads_.resize(slot + 1);   // ring-staged slots stay unallocated
  if (!quads_[slot].data) quads_[slot] = dalloc(H_, 2 * W_, CV_32S);   // 8-byte texels
  // the u8 frame goes through the WTA scratch's device memory, which is free until an
  // output call; uploads and conversions of successive updates are ordered on the stream
  stage_upload(_img_r, wta_tmp_.data, 1 + slot, "update");
  check(pm_make_quad(wta_tmp_.data, reinterpret_cast<uint64_t*>(quads_[slot].data), W_, H_,
                     stream_));
  homs_.push_back(h);
  staged_.push_back(RingFrame{nullptr, -1});
  cost_ready_ = false;
  depth_ready_ = raw_ready_ = false;
}

// update from an ingested frame: only the homography is set up here; the cost pass reads
// the frame's texture from the ring
void StereoMapper::update(const FrameRing& ring, int64_t frame, const Mat& R_l, const Mat& T_l,
                          const Mat& R_r, const Mat& T_r) {
  if (ring.width() != W_ || ring.height() != H_)
    throw Error(PM_ERR_INVALID_ARGUMENT, "update: the ring's frames must be " +
                                             std::to_string(W_) + "x" + std::to_string(H_));
  if (!ring.holds(frame))
    throw Error(PM_ERR_INVALID_STATE, "update: frame " + std::to_string(frame) +
                                          " is not in the ring");
  const pm_homography h = stage_homography(R_l, T_l, R_r, T_r);
  homs_.push_back(h);
  staged_.push_back(RingFrame{&ring, frame});
  cost_ready_ = false;
  depth_ready_ = raw_ready_ = false;
}

void StereoMapper::ensure_cost() {
  if (cost_ready_) return;
  if (!have_reference_)
    throw Error(PM_ERR_INVALID_STATE, "output: initReference has not been called");
  std::vector<const uint64_t*> q(homs_.size());
  hipStream_t st = (hipStream_t)stream_;
  for (size_t s = 0; s < homs_.size(); ++s) {
    const RingFrame& f = staged_[s];
    if (!f.ring) {
      q[s] = reinterpret_cast<const uint64_t*>(quads_[s].data);
      continue;
    }
    if (!f.ring->holds(f.frame))
      throw Error(PM_ERR_INVALID_STATE, "output: staged frame " + std::to_string(f.frame) +
                                            " was evicted from its ring before the output call");
    hip_check(hipStreamWaitEvent(st, (hipEvent_t)f.ring->ready_event(f.frame), 0),
              "hipStreamWaitEvent");
    q[s] = f.ring->quad(f.frame);
  }
  check(pm_calc_cost(pm_shape{W_, H_, D_}, img_l.data, (int32_t)homs_.size(), q.data(),
                     homs_.data(), IDEP_MIN, DEP_SAMPLE, raw_cost.data, raw_cost_type_,
                     stream_));
  cost_ready_ = true;
}

// sgm (stereo_mapper.cpp:118-120) + filter (:125-126) (+ the optional median).  median =
// false leaves a pending median to the caller (outputFusion computes it inside its fusion
// step's launch): the filter's depth is then in med_tmp_ and depth_ready_ stays false.
void StereoMapper::run_sgm_filter(bool median) {
  if (depth_ready_) return;
  if (!raw_ready_) {
    ensure_cost();
    pm_sgm_params p{pi1, pi2, tau_so, sgm_q1, sgm_q2, SGM_PATHS};
    const pm_shape sh{W_, H_, D_};
    raw_in_med_ = MEDIAN_FILTER != 0;
    float* fout = reinterpret_cast<float*>(raw_in_med_ ? med_tmp_.data : dep.data);
    // all paths in flight at once through the per-path workspace (same bits as pm_sgm_filter)
    uint64_t need = 0;
    check(pm_sgm_workspace_size(sh, SGM_PATHS, &need));
    if (sgm_ws_.bytes() < need) {
      hip_check(hipStreamSynchronize((hipStream_t)stream_), "hipStreamSynchronize");
      dfree(sgm_ws_);
      sgm_ws_ = dalloc(SGM_PATHS, (int)((long)W_ * H_ * D_), CV_32F);
    }
    check(pm_sgm_filter_paths(sh, img_l.data, raw_cost.data, raw_cost_type_, p, var_scale,
                              IDEP_MIN, DEP_SAMPLE, sgm_ws_.data, sgm_ws_.bytes(),
                              reinterpret_cast<float*>(sgm_cost.data), fout, stream_));
    raw_ready_ = true;
  }
  if (raw_in_med_) {
    if (!median) return;
    check(pm_median3x3(reinterpret_cast<const float*>(med_tmp_.data),
                       reinterpret_cast<float*>(dep.data), W_, H_, stream_));
  }
  depth_ready_ = true;
}

void StereoMapper::compute_depth() { run_sgm_filter(); }

// stereo_mapper.cpp:93-109
void StereoMapper::output(Mat& result) {
  run_sgm_filter();
  download_result(dep, result);
}

// stereo_mapper.cpp:111-244
void StereoMapper::outputFusion(Mat& result, Matrix3d Rw, Vector3d tw, float fx, float fy,
                                float cx, float cy) {
  run_sgm_filter(/*median=*/false);   // a pending median runs inside the fusion step
  const size_t N = (size_t)W_ * H_;
  if (firstDepth) {   // :128-146
    run_sgm_filter();
    firstDepth = false;
    if (!(result.data && result.rows == H_ && result.cols == W_ && result.type == CV_32F))
      result.create_with(H_, W_, CV_32F, pinned_alloc, pinned_release, &kPinnedTag);
    result.setTo<float>(0.0f);
    R_b_2_w = Rw;
    t_b_2_w = tw;
    pre_index = 0;
    cur_index = 1;
    check(pm_fill_f32(reinterpret_cast<float*>(alpha0.data), N, 0.0f, stream_));
    check(pm_fill_f32(reinterpret_cast<float*>(alpha1.data), N, 0.0f, stream_));
    synchronize();
    return;
  }
  // :149-152 (dense layouts: the fusion step reads dep and writes fuseDepth, no conversion)
  // :155-156, FP64 then cast to float (:167-170)
  const Matrix3d R_p_2_c = Rw.transpose() * R_b_2_w;
  const Vector3d t_p_2_c = Rw.transpose() * (t_b_2_w - tw);
  float r[9], t[3];
  for (int k = 0; k < 9; ++k) r[k] = (float)R_p_2_c.m[k];
  for (int k = 0; k < 3; ++k) t[k] = (float)t_p_2_c.v[k];
  DeviceMat* s0[4] = {&alpha0, &beta0, &mu0, &sigma0};
  DeviceMat* s1[4] = {&alpha1, &beta1, &mu1, &sigma1};
  DeviceMat** pre = pre_index == 0 ? s0 : s1;
  DeviceMat** cur = pre_index == 0 ? s1 : s0;
  auto state = [](DeviceMat** s) {
    return pm_fusion_state{reinterpret_cast<float*>(s[0]->data), reinterpret_cast<float*>(s[1]->data),
                           reinterpret_cast<float*>(s[2]->data), reinterpret_cast<float*>(s[3]->data)};
  };
  // pm_depth_fuse as its two halves, with the pending median (if any) computed inside the
  // second launch: the forward warp fills z-buffer zpar_ and empties the other one for the
  // next keyframe, then hole fill, predict and update (same results as median3x3 +
  // pm_depth_fuse)
  uint64_t* zb = reinterpret_cast<uint64_t*>(zbuf_.data);
  const pm_camera cam{fx, fy, cx, cy};
  check(pm_depth_fuse_warp(W_, H_, cam, r, t, state(pre), zb + N * zpar_, zb + N * (zpar_ ^ 1),
                           stream_));
  const bool med = !depth_ready_;
  check(pm_depth_fuse_step(W_, H_, cam, reinterpret_cast<const float*>(med ? med_tmp_.data : dep.data),
                           med ? 1 : 0, med ? reinterpret_cast<float*>(dep.data) : nullptr,
                           reinterpret_cast<float*>(fuseDepth.data), zb + N * zpar_,
                           reinterpret_cast<int32_t*>(propogate_table.data), state(pre),
                           state(cur), nullptr, nullptr, nullptr, nullptr, stream_));
  depth_ready_ = true;
  zpar_ ^= 1;
  download_result(dep, result);   // :229-230 -- the filtered, not the fused, depth
  R_b_2_w = Rw;            // :237-241
  t_b_2_w = tw;
  pre_index = cur_index;
  cur_index ^= 1;
}

void StereoMapper::wta(int case_mode, Mat& result) {
  const pm_shape sh{W_, H_, D_};
  if (case_mode == 0) {
    ensure_cost();
    check(pm_wta(sh, raw_cost.data, raw_cost_type_, IDEP_MIN, DEP_SAMPLE,
                 reinterpret_cast<float*>(wta_tmp_.data), stream_));
  } else {
    ensure_cost();
    pm_sgm_params p{pi1, pi2, tau_so, sgm_q1, sgm_q2, SGM_PATHS};
    check(pm_sgm(sh, img_l.data, raw_cost.data, raw_cost_type_, p,
                 reinterpret_cast<float*>(sgm_cost.data), stream_));
    check(pm_wta(sh, sgm_cost.data, PM_COST_F32, IDEP_MIN, DEP_SAMPLE,
                 reinterpret_cast<float*>(wta_tmp_.data), stream_));
  }
  download_result(wta_tmp_, result);
}

// stereo_mapper.cpp:719-800, the CASE >= 3 branch of outputFusionCPU: the CPU rules of
// depthFuseCPU (:246-587), run on the GPU by pm_depth_fuse_cpu_rules.  The first keyframe
// only resets (:725-741); the state flags and pose (firstDepth, pre_index, cur_index,
// R_b_2_w, t_b_2_w) are the members outputFusion uses as well, as in the reference.
void StereoMapper::fuse_cpu_rules(Mat& result, Matrix3d Rw, Vector3d tw, float fx, float fy,
                                  float cx, float cy) {
  run_sgm_filter();
  hipStream_t st = (hipStream_t)stream_;
  const size_t N = (size_t)W_ * H_;
  if (!cpu_table.data) {
    for (int h = 0; h < 2; ++h) {
      cpu_alpha[h] = dalloc(H_, W_, CV_32F); cpu_beta[h] = dalloc(H_, W_, CV_32F);
      cpu_mu[h] = dalloc(H_, W_, CV_32F); cpu_sigma[h] = dalloc(H_, W_, CV_32F);
    }
    cpu_table = dalloc(H_, W_, CV_32S);
    cpu_result_ = dalloc(H_, W_, CV_32F);
    uint64_t bytes = 0;
    check(pm_depth_fuse_cpu_rules_workspace_size(W_, H_, &bytes));
    if (bytes > (uint64_t)INT32_MAX)
      throw Error(PM_ERR_INVALID_ARGUMENT, "outputFusionCPU: image too large for the fusion scratch");
    cpu_work_ = dalloc(1, (int)bytes, CV_8U);
    hip_check(hipMemsetAsync(cpu_table.data, 0, cpu_table.bytes(), st), "hipMemsetAsync");
  }
  if (firstDepth) {   // :725-741
    firstDepth = false;
    if (!(result.data && result.rows == H_ && result.cols == W_ && result.type == CV_32F))
      result.create_with(H_, W_, CV_32F, pinned_alloc, pinned_release, &kPinnedTag);
    result.setTo<float>(0.0f);
    R_b_2_w = Rw;
    t_b_2_w = tw;
    pre_index = 0;
    cur_index = 1;
    for (int h = 0; h < 2; ++h)
      for (DeviceMat* m : {&cpu_alpha[h], &cpu_beta[h], &cpu_mu[h], &cpu_sigma[h]})
        check(pm_fill_f32(reinterpret_cast<float*>(m->data), N, 0.0f, stream_));
    synchronize();
    return;
  }
  hip_check(hipMemcpyAsync(cpu_result_.data, dep.data, dep.bytes(), hipMemcpyDeviceToDevice, st),
            "copy depth");
  const Matrix3d R_p_2_c = Rw.transpose() * R_b_2_w;   // :778-779, double
  const Vector3d t_p_2_c = Rw.transpose() * (t_b_2_w - tw);
  auto state = [&](int h) {
    return pm_fusion_state{reinterpret_cast<float*>(cpu_alpha[h].data),
                           reinterpret_cast<float*>(cpu_beta[h].data),
                           reinterpret_cast<float*>(cpu_mu[h].data),
                           reinterpret_cast<float*>(cpu_sigma[h].data)};
  };
  check(pm_depth_fuse_cpu_rules(W_, H_, pm_camera{fx, fy, cx, cy}, R_p_2_c.m, t_p_2_c.v,
                                reinterpret_cast<float*>(cpu_result_.data),
                                reinterpret_cast<int32_t*>(cpu_table.data), cpu_work_.data,
                                cpu_work_.bytes(), state(pre_index), state(pre_index ^ 1),
                                stream_));
  download_result(cpu_result_, result);
  R_b_2_w = Rw;           // :794-798
  t_b_2_w = tw;
  pre_index = cur_index;
  cur_index ^= 1;
}

// stereo_mapper.cpp:625-800
void StereoMapper::outputFusionCPU(Mat& result, Matrix3d Rw, Vector3d tw, float fx, float fy,
                                   float cx, float cy) {
  if (CASE == 0 || CASE == 1) {
    wta(CASE, result);
    return;
  }
  if (CASE == 2) {
    output(result);
    return;
  }
  fuse_cpu_rules(result, Rw, tw, fx, fy, cx, cy);
}

}  // namespace pm

// ---- C wrappers ---------------------------------------------------------------------------

struct pm_mapper {
  pm::StereoMapper* impl;
};

namespace {

template <typename F>
pm_status guard(F&& f) {
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

pm::Mat wrap_image(const uint8_t* img, int H, int W) {
  return pm::Mat(H, W, pm::CV_8U, const_cast<uint8_t*>(img));
}

pm::Mat wrap_d(const double* v, int r, int c) {
  pm::Mat m(r, c, pm::CV_64F);
  std::memcpy(m.data, v, sizeof(double) * r * c);
  return m;
}

float half_to_float(uint16_t h) {
  const uint32_t sign = (uint32_t)(h >> 15) << 31;
  const int exp = (h >> 10) & 0x1F;
  const uint32_t man = h & 0x3FF;
  float f;
  if (exp == 0) {
    f = std::ldexp((float)man, -24);
  } else if (exp == 31) {
    f = man ? NAN : INFINITY;
  } else {
    f = std::ldexp((float)(man | 0x400), exp - 25);
  }
  uint32_t bits;
  std::memcpy(&bits, &f, 4);
  bits |= sign;
  std::memcpy(&f, &bits, 4);
  return f;
}

}  // namespace

extern "C" pm_status pm_mapper_create(int32_t width, int32_t height, int32_t planes,
                                      pm_mapper** out) {
  if (!out) return pm::fail(PM_ERR_INVALID_ARGUMENT, "pm_mapper_create: null out");
  *out = nullptr;
  return guard([&] {
    auto* m = new pm_mapper{nullptr};
    try {
      m->impl = new pm::StereoMapper(width, height, planes);
    } catch (...) {
      delete m;
      throw;
    }
    *out = m;
  });
}

extern "C" void pm_mapper_destroy(pm_mapper* m) {
  if (!m) return;
  delete m->impl;
  delete m;
}

#define PM_MAPPER_CHECK(m) \
  if (!(m) || !(m)->impl) return pm::fail(PM_ERR_INVALID_ARGUMENT, "null mapper")

extern "C" pm_status pm_mapper_init_intrinsic(pm_mapper* m, const double K[9], float bf,
                                              float dep_sample) {
  PM_MAPPER_CHECK(m);
  if (!K) return pm::fail(PM_ERR_INVALID_ARGUMENT, "null K");
  return guard([&] { m->impl->initIntrinsic(wrap_d(K, 3, 3), bf, dep_sample); });
}

extern "C" pm_status pm_mapper_init_reference(pm_mapper* m, const uint8_t* image) {
  PM_MAPPER_CHECK(m);
  if (!image) return pm::fail(PM_ERR_INVALID_ARGUMENT, "null image");
  return guard([&] {
    m->impl->initReference(wrap_image(image, m->impl->height(), m->impl->width()));
  });
}

extern "C" pm_status pm_mapper_update(pm_mapper* m, const uint8_t* image, const double R_l[9],
                                      const double T_l[3], const double R_r[9],
                                      const double T_r[3]) {
  PM_MAPPER_CHECK(m);
  if (!image || !R_l || !T_l || !R_r || !T_r)
    return pm::fail(PM_ERR_INVALID_ARGUMENT, "null argument");
  return guard([&] {
    m->impl->update(wrap_image(image, m->impl->height(), m->impl->width()), wrap_d(R_l, 3, 3),
                    wrap_d(T_l, 3, 1), wrap_d(R_r, 3, 3), wrap_d(T_r, 3, 1));
  });
}

extern "C" pm_status pm_mapper_init_reference_frame(pm_mapper* m, const pm_frame_ring* ring,
                                                    int64_t frame) {
  PM_MAPPER_CHECK(m);
  const pm::FrameRing* r = pm::ring_of(ring);
  if (!r) return pm::fail(PM_ERR_INVALID_ARGUMENT, "null ring");
  return guard([&] { m->impl->initReference(*r, frame); });
}

extern "C" pm_status pm_mapper_update_frame(pm_mapper* m, const pm_frame_ring* ring,
                                            int64_t frame, const double R_l[9],
                                            const double T_l[3], const double R_r[9],
                                            const double T_r[3]) {
  PM_MAPPER_CHECK(m);
  const pm::FrameRing* r = pm::ring_of(ring);
  if (!r || !R_l || !T_l || !R_r || !T_r)
    return pm::fail(PM_ERR_INVALID_ARGUMENT, "null argument");
  return guard([&] {
    m->impl->update(*r, frame, wrap_d(R_l, 3, 3), wrap_d(T_l, 3, 1), wrap_d(R_r, 3, 3),
                    wrap_d(T_r, 3, 1));
  });
}

extern "C" pm_status pm_mapper_output(pm_mapper* m, float* result) {
  PM_MAPPER_CHECK(m);
  if (!result) return pm::fail(PM_ERR_INVALID_ARGUMENT, "null result");
  return guard([&] {
    pm::Mat r(m->impl->height(), m->impl->width(), pm::CV_32F, result);
    m->impl->output(r);
  });
}

extern "C" pm_status pm_mapper_output_fusion(pm_mapper* m, float* result, const double R[9],
                                             const double t[3], float fx, float fy, float cx,
                                             float cy) {
  PM_MAPPER_CHECK(m);
  if (!result || !R || !t) return pm::fail(PM_ERR_INVALID_ARGUMENT, "null argument");
  return guard([&] {
    pm::Matrix3d Rm;
    for (int k = 0; k < 9; ++k) Rm.m[k] = R[k];
    pm::Vector3d tv(t[0], t[1], t[2]);
    pm::Mat r(m->impl->height(), m->impl->width(), pm::CV_32F, result);
    m->impl->outputFusion(r, Rm, tv, fx, fy, cx, cy);
  });
}

extern "C" pm_status pm_mapper_output_fusion_cpu(pm_mapper* m, float* result, const double R[9],
                                                 const double t[3], float fx, float fy,
                                                 float cx, float cy) {
  PM_MAPPER_CHECK(m);
  if (!result || !R || !t) return pm::fail(PM_ERR_INVALID_ARGUMENT, "null argument");
  return guard([&] {
    pm::Matrix3d Rm;
    for (int k = 0; k < 9; ++k) Rm.m[k] = R[k];
    pm::Vector3d tv(t[0], t[1], t[2]);
    pm::Mat r(m->impl->height(), m->impl->width(), pm::CV_32F, result);
    m->impl->outputFusionCPU(r, Rm, tv, fx, fy, cx, cy);
  });
}

extern "C" pm_status pm_mapper_output_case(pm_mapper* m, int32_t case_mode, float* result) {
  PM_MAPPER_CHECK(m);
  if (!result) return pm::fail(PM_ERR_INVALID_ARGUMENT, "null result");
  if (case_mode < 0 || case_mode > 2)
    return pm::fail(PM_ERR_INVALID_ARGUMENT, "case_mode must be 0, 1 or 2");
  return guard([&] {
    pm::Mat r(m->impl->height(), m->impl->width(), pm::CV_32F, result);
    if (case_mode == 2) m->impl->output(r);
    else m->impl->wta(case_mode, r);
  });
}

extern "C" pm_status pm_mapper_set_cost_type(pm_mapper* m, pm_cost_type type) {
  PM_MAPPER_CHECK(m);
  return guard([&] { m->impl->set_cost_type(type); });
}

extern "C" pm_status pm_mapper_download(pm_mapper* m, const char* name, void* host,
                                        size_t bytes) {
  PM_MAPPER_CHECK(m);
  if (!name || !host) return pm::fail(PM_ERR_INVALID_ARGUMENT, "null argument");
  return guard([&] {
    pm::StereoMapper& s = *m->impl;
    const std::string n(name);
    const bool cur0 = s.pre_index == 0;   // after a fusion step the new state sits at pre_index
    const pm::DeviceMat* src = nullptr;
    if (n == "raw_cost") src = &s.raw_cost;
    else if (n == "sgm_cost") src = &s.sgm_cost;
    else if (n == "dep") src = &s.dep;
    else if (n == "fuse_depth") src = &s.fuseDepth;
    else if (n == "table") src = &s.propogate_table;
    else if (n == "img_l") src = &s.img_l;
    else if (n == "alpha") src = cur0 ? &s.alpha0 : &s.alpha1;
    else if (n == "beta") src = cur0 ? &s.beta0 : &s.beta1;
    else if (n == "mu") src = cur0 ? &s.mu0 : &s.mu1;
    else if (n == "sig") src = cur0 ? &s.sigma0 : &s.sigma1;
    else if (n == "alpha0") src = &s.alpha0;
    else if (n == "beta0") src = &s.beta0;
    else if (n == "mu0") src = &s.mu0;
    else if (n == "sigma0") src = &s.sigma0;
    else if (n == "alpha1") src = &s.alpha1;
    else if (n == "beta1") src = &s.beta1;
    else if (n == "mu1") src = &s.mu1;
    else if (n == "sigma1") src = &s.sigma1;
    else if (n == "cpu_alpha") src = &s.cpu_alpha[s.pre_index];
    else if (n == "cpu_beta") src = &s.cpu_beta[s.pre_index];
    else if (n == "cpu_mu") src = &s.cpu_mu[s.pre_index];
    else if (n == "cpu_sig") src = &s.cpu_sigma[s.pre_index];
    else if (n == "cpu_table") src = &s.cpu_table;
    else throw pm::Error(PM_ERR_INVALID_ARGUMENT, "unknown buffer name: " + n);
    s.synchronize();
    const size_t count = (size_t)src->rows * src->cols;
    if (src == &s.raw_cost && src->type != pm::CV_32F) {
      if (bytes != count * sizeof(float))
        throw pm::Error(PM_ERR_INVALID_ARGUMENT, "raw_cost: bytes must be H*W*D*4");
      std::vector<uint16_t> tmp(count);
      pm::hip_check(hipMemcpy(tmp.data(), src->data, count * 2, hipMemcpyDeviceToHost), "download");
      float* out = static_cast<float*>(host);
      for (size_t i = 0; i < count; ++i) {
        if (src->type == pm::PM_16F) {
          out[i] = half_to_float(tmp[i]);
        } else {
          const uint32_t b = (uint32_t)tmp[i] << 16;
          std::memcpy(&out[i], &b, 4);
        }
      }
      return;
    }
    if (bytes != src->bytes())
      throw pm::Error(PM_ERR_INVALID_ARGUMENT,
                      n + ": bytes must be " + std::to_string(src->bytes()));
    pm::hip_check(hipMemcpy(host, src->data, src->bytes(), hipMemcpyDeviceToHost), "download");
  });
}

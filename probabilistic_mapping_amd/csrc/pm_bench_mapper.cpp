// pm_bench_mapper: times a keyframe sequence through the C++ StereoMapper, the drop-in for the
// reference's per-keyframe calls (motion_stereo/stereo_mapper.cpp:63-244, as evaluate.cc:
// 247-260 makes them): per keyframe initReference(image), update(image, poses) per source,
// then outputFusion(R, t, fx, fy, cx, cy), which hands back the filtered depth on the host.
//
//   pm_bench_mapper WORKLOAD OUT_DIR
//
// WORKLOAD is a little-endian binary file:
//   int32  W, H, D, frames, sources, steps, warmup, median (0/1), paths, ring (0/1)
//   double K[9], idep_min, idep_max, pi1, pi2, tau_so, sgm_q1, sgm_q2, var_scale
//   double per frame: R[9] (camera -> world, row-major), T[3]
//   int32  per frame: `sources` source frame indices
//   uint8  per frame: H x W image
// ring = 0 passes host images at every use (each copied into pinned staging and uploaded);
// ring = 1 uploads each frame once per step into a pm::FrameRing and passes frame ids.
// One step runs the whole sequence with the same mapper; its first keyframe resets the fusion
// state as the reference's first outputFusion call does.  Prints one JSON line with the step
// times (and, with the ring, the frames one step uploaded) and writes the last step's fused
// and filtered depth to OUT_DIR/fused.f32 and OUT_DIR/dep.f32 (float32 [H][W]).
#include <chrono>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <iostream>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "pm/frame_ring.h"
#include "pm/parameters.h"
#include "pm/stereo_mapper.h"

namespace {

template <typename T>
void rd(std::ifstream& in, T* v, size_t n) {
  in.read(reinterpret_cast<char*>(v), (std::streamsize)(sizeof(T) * n));
  if (!in) throw std::runtime_error("workload file truncated");
}

void write_raw(const std::string& path, const pm::Mat& m) {
  std::ofstream out(path, std::ios::binary);
  out.write(reinterpret_cast<const char*>(m.data), (std::streamsize)m.total() * 4);
}

}  // namespace

int main(int argc, char** argv) {
  if (argc != 3) {
    std::cerr << "usage: pm_bench_mapper WORKLOAD OUT_DIR\n";
    return 2;
  }
  try {
    std::ifstream in(argv[1], std::ios::binary);
    if (!in) throw std::runtime_error(std::string("cannot open ") + argv[1]);
    int32_t hdr[10];
    rd(in, hdr, 10);
    const int W = hdr[0], H = hdr[1], D = hdr[2], F = hdr[3], S = hdr[4], steps = hdr[5],
              warmup = hdr[6], median = hdr[7], paths = hdr[8], use_ring = hdr[9];
    double kd[9], par[8];
    rd(in, kd, 9);
    rd(in, par, 8);
    std::vector<pm::Matrix3d> Rs(F);
    std::vector<pm::Vector3d> Ts(F);
    std::vector<pm::Mat> Rm(F), Tm(F);
    for (int f = 0; f < F; ++f) {
      double rt[12];
      rd(in, rt, 12);
      Rm[f].create(3, 3, pm::CV_64F);
      Tm[f].create(3, 1, pm::CV_64F);
      for (int k = 0; k < 9; ++k) {
        Rs[f].m[k] = rt[k];
        Rm[f].at<double>(k / 3, k % 3) = rt[k];
      }
      for (int k = 0; k < 3; ++k) {
        Ts[f].v[k] = rt[9 + k];
        Tm[f].at<double>(k, 0) = rt[9 + k];
      }
    }
    std::vector<int32_t> src((size_t)F * S);
    rd(in, src.data(), src.size());
    std::vector<pm::Mat> img(F);
    for (int f = 0; f < F; ++f) {
      img[f].create(H, W, pm::CV_8U);
      rd(in, img[f].data, (size_t)W * H);
    }
    pm::Mat K(3, 3, pm::CV_64F);
    for (int k = 0; k < 9; ++k) K.at<double>(k / 3, k % 3) = kd[k];
    const float fx = (float)kd[0], fy = (float)kd[4], cx = (float)kd[2], cy = (float)kd[5];
    IDEP_MIN = (float)par[0];
    pi1 = (float)par[2]; pi2 = (float)par[3]; tau_so = (float)par[4];
    sgm_q1 = (float)par[5]; sgm_q2 = (float)par[6]; var_scale = (float)par[7];
    SGM_PATHS = paths;
    MEDIAN_FILTER = median;
    const float dep_sample = (float)((par[1] - par[0]) / (D - 1));

    pm::StereoMapper m(W, H, D);
    m.initIntrinsic(K, 0.0f, dep_sample);
    std::unique_ptr<pm::FrameRing> ring(new pm::FrameRing(W, H, F));
    int64_t ring_puts = 0;
    pm::Mat result;
    std::vector<double> step_ms;
    for (int it = 0; it < warmup + steps; ++it) {
      m.firstDepth = true;   // every step starts the sequence over (stereo_mapper.cpp:128)
      const auto t0 = std::chrono::steady_clock::now();
      for (int i = 0; i < F; ++i) {
        if (use_ring) {
          // frames arrive in order; each is uploaded once, when it first appears
          if (!ring->holds(i)) ring->put(i, img[i], nullptr);
          for (int s = 0; s < S; ++s) {
            const int j = src[(size_t)i * S + s];
            if (!ring->holds(j)) ring->put(j, img[j], nullptr);
          }
          m.initReference(*ring, i);
          for (int s = 0; s < S; ++s) {
            const int j = src[(size_t)i * S + s];
            m.update(*ring, j, Rm[i], Tm[i], Rm[j], Tm[j]);
          }
        } else {
          m.initReference(img[i]);
          for (int s = 0; s < S; ++s) {
            const int j = src[(size_t)i * S + s];
            m.update(img[j], Rm[i], Tm[i], Rm[j], Tm[j]);
          }
        }
        m.outputFusion(result, Rs[i], Ts[i], fx, fy, cx, cy);
      }
      const auto t1 = std::chrono::steady_clock::now();
      if (it >= warmup) step_ms.push_back(std::chrono::duration<double, std::milli>(t1 - t0).count());
      if (use_ring) {
        // the next step uploads every frame again (a new ring, outside the timed region)
        m.synchronize();
        ring_puts = ring->puts();
        ring.reset(new pm::FrameRing(W, H, F));
      }
    }
    m.synchronize();
    pm::Mat out;
    m.fuseDepth.download(out);
    write_raw(std::string(argv[2]) + "/fused.f32", out);
    m.dep.download(out);
    write_raw(std::string(argv[2]) + "/dep.f32", out);
    std::printf("{\"steps\": %d, \"warmup\": %d, \"keyframes_per_step\": %d, \"ring\": %d, "
                "\"ring_puts\": %lld, \"step_ms\": [",
                steps, warmup, F, use_ring, (long long)ring_puts);
    for (size_t k = 0; k < step_ms.size(); ++k)
      std::printf("%s%.4f", k ? ", " : "", step_ms[k]);
    std::printf("]}\n");
    return 0;
  } catch (const std::exception& e) {
    std::cerr << "pm_bench_mapper: " << e.what() << "\n";
    return 1;
  }
}

// Runs every oracle function on small seeded inputs, in both instantiations, so that a build
// with -fsanitize=address,undefined checks the oracle's memory accesses and arithmetic.
// TEST INFRASTRUCTURE ONLY: built and run by tests/test_oracle_sanitize.py.  Exit status 0
// when every call returned with finite, in-range results; the sanitizers abort otherwise.
#include <cmath>
#include <cstdio>
#include <random>
#include <vector>

#include "pm_oracle.hpp"

namespace {

template <typename T>
bool run(int W, int H, int D, int S, int paths, double back, unsigned seed) {
  std::mt19937 rng(seed);
  std::uniform_int_distribution<int> pix(0, 255);
  std::uniform_real_distribution<double> u01(0.0, 1.0);
  const long N = (long)W * H;
  std::vector<std::vector<uint8_t>> imgs(S + 1, std::vector<uint8_t>(N));
  for (auto& im : imgs)
    for (auto& p : im) p = (uint8_t)pix(rng);
  std::vector<pmo::Source> src(S);
  for (int s = 0; s < S; ++s) {
    src[s].img = imgs[s + 1].data();
    for (int k = 0; k < 9; ++k) src[s].h[k] = (k % 4 == 0 ? 1.0 : 0.0) + 0.01 * (u01(rng) - 0.5);
    src[s].t[0] = 0.2 * (u01(rng) - 0.5) * W;
    src[s].t[1] = 0.2 * (u01(rng) - 0.5) * H;
    src[s].t[2] = 0.05 * (u01(rng) - 0.5);
  }
  const T idep_min = T(0.2), ds = T(0.8) / T(D > 1 ? D - 1 : 1);
  std::vector<T> cost((size_t)N * D), agg((size_t)N * D), depth(N), med(N);
  pmo::calc_cost<T>(W, H, D, imgs[0].data(), src.data(), S, idep_min, ds, cost.data());
  pmo::Knobs kn;
  kn.paths = paths;
  pmo::sgm<T>(W, H, D, imgs[0].data(), cost.data(), kn, agg.data());
  std::vector<double> gagg((size_t)N * D, 1.0), gcost((size_t)N * D);
  pmo::sgm_adjoint<T>(W, H, D, imgs[0].data(), cost.data(), kn, gagg.data(), gcost.data());
  pmo::filter_cost<T>(W, H, D, agg.data(), T(kn.var_scale), idep_min, ds, depth.data());
  pmo::median3x3<T>(W, H, depth.data(), med.data());
  for (long p = 0; p < N; ++p)
    if (!std::isfinite((double)depth[p]) || !std::isfinite((double)med[p])) return false;

  // two fusion steps of each rule set; `back` moves the camera back along its axis so that
  // many sources share a cell
  const double r[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
  const double t[3] = {0.0, 0.0, back};
  const T fx = T(0.9 * W), fy = T(0.9 * H), cx = T(0.5 * (W - 1)), cy = T(0.5 * (H - 1));
  for (int rules = 0; rules < 2; ++rules) {
    std::vector<T> a[2], b[2], m[2], s[2];
    for (int h = 0; h < 2; ++h) {
      a[h].assign(N, T(0)); b[h].assign(N, T(0)); m[h].assign(N, T(0)); s[h].assign(N, T(0));
    }
    std::vector<int32_t> table(N);
    for (int step = 0; step < 3; ++step) {
      std::vector<T> d(N);
      for (long p = 0; p < N; ++p) d[p] = u01(rng) < 0.1 ? T(0) : T(0.5 + 4.0 * u01(rng));
      const int pre = step & 1;
      pmo::State<T> sp{a[pre].data(), b[pre].data(), m[pre].data(), s[pre].data()};
      pmo::State<T> sc{a[pre ^ 1].data(), b[pre ^ 1].data(), m[pre ^ 1].data(),
                       s[pre ^ 1].data()};
      if (rules)
        pmo::depth_fuse_cpu_rules<T>(W, H, fx, fy, cx, cy, r, t, d.data(), table.data(), sp, sc);
      else
        pmo::depth_fuse<T>(W, H, fx, fy, cx, cy, r, t, d.data(), table.data(), sp, sc);
      for (long p = 0; p < N; ++p) {
        const int32_t e = table[p];
        if (e < 0) return false;
        if (rules) {
          if ((e >> 16) >= H || (e & 0xFFFF) >= W) return false;
        } else if (e >= N) {
          return false;
        }
      }
    }
  }
  return true;
}

}  // namespace

int main() {
  struct Case { int W, H, D, S, paths; double back; };
  const Case cases[] = {{1, 1, 2, 1, 4, 0.5}, {2, 3, 5, 2, 8, 0.0}, {17, 9, 16, 3, 8, 3.0},
                        {33, 20, 33, 4, 4, 20.0}, {64, 1, 8, 2, 8, 1.0}};
  int failed = 0;
  unsigned seed = 1;
  for (const Case& c : cases) {
    const bool ok64 = run<double>(c.W, c.H, c.D, c.S, c.paths, c.back, seed++);
    const bool ok32 = run<float>(c.W, c.H, c.D, c.S, c.paths, c.back, seed++);
    std::printf("%dx%dx%d S=%d P=%d back=%.1f: f64 %s, f32 %s\n", c.W, c.H, c.D, c.S, c.paths,
                c.back, ok64 ? "ok" : "FAIL", ok32 ? "ok" : "FAIL");
    failed += !ok64 + !ok32;
  }
  return failed ? 1 : 0;
}

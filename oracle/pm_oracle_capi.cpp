// C entry points over pm_oracle.hpp for the Python test harness (ctypes).
// TEST INFRASTRUCTURE ONLY: called by tests/, the smoke check and bench.py's cpu_baseline leg.
// Suffix _f64 = the FP64 oracle, _f32 = the float/OpenMP CPU baseline.
#include "pm_oracle.hpp"

#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

template <typename T, typename RT = uint8_t>
void cost_impl(int W, int H, int D, const RT* ref, int S, const uint8_t* const* imgs,
               const double* h, const double* t, double idep_min, double dep_sample, T* cost,
               float* margin = nullptr) {
  std::vector<pmo::Source> src(S);
  for (int s = 0; s < S; ++s) {
    src[s].img = imgs[s];
    for (int k = 0; k < 9; ++k) src[s].h[k] = h[s * 9 + k];
    for (int k = 0; k < 3; ++k) src[s].t[k] = t[s * 3 + k];
  }
  pmo::calc_cost<T, RT>(W, H, D, ref, src.data(), S, T(idep_min), T(dep_sample), cost, margin);
}

pmo::Knobs knobs_from(const double* k, int paths) {
  pmo::Knobs kn;
  kn.pi1 = k[0]; kn.pi2 = k[1]; kn.tau_so = k[2]; kn.sgm_q1 = k[3]; kn.sgm_q2 = k[4];
  kn.var_scale = k[5];
  kn.paths = paths;
  return kn;
}

}  // namespace

extern "C" {

int oracle_max_threads() {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

void oracle_set_threads(int n) {
#ifdef _OPENMP
  if (n > 0) omp_set_num_threads(n);
#else
  (void)n;
#endif
}

void oracle_calc_cost_f64(int W, int H, int D, const uint8_t* ref, int S,
                          const uint8_t* const* imgs, const double* h, const double* t,
                          double idep_min, double dep_sample, double* cost) {
  cost_impl<double>(W, H, D, ref, S, imgs, h, t, idep_min, dep_sample, cost);
}
// FP64 cost and aggregation of a real-valued reference image (the intensity finite
// differences of the differentiable readout's checks)
void oracle_calc_cost_fref_f64(int W, int H, int D, const double* ref, int S,
                               const uint8_t* const* imgs, const double* h, const double* t,
                               double idep_min, double dep_sample, double* cost) {
  cost_impl<double, double>(W, H, D, ref, S, imgs, h, t, idep_min, dep_sample, cost);
}
void oracle_sgm_fref_f64(int W, int H, int D, const double* ref, const double* cost,
                         const double* knobs, int paths, double* agg) {
  pmo::sgm<double, double>(W, H, D, ref, cost, knobs_from(knobs, paths), agg);
}
// the FP64 cost with each entry's validity margin (pmo::calc_cost)
void oracle_calc_cost_margin_f64(int W, int H, int D, const uint8_t* ref, int S,
                                 const uint8_t* const* imgs, const double* h, const double* t,
                                 double idep_min, double dep_sample, double* cost,
                                 float* margin) {
  cost_impl<double>(W, H, D, ref, S, imgs, h, t, idep_min, dep_sample, cost, margin);
}
void oracle_calc_cost_f32(int W, int H, int D, const uint8_t* ref, int S,
                          const uint8_t* const* imgs, const double* h, const double* t,
                          double idep_min, double dep_sample, float* cost) {
  cost_impl<float>(W, H, D, ref, S, imgs, h, t, idep_min, dep_sample, cost);
}

// knobs = {pi1, pi2, tau_so, sgm_q1, sgm_q2, var_scale}
void oracle_sgm_f64(int W, int H, int D, const uint8_t* ref, const double* cost,
                    const double* knobs, int paths, double* agg) {
  pmo::sgm<double>(W, H, D, ref, cost, knobs_from(knobs, paths), agg);
}
void oracle_sgm_f32(int W, int H, int D, const uint8_t* ref, const float* cost,
                    const double* knobs, int paths, float* agg) {
  pmo::sgm<float>(W, H, D, ref, cost, knobs_from(knobs, paths), agg);
}

// dL/dC from dL/dagg; decisions in the cost's precision, gradients in double
void oracle_sgm_adjoint_f64(int W, int H, int D, const uint8_t* ref, const double* cost,
                            const double* knobs, int paths, const double* gagg, double* gcost) {
  pmo::sgm_adjoint<double>(W, H, D, ref, cost, knobs_from(knobs, paths), gagg, gcost);
}
void oracle_sgm_adjoint_f32(int W, int H, int D, const uint8_t* ref, const float* cost,
                            const double* knobs, int paths, const double* gagg, double* gcost) {
  pmo::sgm_adjoint<float>(W, H, D, ref, cost, knobs_from(knobs, paths), gagg, gcost);
}

void oracle_filter_f64(int W, int H, int D, const double* agg, double var_scale,
                       double idep_min, double dep_sample, double* depth) {
  pmo::filter_cost<double>(W, H, D, agg, var_scale, idep_min, dep_sample, depth);
}
void oracle_filter_f32(int W, int H, int D, const float* agg, double var_scale,
                       double idep_min, double dep_sample, float* depth) {
  pmo::filter_cost<float>(W, H, D, agg, float(var_scale), float(idep_min), float(dep_sample),
                          depth);
}

void oracle_median_f64(int W, int H, const double* in, double* out) {
  pmo::median3x3<double>(W, H, in, out);
}
void oracle_median_f32(int W, int H, const float* in, float* out) {
  pmo::median3x3<float>(W, H, in, out);
}

// intr = {fx, fy, cx, cy}; r = R_p2c row-major 3x3, t = t_p2c;
// pre / cur = {alpha, beta, mu, sig} plane pointers.
void oracle_fuse_f64(int W, int H, const double* intr, const double* r, const double* t,
                     double* depth, int32_t* table, double* const* pre, double* const* cur) {
  pmo::State<double> p{pre[0], pre[1], pre[2], pre[3]}, c{cur[0], cur[1], cur[2], cur[3]};
  pmo::depth_fuse<double>(W, H, intr[0], intr[1], intr[2], intr[3], r, t, depth, table, p, c);
}
void oracle_fuse_f32(int W, int H, const double* intr, const double* r, const double* t,
                     float* depth, int32_t* table, float* const* pre, float* const* cur) {
  pmo::State<float> p{pre[0], pre[1], pre[2], pre[3]}, c{cur[0], cur[1], cur[2], cur[3]};
  pmo::depth_fuse<float>(W, H, float(intr[0]), float(intr[1]), float(intr[2]), float(intr[3]),
                         r, t, depth, table, p, c);
}

// The CPU-rules step of outputFusionCPU (CASE >= 3); same arguments as oracle_fuse_*.
void oracle_fuse_cpu_rules_f64(int W, int H, const double* intr, const double* r,
                               const double* t, double* depth, int32_t* table,
                               double* const* pre, double* const* cur) {
  pmo::State<double> p{pre[0], pre[1], pre[2], pre[3]}, c{cur[0], cur[1], cur[2], cur[3]};
  pmo::depth_fuse_cpu_rules<double>(W, H, intr[0], intr[1], intr[2], intr[3], r, t, depth,
                                    table, p, c);
}
void oracle_fuse_cpu_rules_f32(int W, int H, const double* intr, const double* r,
                               const double* t, float* depth, int32_t* table,
                               float* const* pre, float* const* cur) {
  pmo::State<float> p{pre[0], pre[1], pre[2], pre[3]}, c{cur[0], cur[1], cur[2], cur[3]};
  pmo::depth_fuse_cpu_rules<float>(W, H, float(intr[0]), float(intr[1]), float(intr[2]),
                                   float(intr[3]), r, t, depth, table, p, c);
}

}  // extern "C"

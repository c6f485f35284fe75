// C-ABI plumbing: per-thread error messages, version, and the FP64 homography setup.
#include <cmath>
#include <string>

#include "common.hpp"

namespace pm {

namespace {
thread_local std::string g_last_error;
}

void set_error(const std::string& msg) { g_last_error = msg; }
void clear_error() { g_last_error.clear(); }

pm_status fail(pm_status code, const std::string& msg) {
  g_last_error = msg;
  return code;
}

pm_status check_launch(const char* what) {
  const hipError_t e = hipGetLastError();
  if (e == hipSuccess) return PM_OK;
  return fail(PM_ERR_LAUNCH, std::string(what) + ": " + hipGetErrorString(e));
}

}  // namespace pm

using namespace pm;

extern "C" const char* pm_last_error(void) { return g_last_error.c_str(); }

extern "C" const char* pm_version(void) { return "probabilistic_mapping_amd 0.1 (gfx950)"; }

namespace pm {

bool homography_fp64(const double K[9], const double R_l[9], const double T_l[3],
                     const double R_r[9], const double T_r[3], double Hm[9], double t[3]) {
  for (int k = 0; k < 9; ++k)
    if (!std::isfinite(K[k]) || !std::isfinite(R_l[k]) || !std::isfinite(R_r[k])) return false;
  for (int k = 0; k < 3; ++k)
    if (!std::isfinite(T_l[k]) || !std::isfinite(T_r[k])) return false;
  // K^-1 by the adjugate
  const double a = K[0], b = K[1], c = K[2], d = K[3], e = K[4], f = K[5], g = K[6], h = K[7],
               i = K[8];
  const double det = a * (e * i - f * h) - b * (d * i - f * g) + c * (d * h - e * g);
  if (!(std::fabs(det) > 0.0) || !std::isfinite(det)) return false;
  const double Ki[9] = {(e * i - f * h) / det, (c * h - b * i) / det, (b * f - c * e) / det,
                        (f * g - d * i) / det, (a * i - c * g) / det, (c * d - a * f) / det,
                        (d * h - e * g) / det, (b * g - a * h) / det, (a * e - b * d) / det};
  double M[9];   // K R_r^T
  for (int r = 0; r < 3; ++r)
    for (int col = 0; col < 3; ++col) {
      double s = 0;
      for (int k = 0; k < 3; ++k) s += K[r * 3 + k] * R_r[col * 3 + k];
      M[r * 3 + col] = s;
    }
  double MR[9];  // K R_r^T R_l
  for (int r = 0; r < 3; ++r)
    for (int col = 0; col < 3; ++col) {
      double s = 0;
      for (int k = 0; k < 3; ++k) s += M[r * 3 + k] * R_l[k * 3 + col];
      MR[r * 3 + col] = s;
    }
  for (int r = 0; r < 3; ++r)
    for (int col = 0; col < 3; ++col) {
      double s = 0;
      for (int k = 0; k < 3; ++k) s += MR[r * 3 + k] * Ki[k * 3 + col];
      Hm[r * 3 + col] = s;
    }
  const double dT[3] = {T_l[0] - T_r[0], T_l[1] - T_r[1], T_l[2] - T_r[2]};
  for (int r = 0; r < 3; ++r) {
    double s = 0;
    for (int k = 0; k < 3; ++k) s += M[r * 3 + k] * dT[k];
    t[r] = s;
  }
  return true;
}

}  // namespace pm

// FP64 setup rounded to float, as the reference passes it to ad_calc_cost
// (motion_stereo/stereo_mapper.cpp:79-90).
extern "C" pm_status pm_homography_setup(const double K[9], const double R_l[9],
                                         const double T_l[3], const double R_r[9],
                                         const double T_r[3], pm_homography* out) {
  clear_error();
  if (!K || !R_l || !T_l || !R_r || !T_r || !out)
    return fail(PM_ERR_INVALID_ARGUMENT, "pm_homography_setup: null pointer");
  double Hm[9], t[3];
  if (!homography_fp64(K, R_l, T_l, R_r, T_r, Hm, t))
    return fail(PM_ERR_INVALID_ARGUMENT, "pm_homography_setup: non-finite input or singular K");
  for (int k = 0; k < 9; ++k) out->h[k] = (float)Hm[k];
  for (int k = 0; k < 3; ++k) out->t[k] = (float)t[k];
  return PM_OK;
}

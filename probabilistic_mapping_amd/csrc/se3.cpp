// se3 pose perturbations of a measurement frame and the chain rule from the plane-sweep
// homography (H, t) to the pose (host code, FP64).  No counterpart in the reference, whose
// poses enter only through the homography setup of StereoMapper::update
// (motion_stereo/stereo_mapper.cpp:79-80): with the relative pose (R, t) taking reference
// camera coordinates to the measurement camera, H = K R K^-1 and t_h = K t.
//
// Convention: xi = (rho, phi) in R^6 (translation part first); Exp(xi) = [[C, J rho], [0, 1]]
// with C = exp(phi^) (Rodrigues) and J the SO(3) left Jacobian; the perturbed pose is
// (R, t) = Exp(xi) (R0, t0).  The gradient with respect to xi is J_se3(xi)^T times the
// gradient with respect to a left perturbation of (R, t), where J_se3 = [[J, Q], [0, J]] is the
// SE(3) left Jacobian (Barfoot, "State Estimation for Robotics", 2017, eqs. 7.83-7.86).
#include <cmath>
#include <cstring>

#include "common.hpp"

namespace pm {
namespace {

void hat(const double v[3], double out[9]) {
  out[0] = 0.0;   out[1] = -v[2]; out[2] = v[1];
  out[3] = v[2];  out[4] = 0.0;   out[5] = -v[0];
  out[6] = -v[1]; out[7] = v[0];  out[8] = 0.0;
}

void mul(const double a[9], const double b[9], double out[9]) {
  double t[9];
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) {
      double s = 0.0;
      for (int k = 0; k < 3; ++k) s += a[r * 3 + k] * b[k * 3 + c];
      t[r * 3 + c] = s;
    }
  std::memcpy(out, t, sizeof(t));
}

void mulv(const double a[9], const double v[3], double out[3]) {
  double t[3];
  for (int r = 0; r < 3; ++r) t[r] = a[r * 3] * v[0] + a[r * 3 + 1] * v[1] + a[r * 3 + 2] * v[2];
  std::memcpy(out, t, sizeof(t));
}

void mulTv(const double a[9], const double v[3], double out[3]) {   // a^T v
  double t[3];
  for (int c = 0; c < 3; ++c) t[c] = a[c] * v[0] + a[3 + c] * v[1] + a[6 + c] * v[2];
  std::memcpy(out, t, sizeof(t));
}

// Coefficients of exp(phi^) = I + s1 phi^ + s2 phi^ phi^, of the left Jacobian
// J = I + s2 phi^ + s3 phi^ phi^, and of Q; Taylor series below theta = 0.1 where the closed
// forms cancel.
struct Coef {
  double s1, s2, s3, q1, q2, q3;
};
Coef coefficients(double th) {
  Coef c;
  const double t2 = th * th;
  if (th < 0.1) {
    const double t4 = t2 * t2, t6 = t4 * t2;
    c.s1 = 1.0 - t2 / 6.0 + t4 / 120.0 - t6 / 5040.0;              // sin t / t
    c.s2 = 0.5 - t2 / 24.0 + t4 / 720.0 - t6 / 40320.0;             // (1 - cos t) / t^2
    c.s3 = 1.0 / 6.0 - t2 / 120.0 + t4 / 5040.0 - t6 / 362880.0;    // (t - sin t) / t^3
    c.q1 = c.s3;
    c.q2 = 1.0 / 24.0 - t2 / 720.0 + t4 / 40320.0 - t6 / 3628800.0;  // (t^2 + 2 cos t - 2) / (2 t^4)
    c.q3 = 1.0 / 120.0 - t2 / 2520.0 + t4 / 120960.0;               // (2t - 3 sin t + t cos t) / (2 t^5)
  } else {
    const double s = std::sin(th), co = std::cos(th);
    c.s1 = s / th;
    c.s2 = (1.0 - co) / t2;
    c.s3 = (th - s) / (t2 * th);
    c.q1 = c.s3;
    c.q2 = (t2 + 2.0 * co - 2.0) / (2.0 * t2 * t2);
    c.q3 = (2.0 * th - 3.0 * s + th * co) / (2.0 * t2 * t2 * th);
  }
  return c;
}

void identity(double out[9]) {
  for (int k = 0; k < 9; ++k) out[k] = (k % 4 == 0) ? 1.0 : 0.0;
}

// C = exp(phi^), J = SO(3) left Jacobian
void so3(const double phi[3], double C[9], double J[9]) {
  const double th = std::sqrt(phi[0] * phi[0] + phi[1] * phi[1] + phi[2] * phi[2]);
  const Coef c = coefficients(th);
  double P[9], P2[9];
  hat(phi, P);
  mul(P, P, P2);
  identity(C);
  identity(J);
  for (int k = 0; k < 9; ++k) {
    C[k] += c.s1 * P[k] + c.s2 * P2[k];
    J[k] += c.s2 * P[k] + c.s3 * P2[k];
  }
}

// Q(rho, phi) of the SE(3) left Jacobian
void se3_q(const double rho[3], const double phi[3], double Q[9]) {
  const double th = std::sqrt(phi[0] * phi[0] + phi[1] * phi[1] + phi[2] * phi[2]);
  const Coef c = coefficients(th);
  double Pp[9], Pr[9], PpPr[9], PrPp[9], PpPrPp[9], PpPp[9], PpPpPr[9], PrPpPp[9],
      PpPrPpPp[9], PpPpPrPp[9];
  hat(phi, Pp);
  hat(rho, Pr);
  mul(Pp, Pr, PpPr);
  mul(Pr, Pp, PrPp);
  mul(PpPr, Pp, PpPrPp);
  mul(Pp, Pp, PpPp);
  mul(PpPp, Pr, PpPpPr);
  mul(PrPp, Pp, PrPpPp);
  mul(PpPrPp, Pp, PpPrPpPp);
  mul(PpPp, PrPp, PpPpPrPp);
  for (int k = 0; k < 9; ++k)
    Q[k] = 0.5 * Pr[k] + c.q1 * (PpPr[k] + PrPp[k] + PpPrPp[k]) +
           c.q2 * (PpPpPr[k] + PrPpPp[k] - 3.0 * PpPrPp[k]) + c.q3 * (PpPrPpPp[k] + PpPpPrPp[k]);
}

bool finite_all(const double* v, int n) {
  for (int k = 0; k < n; ++k)
    if (!std::isfinite(v[k])) return false;
  return true;
}

// (R, t) = Exp(xi) (R0, t0)
void apply(const double xi[6], const double R0[9], const double t0[3], double R[9], double t[3]) {
  double C[9], J[9], Jr[3], Ct[3];
  so3(xi + 3, C, J);
  mulv(J, xi, Jr);
  mul(C, R0, R);
  mulv(C, t0, Ct);
  for (int k = 0; k < 3; ++k) t[k] = Ct[k] + Jr[k];
}

}  // namespace
}  // namespace pm

using namespace pm;

extern "C" pm_status pm_se3_homography(const double K[9], const double R0[9], const double t0[3],
                                       const double xi[6], double H[9], double t[3]) {
  clear_error();
  if (!K || !R0 || !t0 || !xi || !H || !t)
    return fail(PM_ERR_INVALID_ARGUMENT, "pm_se3_homography: null pointer");
  if (!finite_all(xi, 6) || !finite_all(R0, 9) || !finite_all(t0, 3))
    return fail(PM_ERR_INVALID_ARGUMENT, "pm_se3_homography: non-finite input");
  double R[9], tr[3];
  apply(xi, R0, t0, R, tr);
  const double I[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1}, Z[3] = {0, 0, 0};
  // H = K R_r^T R_l K^-1, t = K R_r^T (T_l - T_r) with R_l = R, T_l = t, R_r = I, T_r = 0
  if (!homography_fp64(K, R, tr, I, Z, H, t))
    return fail(PM_ERR_INVALID_ARGUMENT, "pm_se3_homography: singular K");
  return PM_OK;
}

extern "C" pm_status pm_se3_pose_grad(const double K[9], const double R0[9], const double t0[3],
                                      const double xi[6], const double grad_hom[12],
                                      double grad_xi[6]) {
  clear_error();
  if (!K || !R0 || !t0 || !xi || !grad_hom || !grad_xi)
    return fail(PM_ERR_INVALID_ARGUMENT, "pm_se3_pose_grad: null pointer");
  if (!finite_all(K, 9) || !finite_all(xi, 6) || !finite_all(R0, 9) || !finite_all(t0, 3) ||
      !finite_all(grad_hom, 12))
    return fail(PM_ERR_INVALID_ARGUMENT, "pm_se3_pose_grad: non-finite input");
  // K^-1 by the adjugate
  double Kinv[9];
  {
    const double a = K[0], b = K[1], c = K[2], d = K[3], e = K[4], f = K[5], g = K[6],
                 h = K[7], i = K[8];
    const double det = a * (e * i - f * h) - b * (d * i - f * g) + c * (d * h - e * g);
    if (!(std::fabs(det) > 0.0))
      return fail(PM_ERR_INVALID_ARGUMENT, "pm_se3_pose_grad: singular K");
    const double adj[9] = {(e * i - f * h), (c * h - b * i), (b * f - c * e),
                           (f * g - d * i), (a * i - c * g), (c * d - a * f),
                           (d * h - e * g), (b * g - a * h), (a * e - b * d)};
    for (int k = 0; k < 9; ++k) Kinv[k] = adj[k] / det;
  }
  double R[9], t[3];
  apply(xi, R0, t0, R, t);
  // dL/dR = K^T G K^-T, dL/dt = K^T g   (H = K R K^-1, t_h = K t)
  double KT[9], KinvT[9], G[9], GR[9], gt[3];
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) {
      KT[r * 3 + c] = K[c * 3 + r];
      KinvT[r * 3 + c] = Kinv[c * 3 + r];
    }
  std::memcpy(G, grad_hom, sizeof(G));
  mul(KT, G, GR);
  mul(GR, KinvT, GR);
  mulTv(K, grad_hom + 9, gt);
  // left perturbation: dR = [w]x R, dt = [w]x t + v
  double RT[9], M[9];
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) RT[r * 3 + c] = R[c * 3 + r];
  mul(GR, RT, M);
  const double g_rho[3] = {gt[0], gt[1], gt[2]};
  const double g_phi[3] = {M[7] - M[5] + (t[1] * gt[2] - t[2] * gt[1]),
                           M[2] - M[6] + (t[2] * gt[0] - t[0] * gt[2]),
                           M[3] - M[1] + (t[0] * gt[1] - t[1] * gt[0])};
  // grad_xi = J_se3(xi)^T (g_rho, g_phi) = (J^T g_rho, Q^T g_rho + J^T g_phi)
  double C[9], J[9], Q[9], a[3], b[3], c3[3];
  so3(xi + 3, C, J);
  se3_q(xi, xi + 3, Q);
  mulTv(J, g_rho, a);
  mulTv(Q, g_rho, b);
  mulTv(J, g_phi, c3);
  for (int k = 0; k < 3; ++k) {
    grad_xi[k] = a[k];
    grad_xi[3 + k] = b[k] + c3[k];
  }
  return PM_OK;
}

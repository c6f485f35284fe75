// Knob globals with the reference's defaults (motion_stereo/parameters.cpp:10-18), plus the
// additive knobs and their C accessors.
#include "pm/parameters.h"

#include "pm_capi.h"

int CASE = 0;

float pi1 = 8;
float pi2 = 24;
float tau_so = 4;
float sgm_q1 = 2;
float sgm_q2 = 3;

float var_scale = 1.02f;

float IDEP_MIN = 0.0f;
int SGM_PATHS = 4;
int MEDIAN_FILTER = 0;

extern "C" void pm_get_knobs(pm_knobs* out) {
  if (!out) return;
  out->pi1 = pi1; out->pi2 = pi2; out->tau_so = tau_so; out->sgm_q1 = sgm_q1;
  out->sgm_q2 = sgm_q2; out->var_scale = var_scale; out->case_mode = CASE;
  out->idep_min = IDEP_MIN; out->sgm_paths = SGM_PATHS; out->median = MEDIAN_FILTER;
}

extern "C" void pm_set_knobs(const pm_knobs* in) {
  if (!in) return;
  pi1 = in->pi1; pi2 = in->pi2; tau_so = in->tau_so; sgm_q1 = in->sgm_q1;
  sgm_q2 = in->sgm_q2; var_scale = in->var_scale; CASE = in->case_mode;
  IDEP_MIN = in->idep_min; SGM_PATHS = in->sgm_paths; MEDIAN_FILTER = in->median;
}

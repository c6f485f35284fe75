// Knobs and constants of the depth pipeline.  Same names, types and defaults as
// motion_stereo/parameters.h:6-30 and parameters.cpp:10-18 of the reference, so code written
// against the reference's globals compiles unchanged.  WIDTH / HEIGHT / DEP_CNT are only the
// defaults of StereoMapper's default constructor: shapes are runtime values here.
#ifndef PM_PARAMETERS_H
#define PM_PARAMETERS_H

#define FLOAT_EPS 1e-5

const int WIDTH = 640;
const int HEIGHT = 480;

extern int CASE;   // outputFusionCPU mode: 0 raw-cost WTA, 1 SGM WTA, 2 filtered depth

const int DEP_CNT = 64;
const float SAD_WEIGHT = 1.0f;

const int MARGIN = 2;
const int PATCH_SIZE = (2 * MARGIN + 1) * (2 * MARGIN + 1);

extern float pi1;        // SGM P1 (cost units)
extern float pi2;        // SGM P2
extern float tau_so;     // intensity step below which P1, P2 are scaled
extern float sgm_q1;     // P1 scale on smooth intensity
extern float sgm_q2;     // P2 scale on smooth intensity
extern float var_scale;  // curvature test of the subpixel fit

const float DEP_INF_1 = 2000.0f;
const float DEP_INF = 1000.0f;
const float COST_INF = 0.0f;

// Additive knobs (no reference counterpart; the defaults reproduce the reference):
extern float IDEP_MIN;     // inverse depth of plane 0 (reference: 0, plane 0 at infinity)
extern int SGM_PATHS;      // 4 = the reference's L->R, R->L, T->B, B->T; 8 adds diagonals
extern int MEDIAN_FILTER;  // 1 = 3x3 median of valid depths after the filter

#endif  // PM_PARAMETERS_H

// Plane-sweep matching cost, fused over all measurement frames of a keyframe.
//
// Reference: ADCalcCostKernel + ad_calc_cost, motion_stereo/calc_cost.cu:11-253, called once
// per measurement frame by StereoMapper::update (stereo_mapper.cpp:73-91) on a volume reset
// to -1 (stereo_mapper.cpp:69-70).  The reference keeps a running mean in the cost volume and
// a second float volume of counts, re-reading and re-writing both per source (16 B per element
// per source).  Here one launch visits every source while the per-(pixel, plane) sum and count
// stay in registers, and the volume is written once.
//
// Mapping: a wave owns a run of PX pixels of one row; lane l handles plane d = dbase + l.
// Everything that depends only on the pixel (the 9 reference samples, the 9 homography
// triples) is wave-uniform; each lane evaluates the 9 warps for its inverse depth.
//
// Bilinear sampling reads the "quad" texture: one uint32 per texel packing
// I(x,y), I(x+1,y), I(x,y+1), I(x+1,y+1) (border 0), so a sample is a single dword gather.
#include "common.hpp"

namespace pm {

struct CostArgs {
  int W, H, D, S;
  float idep_min, dep_sample;
  const uint8_t* ref;
  const uint32_t* quad[PM_MAX_SOURCES];
  float hm[PM_MAX_SOURCES][12];   // h11..h33, t1..t3
};

__global__ __launch_bounds__(256) void make_quad_kernel(const uint8_t* __restrict__ img,
                                                        uint32_t* __restrict__ quad, int W,
                                                        int H) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x;
  const int y = blockIdx.y;
  if (x >= W) return;
  const uint32_t i00 = img[(long)y * W + x];
  const uint32_t i10 = x + 1 < W ? img[(long)y * W + x + 1] : 0u;
  const uint32_t i01 = y + 1 < H ? img[(long)(y + 1) * W + x] : 0u;
  const uint32_t i11 = (x + 1 < W && y + 1 < H) ? img[(long)(y + 1) * W + x + 1] : 0u;
  quad[(long)y * W + x] = i00 | (i10 << 8) | (i01 << 16) | (i11 << 24);
}

// |ref - bilinear(meas, U, V)| with the warp's validity (calc_cost.cu:85-93).  The validity
// test is written positively so a NaN warp (0/0) is invalid; invalid warps read texel 0 so
// the gather stays in bounds.
__device__ __forceinline__ float warp_term(const uint32_t* __restrict__ quad, int W, float Wm1,
                                           float Hm1, float X, float Y, float Z, float t1,
                                           float t2, float t3, float idep, float r,
                                           bool& ok) {
  const float w = Z + t3 * idep;
  const float rw = __builtin_amdgcn_rcpf(w);
  const float U = (X + t1 * idep) * rw;
  const float V = (Y + t2 * idep) * rw;
  const bool valid = (w >= 0.0f) && (U >= 0.0f) && (U <= Wm1) && (V >= 0.0f) && (V <= Hm1);
  ok = ok && valid;
  const float Uc = valid ? U : 0.0f, Vc = valid ? V : 0.0f;
  const int x0 = (int)Uc, y0 = (int)Vc;   // non-negative, so truncation is floor
  const float ax = Uc - (float)x0, ay = Vc - (float)y0;
  const uint32_t q = quad[y0 * W + x0];
  const float i00 = (float)(q & 0xFFu), i10 = (float)((q >> 8) & 0xFFu);
  const float i01 = (float)((q >> 16) & 0xFFu), i11 = (float)(q >> 24);
  const float top = i00 + ax * (i10 - i00);
  const float bot = i01 + ax * (i11 - i01);
  return fabsf(r - (top + ay * (bot - top)));
}

template <typename CT, int PX>
__global__ __launch_bounds__(256) void cost_kernel(const CostArgs a, CT* __restrict__ cost) {
  const int lane = threadIdx.x & 63;
  const int y = blockIdx.y;
  const int xs = (blockIdx.x * 4 + (threadIdx.x >> 6)) * PX;
  if (xs >= a.W) return;
  const int W = a.W, H = a.H, D = a.D;
  const float Wm1 = (float)(W - 1), Hm1 = (float)(H - 1);
  for (int dbase = 0; dbase < D; dbase += kWave) {
    const int d = dbase + lane;
    const float idep = a.idep_min + (float)d * a.dep_sample;
    for (int i = 0; i < PX; ++i) {
      const int u = xs + i;
      if (u >= W) break;
      // reference samples, calc_cost.cu:93-198 (note the swapped up/down rows of terms 1, 2)
      const float r[9] = {texel_u8(a.ref, W, H, u, y),         texel_u8(a.ref, W, H, u, y + 1),
                          texel_u8(a.ref, W, H, u, y - 1),     texel_u8(a.ref, W, H, u - 1, y),
                          texel_u8(a.ref, W, H, u + 1, y),     texel_u8(a.ref, W, H, u - 1, y - 1),
                          texel_u8(a.ref, W, H, u + 1, y + 1), texel_u8(a.ref, W, H, u - 1, y + 1),
                          texel_u8(a.ref, W, H, u + 1, y - 1)};
      float sum = 0.0f, cnt = 0.0f;
      for (int s = 0; s < a.S; ++s) {
        const float* hm = a.hm[s];
        const float h11 = hm[0], h12 = hm[1], h13 = hm[2], h21 = hm[3], h22 = hm[4],
                    h23 = hm[5], h31 = hm[6], h32 = hm[7], h33 = hm[8];
        const float t1 = hm[9], t2 = hm[10], t3 = hm[11];
        const float x = h11 * u + h12 * y + h13;   // calc_cost.cu:28-30
        const float yy = h21 * u + h22 * y + h23;
        const float z = h31 * u + h32 * y + h33;
        const float xr = x + h11;
        // warp triples, calc_cost.cu:34-71, including the garbled right (:50-51) and
        // up-right (:70-71) neighbours
        const float X[9] = {x, x - h12, x + h12, x - h11, xr, x - h12 - h11, x + h12 + h11,
                            x - h11 + h12, xr - h12};
        const float Y[9] = {yy, yy - h22, yy + h22, yy - h21, x + h21, yy - h22 - h21,
                            yy + h22 + h21, yy - h21 + h22, xr - h22};
        const float Z[9] = {z, z - h32, z + h32, z - h31, x + h31, z - h32 - h31,
                            z + h32 + h31, z - h31 + h32, xr - h32};
        const uint32_t* q = a.quad[s];
        bool ok = true;
        float tmp = 0.0f;
#pragma unroll
        for (int k = 0; k < 9; ++k)
          tmp += warp_term(q, W, Wm1, Hm1, X[k], Y[k], Z[k], t1, t2, t3, idep, r[k], ok);
        if (ok) {   // a source counts only if all 9 warps are valid (calc_cost.cu:89-91)
          sum += tmp;
          cnt += 1.0f;
        }
      }
      if (d < D) {
        // mean over sources of (sum of 9 terms)/9, or -1 if no source saw it
        const float c = cnt > 0.0f ? sum / (9.0f * cnt) : -1.0f;
        cost[((long)y * W + u) * D + d] = store_cost<CT>(c);
      }
    }
  }
}

}  // namespace pm

using namespace pm;

extern "C" pm_status pm_make_quad(const uint8_t* image, uint32_t* quad, int32_t width,
                                  int32_t height, pm_stream stream) {
  clear_error();
  if (!image || !quad) return fail(PM_ERR_INVALID_ARGUMENT, "pm_make_quad: null pointer");
  if (width <= 0 || height <= 0 || width > 65535 || height > 65535)
    return fail(PM_ERR_INVALID_ARGUMENT, "pm_make_quad: bad image size");
  dim3 grid((width + 255) / 256, height);
  make_quad_kernel<<<grid, 256, 0, (hipStream_t)stream>>>(image, quad, width, height);
  return check_launch("make_quad_kernel");
}

extern "C" pm_status pm_calc_cost(pm_shape shape, const uint8_t* ref, int32_t num_sources,
                                  const uint32_t* const* src_quads, const pm_homography* homs,
                                  float idep_min, float dep_sample, void* cost,
                                  pm_cost_type cost_type, pm_stream stream) {
  clear_error();
  if (!ref || !cost) return fail(PM_ERR_INVALID_ARGUMENT, "pm_calc_cost: null pointer");
  if (shape.width <= 0 || shape.height <= 0 || shape.width > 65535 || shape.height > 65535)
    return fail(PM_ERR_INVALID_ARGUMENT, "pm_calc_cost: bad image size");
  if (shape.planes <= 0 || shape.planes > PM_MAX_PLANES)
    return fail(PM_ERR_INVALID_ARGUMENT, "pm_calc_cost: planes must be in 1..256");
  if (num_sources < 0 || num_sources > PM_MAX_SOURCES)
    return fail(PM_ERR_INVALID_ARGUMENT, "pm_calc_cost: num_sources must be in 0..16");
  if (num_sources > 0 && (!src_quads || !homs))
    return fail(PM_ERR_INVALID_ARGUMENT, "pm_calc_cost: null source arrays");
  if (!std::isfinite(idep_min) || !std::isfinite(dep_sample))
    return fail(PM_ERR_INVALID_ARGUMENT, "pm_calc_cost: non-finite plane spacing");
  CostArgs a{};
  a.W = shape.width; a.H = shape.height; a.D = shape.planes; a.S = num_sources;
  a.idep_min = idep_min; a.dep_sample = dep_sample; a.ref = ref;
  for (int s = 0; s < num_sources; ++s) {
    if (!src_quads[s]) return fail(PM_ERR_INVALID_ARGUMENT, "pm_calc_cost: null source quad");
    a.quad[s] = src_quads[s];
    for (int k = 0; k < 9; ++k) a.hm[s][k] = homs[s].h[k];
    for (int k = 0; k < 3; ++k) a.hm[s][9 + k] = homs[s].t[k];
    for (int k = 0; k < 12; ++k)
      if (!std::isfinite(a.hm[s][k]))
        return fail(PM_ERR_INVALID_ARGUMENT, "pm_calc_cost: non-finite homography");
  }
  constexpr int PX = 8;
  dim3 grid((a.W + 4 * PX - 1) / (4 * PX), a.H);
  hipStream_t st = (hipStream_t)stream;
  switch (cost_type) {
    case PM_COST_F32:
      cost_kernel<float, PX><<<grid, 256, 0, st>>>(a, (float*)cost);
      break;
    case PM_COST_F16:
      cost_kernel<__half, PX><<<grid, 256, 0, st>>>(a, (__half*)cost);
      break;
    case PM_COST_BF16:
      cost_kernel<__hip_bfloat16, PX><<<grid, 256, 0, st>>>(a, (__hip_bfloat16*)cost);
      break;
    default:
      return fail(PM_ERR_INVALID_ARGUMENT, "pm_calc_cost: unknown cost type");
  }
  return check_launch("cost_kernel");
}

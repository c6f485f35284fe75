// Semi-global aggregation over 4 or 8 scanline directions.
//
// Reference: sgm2<idx,start,dx,dy,n> and its host driver, motion_stereo/sgm.cu:5-135.  The
// reference runs one block per scanline with one thread per plane, and finds min_d L with a
// shared-memory tree and log2(D) __syncthreads per step (sgm.cu:55-66).  Here a scanline is
// one wave64: lane l holds planes d = NV*l .. NV*l+NV-1, the minimum over planes is a DPP
// row reduction plus four readlanes, and the d-1 / d+1 neighbours come from DPP wave shifts,
// so a step has no LDS traffic and no barrier.  Cost and aggregate rows of the next B steps
// are loaded while the current B steps run.
//
// Recurrence (sgm.cu:75-98):
//   L(p,d) = C'(p,d) + min(L(q,d), L(q,d-1)+P1, L(q,d+1)+P1, m+P2) - m,  m = min_d L(q,d)
//   with P1, P2 scaled by q1, q2 where |I(p)-I(q)| < tau_so.
// A pixel is invalid if any plane's cost is < 0: there C' = 0 and the aggregate is set to 0
// (sgm.cu:33-38, 67-71, 101-104).  Every L is >= 0, so mins run on the unsigned bit patterns.
#include <algorithm>
#include <cmath>
#include <string>

#include "common.hpp"

#ifndef PM_SGM_PREFETCH
#define PM_SGM_PREFETCH 16
#endif

namespace pm {


__device__ __forceinline__ void scanline_start(const SgmArgs& a, int i, int& x, int& y) {
  if (a.dx != 0) {
    const int xs = a.dx > 0 ? 0 : a.W - 1;
    if (i < a.H) { x = xs; y = i; return; }
    const int j = i - a.H;
    x = a.dx > 0 ? 1 + j : j;
    y = a.dy > 0 ? 0 : a.H - 1;
    return;
  }
  x = i;
  y = a.dy > 0 ? 0 : a.H - 1;
}

// One wave per scanline.  All scanline geometry is wave-uniform (scalar registers); every
// load address is clamped in bounds (past-the-end steps re-read the last pixel, lanes past D
// re-read plane D-1 of the same pixel), so the prefetch has no branches and the B loads of
// the next block stay in flight while the current block computes.
template <int NV, typename CT, bool FIRST, bool FULL, int B>
__device__ __forceinline__ void sgm_scanline(const SgmArgs& a, int line,
                                             const CT* __restrict__ cost,
                                             float* __restrict__ agg) {
  const int lane = threadIdx.x & 63;
  int x, y;
  scanline_start(a, line, x, y);
  int n = 1 << 30;
  if (a.dx > 0) n = min(n, a.W - x);
  if (a.dx < 0) n = min(n, x + 1);
  if (a.dy > 0) n = min(n, a.H - y);
  if (a.dy < 0) n = min(n, y + 1);
  const int D = a.D;
  const int pstep = a.dy * a.W + a.dx;
  const int p0 = y * a.W + x;
  int off[NV];
  bool act[NV];
#pragma unroll
  for (int j = 0; j < NV; ++j) {
    const int d = lane * NV + j;
    act[j] = FULL || d < D;
    off[j] = FULL ? d : min(d, D - 1);
  }
  const float INF = __uint_as_float(0x7F800000u);

  // two register buffers of B steps each, used ping-pong: while one block computes, the
  // loads of the next are in flight (addresses past the end re-read the last pixel)
  float c0[B][NV], g0[B][NV], c1[B][NV], g1[B][NV];
  uint32_t r0, r1;
  auto load_block = [&](int base, float (&c)[B][NV], float (&g)[B][NV], uint32_t& r) {
    // lane i fetches the reference intensity of step base+i; issued first so that reading it
    // does not wait for the block's cost and aggregate loads
    r = a.ref[p0 + min(base + lane, n - 1) * pstep];
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int i = 0; i < B; ++i) {
      const long p = (long)(p0 + min(base + i, n - 1) * pstep) * D;
#pragma unroll
      for (int j = 0; j < NV; ++j) {
        c[i][j] = load_cost<CT>(cost + p + off[j]);
        if (!FIRST) g[i][j] = agg[p + off[j]];
      }
    }
  };

  float L[NV];
  uint32_t iq;        // intensity of the last pixel of the previous block
  uint64_t smooth;    // bit i: |I(step i) - I(step i-1)| < tau_so for the current block
  // gradient test of a whole block at once (sgm.cu:75-78): lane i compares step i with step
  // i-1 (lane 0 with the carried intensity) and a ballot packs the answers
  auto smooth_mask = [&](uint32_t r) {
    const int prev = (int)dpp_u32<0x138>(iq, r);   // wave_shr:1, lane 0 takes iq
    const int d1 = abs((int)r - prev);
    iq = __builtin_amdgcn_readlane(r, B - 1);
    return (uint64_t)__ballot(d1 < a.tau_i);
  };
  // one step k >= 1 from slot i of a block buffer
  auto step = [&](int k, int i, const float (&c)[B][NV], const float (&g)[B][NV]) {
    const long p = (long)(p0 + k * pstep) * D;
    bool neg = false;
#pragma unroll
    for (int j = 0; j < NV; ++j) neg = neg || c[i][j] < 0.0f;
    const bool invalid = __ballot(neg) != 0;
    // m = min_d L(q, d) (sgm.cu:55-66)
    float lm = L[0];
#pragma unroll
    for (int j = 1; j < NV; ++j) lm = umin(lm, L[j]);
    const float m = wave_min_nonneg(lm);
    // gradient-adaptive penalties (sgm.cu:78-82), scalar
    const bool sm = (smooth >> i) & 1ull;
    const float P1 = sm ? a.P1s : a.P1n;
    const float P2 = sm ? a.P2s : a.P2n;
    const float lo = wave_shr1(L[NV - 1], INF);   // L(q, d0-1)
    const float hi = wave_shl1(L[0], INF);        // L(q, d0+NV)
    const float mP2 = m + P2;
    float Ln[NV];
#pragma unroll
    for (int j = 0; j < NV; ++j) {
      const float left = j > 0 ? L[j - 1] : lo;
      const float right = j + 1 < NV ? L[j + 1] : hi;
      float val = umin(L[j], mP2);                 // sgm.cu:85
      val = umin(val, umin(left, right) + P1);     // sgm.cu:86-93
      const float cc = invalid ? kCostInf : c[i][j];
      Ln[j] = act[j] ? (cc + val) - m : INF;       // sgm.cu:98
      const float out = invalid ? kCostInf : (FIRST ? Ln[j] : g[i][j] + Ln[j]);
      if (act[j]) agg[p + off[j]] = out;           // sgm.cu:101-108
    }
#pragma unroll
    for (int j = 0; j < NV; ++j) L[j] = Ln[j];
  };
  // a block whose B steps all exist: no exits, so the loads of the other buffer stay in flight
  auto run_full = [&](int base, const float (&c)[B][NV], const float (&g)[B][NV], uint32_t r) {
    smooth = smooth_mask(r);
#pragma unroll
    for (int i = 0; i < B; ++i) step(base + i, i, c, g);
  };
  auto run_checked = [&](int base, const float (&c)[B][NV], const float (&g)[B][NV],
                         uint32_t r) {
    smooth = smooth_mask(r);
#pragma unroll
    for (int i = 0; i < B; ++i) {
      if (base + i >= n) break;
      step(base + i, i, c, g);
    }
  };

  // step 0: L = C' (sgm.cu:21-43)
  {
    const long p = (long)p0 * D;
    float c[NV], g[NV];
#pragma unroll
    for (int j = 0; j < NV; ++j) {
      c[j] = load_cost<CT>(cost + p + off[j]);
      g[j] = FIRST ? 0.0f : agg[p + off[j]];
    }
    iq = a.ref[p0];   // intensity of step 0
    bool neg = false;
#pragma unroll
    for (int j = 0; j < NV; ++j) neg = neg || c[j] < 0.0f;
    const bool invalid = __ballot(neg) != 0;
#pragma unroll
    for (int j = 0; j < NV; ++j) {
      L[j] = act[j] ? (invalid ? kCostInf : c[j]) : INF;
      const float out = invalid ? kCostInf : (FIRST ? L[j] : g[j] + L[j]);
      if (act[j]) agg[p + off[j]] = out;
    }
  }
  int base = 1;
  load_block(base, c0, g0, r0);
  while (base + 2 * B <= n) {
    load_block(base + B, c1, g1, r1);
    run_full(base, c0, g0, r0);
    load_block(base + 2 * B, c0, g0, r0);
    run_full(base + B, c1, g1, r1);
    base += 2 * B;
  }
  run_checked(base, c0, g0, r0);
  if (base + B < n) {
    load_block(base + B, c1, g1, r1);
    run_checked(base + B, c1, g1, r1);
  }
}

template <int NV, typename CT, bool FIRST, bool FULL, int B>
__global__ __launch_bounds__(256) void sgm_path_kernel(const SgmArgs a,
                                                       const CT* __restrict__ cost,
                                                       float* __restrict__ agg) {
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int line = blockIdx.x * 4 + wave;
  if (line >= a.nlines) return;
  sgm_scanline<NV, CT, FIRST, FULL, B>(a, line, cost, agg);
}

// Batched form of one path direction: the scanlines of every keyframe of a class (same NV,
// FULL) form one work list of 4-scanline items; keyframe kf[j] owns items first[j] ..
// first[j+1]-1.  Persistent blocks stride through the list.
template <int NV, typename CT, bool FIRST, bool FULL, int B>
__global__ __launch_bounds__(256) void sgm_batch_kernel(const KfDesc* __restrict__ kfs,
                                                        const int* __restrict__ kf,
                                                        const int* __restrict__ first, int n,
                                                        int total, int dx, int dy) {
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  for (int i = blockIdx.x; i < total; i += gridDim.x) {
    const int j = owner_of(first, n, i);
    const KfDesc& k = kfs[kf[j]];
    SgmArgs a;
    a.W = k.c.W; a.H = k.c.H; a.D = k.c.D; a.dx = dx; a.dy = dy;
    a.nlines = sgm_lines(a.W, a.H, dx, dy);
    a.ref = k.c.ref;
    a.P1n = k.P1n; a.P2n = k.P2n; a.P1s = k.P1s; a.P2s = k.P2s; a.tau_i = k.tau_i;
    const int line = (i - first[j]) * 4 + wave;
    if (line < a.nlines)
      sgm_scanline<NV, CT, FIRST, FULL, B>(a, line, static_cast<const CT*>(k.cost), k.agg);
  }
}

template <int NV, typename CT, bool FULL, int B>
void sgm_batch_path_t(const KfDesc* kfs, const int* kf, const int* first, int n, int total,
                      int grid, bool first_path, int dx, int dy, hipStream_t st) {
  if (first_path)
    sgm_batch_kernel<NV, CT, true, FULL, B><<<grid, 256, 0, st>>>(kfs, kf, first, n, total, dx, dy);
  else
    sgm_batch_kernel<NV, CT, false, FULL, B><<<grid, 256, 0, st>>>(kfs, kf, first, n, total, dx, dy);
}

template <typename CT, int B>
bool sgm_batch_path_b(int nv, bool full, const KfDesc* kfs, const int* kf, const int* first,
                      int n, int total, int grid, bool first_path, int dx, int dy,
                      hipStream_t st) {
  if (nv == 1 && full) sgm_batch_path_t<1, CT, true, B>(kfs, kf, first, n, total, grid, first_path, dx, dy, st);
  else if (nv == 1) sgm_batch_path_t<1, CT, false, B>(kfs, kf, first, n, total, grid, first_path, dx, dy, st);
  else if (nv == 2 && full) sgm_batch_path_t<2, CT, true, B>(kfs, kf, first, n, total, grid, first_path, dx, dy, st);
  else if (nv == 2) sgm_batch_path_t<2, CT, false, B>(kfs, kf, first, n, total, grid, first_path, dx, dy, st);
  else if (nv == 4) sgm_batch_path_t<4, CT, false, B>(kfs, kf, first, n, total, grid, first_path, dx, dy, st);
  else return false;
  return true;
}

template <typename CT>
bool sgm_batch_path_c(int prefetch, int nv, bool full, const KfDesc* kfs, const int* kf,
                      const int* first, int n, int total, int grid, bool first_path, int dx,
                      int dy, hipStream_t st) {
  if (prefetch == 8)
    return sgm_batch_path_b<CT, 8>(nv, full, kfs, kf, first, n, total, grid, first_path, dx, dy, st);
  if (prefetch == 16)
    return sgm_batch_path_b<CT, 16>(nv, full, kfs, kf, first, n, total, grid, first_path, dx, dy, st);
  return false;
}

bool launch_sgm_batch_path(pm_cost_type ct, int prefetch, int nv, bool full, const KfDesc* kfs,
                           const int* kf, const int* first, int n, int total, int grid,
                           bool first_path, int dx, int dy, hipStream_t st) {
  switch (ct) {
    case PM_COST_F32:
      return sgm_batch_path_c<float>(prefetch, nv, full, kfs, kf, first, n, total, grid, first_path, dx, dy, st);
    case PM_COST_F16:
      return sgm_batch_path_c<__half>(prefetch, nv, full, kfs, kf, first, n, total, grid, first_path, dx, dy, st);
    case PM_COST_BF16:
      return sgm_batch_path_c<__hip_bfloat16>(prefetch, nv, full, kfs, kf, first, n, total, grid, first_path, dx, dy, st);
    default:
      return false;
  }
}

template <int NV, typename CT, bool FULL>
pm_status launch_paths(const SgmArgs& base, int paths, const CT* cost, float* agg,
                       hipStream_t st) {
  constexpr int B = PM_SGM_PREFETCH;
  for (int path = 0; path < paths; ++path) {
    SgmArgs a = base;
    a.dx = kPathDx[path];
    a.dy = kPathDy[path];
    a.nlines = sgm_lines(a.W, a.H, a.dx, a.dy);
    const unsigned blocks = (unsigned)((a.nlines + 3) / 4);
    if (path == 0)
      sgm_path_kernel<NV, CT, true, FULL, B><<<blocks, 256, 0, st>>>(a, cost, agg);
    else
      sgm_path_kernel<NV, CT, false, FULL, B><<<blocks, 256, 0, st>>>(a, cost, agg);
    pm_status s = check_launch("sgm_path_kernel");
    if (s != PM_OK) return s;
  }
  return PM_OK;
}

template <typename CT>
pm_status launch_sgm(const SgmArgs& a, int paths, const CT* cost, float* agg, hipStream_t st) {
  if (a.D == 64) return launch_paths<1, CT, true>(a, paths, cost, agg, st);
  if (a.D == 128) return launch_paths<2, CT, true>(a, paths, cost, agg, st);
  if (a.D < 64) return launch_paths<1, CT, false>(a, paths, cost, agg, st);
  if (a.D < 128) return launch_paths<2, CT, false>(a, paths, cost, agg, st);
  return launch_paths<4, CT, false>(a, paths, cost, agg, st);
}

pm_status sgm_entry(const char* who, pm_shape shape, const uint8_t* ref, const void* cost,
                    pm_cost_type cost_type, pm_sgm_params params, float* agg, hipStream_t st) {
  const std::string w(who);
  if (!ref || !cost || !agg) return fail(PM_ERR_INVALID_ARGUMENT, w + ": null pointer");
  if (shape.width <= 0 || shape.height <= 0 || shape.width > 65535 || shape.height > 65535)
    return fail(PM_ERR_INVALID_ARGUMENT, w + ": bad image size");
  if (shape.planes <= 0 || shape.planes > PM_MAX_PLANES)
    return fail(PM_ERR_INVALID_ARGUMENT, w + ": planes must be in 1..256");
  if (params.paths != 4 && params.paths != 8)
    return fail(PM_ERR_INVALID_ARGUMENT, w + ": paths must be 4 or 8");
  const float kn[5] = {params.pi1, params.pi2, params.tau_so, params.sgm_q1, params.sgm_q2};
  for (float v : kn)
    if (!(std::isfinite(v) && v >= 0.0f))
      return fail(PM_ERR_INVALID_ARGUMENT,
                  w + ": pi1, pi2, tau_so, sgm_q1, sgm_q2 must be finite and >= 0");
  SgmArgs a{};
  a.W = shape.width; a.H = shape.height; a.D = shape.planes; a.ref = ref;
  a.P1n = params.pi1;
  a.P2n = params.pi2;
  a.P1s = params.pi1 * params.sgm_q1;   // the same float products as sgm.cu:80-81
  a.P2s = params.pi2 * params.sgm_q2;
  a.tau_i = (int)std::min(256.0f, std::ceil(params.tau_so));
  switch (cost_type) {
    case PM_COST_F32: return launch_sgm<float>(a, params.paths, (const float*)cost, agg, st);
    case PM_COST_F16: return launch_sgm<__half>(a, params.paths, (const __half*)cost, agg, st);
    case PM_COST_BF16:
      return launch_sgm<__hip_bfloat16>(a, params.paths, (const __hip_bfloat16*)cost, agg, st);
    default: return fail(PM_ERR_INVALID_ARGUMENT, w + ": unknown cost type");
  }
}

}  // namespace pm

using namespace pm;

extern "C" pm_status pm_sgm(pm_shape shape, const uint8_t* ref, const void* cost,
                            pm_cost_type cost_type, pm_sgm_params params, float* agg,
                            pm_stream stream) {
  clear_error();
  return sgm_entry("pm_sgm", shape, ref, cost, cost_type, params, agg, (hipStream_t)stream);
}

extern "C" pm_status pm_sgm_filter(pm_shape shape, const uint8_t* ref, const void* cost,
                                   pm_cost_type cost_type, pm_sgm_params params,
                                   float var_scale, float idep_min, float dep_sample,
                                   float* agg, float* depth, pm_stream stream) {
  clear_error();
  if (!depth) return fail(PM_ERR_INVALID_ARGUMENT, "pm_sgm_filter: null depth");
  const pm_status s = sgm_entry("pm_sgm_filter", shape, ref, cost, cost_type, params, agg,
                                (hipStream_t)stream);
  if (s != PM_OK) return s;
  return pm_filter_cost(shape, agg, var_scale, idep_min, dep_sample, depth, stream);
}

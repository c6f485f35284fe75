// Differentiable form of the depth pipeline: a soft-argmin depth readout over the SGM
// aggregate, a masked L1 loss, and the hand-written adjoint of the whole chain with respect to
// the reference intensities and the per-source plane-sweep homographies (H, t).  The host
// maps dL/d(H, t) to the se3 pose of each measurement frame (pm_se3_pose_grad).
//
// There is no counterpart in the reference; the forward stages are the reference's
// (calc_cost.cu:11-210, sgm.cu:5-113) run on float reference intensities.  Readout:
//   z_d = -agg(p,d) / (cost_scale * T),  sigma = softmax_d(z),  depth_hat(p) = sum_d sigma_d / idep_d
//   loss = mean over M of |depth_hat - gt|,  M = {pixels whose cost is valid and gt > 0}.
//
// Adjoint, in reverse order of the forward:
//   1. readout: dL/dagg(p,d) = g(p) sigma_d (1/idep_d - depth_hat) (-1 / (cost_scale T)),
//      g(p) = sign(depth_hat - gt) / |M| on M, written over the aggregate in place;
//   2. SGM: each path is swept backwards.  L(q,d) = C'(q,d) + min(L(p,d), L(p,d-1)+P1,
//      L(p,d+1)+P1, m_p+P2) - m_p routes dL/dL(q,d) to the argument that won the min (ties:
//      L(p,d), then L(p,d-1), then L(p,d+1), then m_p), and -m_p and the m_p+P2 branch to
//      argmin_d L(p,d) (lowest index).  dL/dC(q,d) accumulates over paths where q is valid;
//   3. cost: each valid (pixel, plane, source) term |I(r) - B(pi(H_d x))| / (9 n) passes
//      -sign(I - B) / (9 n) dL/dC to the sample and +sign(I - B) / (9 n) dL/dC to I(r); the
//      sample's derivative goes through the bilinear cell, the perspective division and the
//      (garbled, calc_cost.cu:49-51, 69-71) warp expressions into (H, t).
// The per-path L volumes are stored by the forward (8 volumes), or recomputed one path at a
// time in the backward ("recompute" mode: the forward keeps only the cost, the source counts
// and the readout).  Every reduction has a fixed order, so results are deterministic.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "common.hpp"
#include "cost_tile.hpp"
#include "sgm_scan.hpp"

namespace pm {

namespace {

constexpr int kCostTy = 4, kCostDc = 8;   // forward cost tile
static_assert(cost_tile_max_sources(kCostTy) >= PM_MAX_SOURCES, "tile too short for the sources");
constexpr int kAdjTy = 2, kAdjDc = 8;     // adjoint cost tile
constexpr int kReadoutBlocks = 1024;

__device__ __forceinline__ float sgnf(float x) {
  return x > 0.0f ? 1.0f : (x < 0.0f ? -1.0f : 0.0f);
}

// Sum over the 64 lanes in a fixed order, wave-uniform: DPP butterflies inside each row of 16
// lanes, then the four row sums (read from lanes 0, 16, 32, 48) added pairwise.
template <int CTRL>
__device__ __forceinline__ float dpp_f32(float v) {
  return __uint_as_float(dpp_u32<CTRL>(__float_as_uint(v), __float_as_uint(v)));
}
__device__ __forceinline__ float wave_sum(float v) {
  v += dpp_f32<0xB1>(v);    // quad_perm [1,0,3,2]
  v += dpp_f32<0x4E>(v);    // quad_perm [2,3,0,1]
  v += dpp_f32<0x124>(v);   // row_ror:4
  v += dpp_f32<0x128>(v);   // row_ror:8
  const uint32_t u = __float_as_uint(v);
  const float a = __uint_as_float(__builtin_amdgcn_readlane(u, 0));
  const float b = __uint_as_float(__builtin_amdgcn_readlane(u, 16));
  const float c = __uint_as_float(__builtin_amdgcn_readlane(u, 32));
  const float d = __uint_as_float(__builtin_amdgcn_readlane(u, 48));
  return (a + b) + (c + d);
}

// ---- forward ------------------------------------------------------------------------------

__global__ __launch_bounds__(256) void cost_diff_kernel(const CostArgs a, float* __restrict__ cost,
                                                        uint8_t* __restrict__ cnt) {
  __shared__ float stage[4][kCostTy * kCostDc * 64];
  __shared__ uint8_t cstage[4][kCostTy * kCostDc * 64];
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  cost_tile<float, kCostTy, kCostDc, float, true>(a, xcd_block(blockIdx.x, gridDim.x), cost,
                                                  stage[wave], cnt, cstage[wave]);
}

// One path of the float-reference SGM: AGG accumulates into agg, LOUT stores the path's L.
template <int NV, bool FIRST, bool FULL, bool AGG, bool LOUT>
__global__ __launch_bounds__(256) void sgm_diff_kernel(const SgmArgs a, const float* __restrict__ cost,
                                                       float* __restrict__ agg) {
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int line = blockIdx.x * 4 + wave;
  if (line >= a.nlines) return;
  sgm_scanline<NV, float, FIRST, FULL, 16 / NV, float, AGG, LOUT>(a, line, cost, agg);
}

// Every path of the float-reference SGM in one launch (store mode).  Global line g counts the
// lines of path 0, then path 1, ...; path r writes its L (unmasked: the backward routes
// through it) to lbuf + r V, so no two waves touch the same output and all paths run at once.
// agg_sum_kernel then forms the aggregate.
template <int NV, bool FULL>
__global__ __launch_bounds__(256) void sgm_diff_paths_kernel(const SgmArgs base, int paths,
                                                             const float* __restrict__ cost,
                                                             float* __restrict__ lbuf) {
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  int g = blockIdx.x * 4 + wave, path = 0, n = 0;
  for (; path < paths; ++path, g -= n) {
    n = sgm_lines(base.W, base.H, path_dx(path), path_dy(path));
    if (g < n) break;
  }
  if (path == paths) return;
  SgmArgs a = base;
  a.dx = path_dx(path);
  a.dy = path_dy(path);
  a.nlines = n;
  a.lout = lbuf + (size_t)path * ((size_t)base.W * base.H * base.D);
  sgm_scanline<NV, float, true, FULL, 16 / NV, float, false, true>(a, g, cost, nullptr);
}

// agg = the paths' L volumes summed in path order, 0 at a pixel with a negative cost on any
// plane: what the path-after-path accumulation (sgm.cu:101-108, sgm_diff_kernel with AGG)
// leaves, in the same order.  16 lanes per pixel, lane k holding planes k, k+16, ...
template <int PPL>
__global__ __launch_bounds__(256) void agg_sum_kernel(long N, int D, int paths,
                                                      const float* __restrict__ cost,
                                                      const float* __restrict__ lbuf,
                                                      float* __restrict__ agg) {
  const long V = N * D;
  const long p = ((long)blockIdx.x * 256 + threadIdx.x) >> 4;
  const int k = threadIdx.x & 15;
  if (p >= N) return;   // the whole 16-lane group
  float sum[PPL];
  bool bad = false;
#pragma unroll
  for (int j = 0; j < PPL; ++j) {
    const int d = k + 16 * j;
    sum[j] = 0.0f;
    if (d < D) {
      const long e = p * D + d;
      bad = bad || cost[e] < 0.0f;
      float v = lbuf[e];
      for (int r = 1; r < paths; ++r) v += lbuf[r * V + e];
      sum[j] = v;
    }
  }
  float b = bad ? 1.0f : 0.0f;
#pragma unroll
  for (int off = 8; off >= 1; off >>= 1) b = fmaxf(b, __shfl_xor(b, off, 16));
#pragma unroll
  for (int j = 0; j < PPL; ++j) {
    const int d = k + 16 * j;
    if (d < D) agg[p * D + d] = b > 0.0f ? kCostInf : sum[j];
  }
}

// The same with D a multiple of 4: lane k of a pixel's 16 sums the 4-plane chunks k, k+16, ...
// with 16-byte loads, so a group reads whole 256-byte runs of the pixel's planes (per-plane
// lanes read each 128-byte line in two halves at different times: 1.7x the bytes).
template <int CPL>
__global__ __launch_bounds__(256) void agg_sum4_kernel(long N, int D, int paths,
                                                       const float* __restrict__ cost,
                                                       const float* __restrict__ lbuf,
                                                       float* __restrict__ agg) {
  const long V = N * D;
  const int D4 = D / 4;
  const long p = ((long)blockIdx.x * 256 + threadIdx.x) >> 4;
  const int k = threadIdx.x & 15;
  if (p >= N) return;   // the whole 16-lane group
  float4 sum[CPL];
  bool bad = false;
#pragma unroll
  for (int j = 0; j < CPL; ++j) {
    const int c = k + 16 * j;
    sum[j] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    if (c < D4) {
      const long e = p * D + 4 * c;
      const float4 cv = *reinterpret_cast<const float4*>(cost + e);
      bad = bad || cv.x < 0.0f || cv.y < 0.0f || cv.z < 0.0f || cv.w < 0.0f;
      float4 v = *reinterpret_cast<const float4*>(lbuf + e);
      for (int r = 1; r < paths; ++r) {
        const float4 w = *reinterpret_cast<const float4*>(lbuf + r * V + e);
        v.x += w.x; v.y += w.y; v.z += w.z; v.w += w.w;
      }
      sum[j] = v;
    }
  }
  float b = bad ? 1.0f : 0.0f;
#pragma unroll
  for (int off = 8; off >= 1; off >>= 1) b = fmaxf(b, __shfl_xor(b, off, 16));
#pragma unroll
  for (int j = 0; j < CPL; ++j) {
    const int c = k + 16 * j;
    if (c < D4)
      *reinterpret_cast<float4*>(agg + p * D + 4 * c) =
          b > 0.0f ? make_float4(kCostInf, kCostInf, kCostInf, kCostInf) : sum[j];
  }
}

// dL/dC = the paths' adjoint volumes summed in path order (each already 0 at invalid pixels):
// what the path-after-path accumulation of sgm_adjoint_kernel leaves, in the same order.
// Element e4 of the float4 view; V is a multiple of 4 when vec4 is set.
__global__ __launch_bounds__(256) void grad_cost_sum_kernel(long V, int paths, bool vec4,
                                                            const float* __restrict__ gbuf,
                                                            float* __restrict__ gC) {
  const long stride = (long)gridDim.x * 256;
  if (vec4) {
    for (long e = ((long)blockIdx.x * 256 + threadIdx.x) * 4; e < V; e += stride * 4) {
      float4 v = *reinterpret_cast<const float4*>(gbuf + e);
      for (int r = 1; r < paths; ++r) {
        const float4 w = *reinterpret_cast<const float4*>(gbuf + r * V + e);
        v.x += w.x; v.y += w.y; v.z += w.z; v.w += w.w;
      }
      *reinterpret_cast<float4*>(gC + e) = v;
    }
    return;
  }
  for (long e = (long)blockIdx.x * 256 + threadIdx.x; e < V; e += stride) {
    float v = gbuf[e];
    for (int r = 1; r < paths; ++r) v += gbuf[r * V + e];
    gC[e] = v;
  }
}

// 16 lanes per pixel, lane k holding planes k, k+16, ...; four pixels per wave.
struct ReadoutArgs {
  long N;
  int D;
  float idep_min, dep_sample, inv_ts;
  const float* agg;
  const uint8_t* cnt;
  const float* gt;
  float* depth;
  uint8_t* pvalid;
  double* part;        // [kReadoutBlocks][2]: sum |depth_hat - gt| over M, |M|
  const double* sums;  // the reduced pair (backward)
  float grad_loss;
};

// Group-of-16 reductions (butterflies stay inside the 16-lane group); every lane of the group
// gets the group's first lane's result, so the value does not depend on the lane.
__device__ __forceinline__ float grp_min(float v) {
#pragma unroll
  for (int off = 8; off >= 1; off >>= 1) v = fminf(v, __shfl_xor(v, off, 16));
  return __shfl(v, 0, 16);
}
__device__ __forceinline__ float grp_sum(float v) {
#pragma unroll
  for (int off = 8; off >= 1; off >>= 1) v += __shfl_xor(v, off, 16);
  return __shfl(v, 0, 16);
}

// Softmax statistics of one pixel; e[j] = exp(-(a_j - amin) inv_ts) for the lane's planes.
template <int PPL>
__device__ __forceinline__ void pixel_softmax(const ReadoutArgs& r, long p, int k, float (&a)[PPL],
                                              float (&e)[PPL], float (&dep)[PPL], float& Z,
                                              float& num, bool& valid) {
  const float INF = __uint_as_float(0x7F800000u);
  float amin = INF;
  bool bad = false;
#pragma unroll
  for (int j = 0; j < PPL; ++j) {
    const int d = k + 16 * j;
    const bool in = d < r.D;
    a[j] = in ? r.agg[p * r.D + d] : INF;
    if (r.cnt) bad = bad || (in && r.cnt[p * r.D + d] == 0);
    dep[j] = in ? 1.0f / (r.idep_min + (float)d * r.dep_sample) : 0.0f;
    amin = fminf(amin, a[j]);
  }
  amin = grp_min(amin);
  float z = 0.0f, n = 0.0f;
#pragma unroll
  for (int j = 0; j < PPL; ++j) {
    e[j] = (k + 16 * j < r.D) ? expf(-(a[j] - amin) * r.inv_ts) : 0.0f;
    z += e[j];
    n = fmaf(e[j], dep[j], n);
  }
  Z = grp_sum(z);
  num = grp_sum(n);
  // a pixel is valid when every plane has a source (group AND through a 16-lane OR)
  float b = bad ? 1.0f : 0.0f;
#pragma unroll
  for (int off = 8; off >= 1; off >>= 1) b = fmaxf(b, __shfl_xor(b, off, 16));
  valid = b == 0.0f;
}

template <int PPL>
__global__ __launch_bounds__(256) void readout_kernel(const ReadoutArgs r) {
  __shared__ double red[2][4];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int grp = lane >> 4, k = lane & 15;
  double lsum = 0.0, lcnt = 0.0;
  const long groups = (r.N + 3) / 4;   // 4 pixels per wave-iteration
  for (long g = (long)blockIdx.x * 4 + wave; g < groups; g += (long)gridDim.x * 4) {
    const long p = g * 4 + grp;
    if (p >= r.N) continue;   // whole 16-lane group skips together
    float a[PPL], e[PPL], dep[PPL], Z, num;
    bool valid;
    pixel_softmax<PPL>(r, p, k, a, e, dep, Z, num, valid);
    const float dh = num / Z;
    const float gt = r.gt[p];
    const bool m = valid && gt > 0.0f && gt < __uint_as_float(0x7F800000u);
    if (k == 0) {
      r.depth[p] = dh;
      r.pvalid[p] = valid ? 1 : 0;
      if (m) { lsum += fabs((double)dh - (double)gt); lcnt += 1.0; }
    }
  }
  // fixed-order block reduction: lanes, then waves
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) {
    lsum += __shfl_xor(lsum, off, 64);
    lcnt += __shfl_xor(lcnt, off, 64);
  }
  if (lane == 0) { red[0][wave] = lsum; red[1][wave] = lcnt; }
  __syncthreads();
  if (threadIdx.x == 0) {
    r.part[blockIdx.x * 2 + 0] = ((red[0][0] + red[0][1]) + red[0][2]) + red[0][3];
    r.part[blockIdx.x * 2 + 1] = ((red[1][0] + red[1][1]) + red[1][2]) + red[1][3];
  }
}

// Sums n rows of values (row stride `stride`) column-wise into out[] in FP64, one block per
// column (launched with `width` blocks), in a fixed order.
template <typename T>
__global__ __launch_bounds__(256) void reduce_columns_kernel(const T* __restrict__ part, int n,
                                                            double* __restrict__ out,
                                                            int stride) {
  __shared__ double red[256];
  const int c = blockIdx.x;
  double s = 0.0;
  for (int i = threadIdx.x; i < n; i += 256) s += (double)part[(long)i * stride + c];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int h = 128; h >= 1; h >>= 1) {
    if ((int)threadIdx.x < h) red[threadIdx.x] += red[threadIdx.x + h];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[c] = red[0];
}

__global__ void finish_loss_kernel(double* sums, double* loss) {
  if (threadIdx.x == 0 && blockIdx.x == 0) loss[0] = sums[1] > 0.0 ? sums[0] / sums[1] : 0.0;
}

// ---- backward -----------------------------------------------------------------------------

// dL/dagg over the aggregate in place (readout adjoint).
template <int PPL>
__global__ __launch_bounds__(256) void readout_adjoint_kernel(ReadoutArgs r, float* __restrict__ gagg) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int grp = lane >> 4, k = lane & 15;
  const double cntM = r.sums[1];
  const float scale = cntM > 0.0 ? (float)((double)r.grad_loss / cntM) : 0.0f;
  const long groups = (r.N + 3) / 4;
  for (long g = (long)blockIdx.x * 4 + wave; g < groups; g += (long)gridDim.x * 4) {
    const long p = g * 4 + grp;
    if (p >= r.N) continue;
    float a[PPL], e[PPL], dep[PPL], Z, num;
    bool valid;
    ReadoutArgs rr = r;
    rr.cnt = nullptr;   // validity comes from pvalid
    pixel_softmax<PPL>(rr, p, k, a, e, dep, Z, num, valid);
    const float dh = r.depth[p];
    const float gt = r.gt[p];
    const bool m = r.pvalid[p] && gt > 0.0f && gt < __uint_as_float(0x7F800000u);
    const float gd = m ? sgnf(dh - gt) * scale : 0.0f;
    const float f = -gd * r.inv_ts / Z;
#pragma unroll
    for (int j = 0; j < PPL; ++j) {
      const int d = k + 16 * j;
      if (d < r.D) gagg[p * r.D + d] = (e[j] * f) * (dep[j] - dh);
    }
  }
}

// Reverse sweep of one SGM path.  One wave per scanline, lane l holding planes
// l*NV .. l*NV+NV-1 as in the forward.  carry = dL/dL(q, .) flowing in from q's successor.
// Loads of the next B steps (walking backwards) are issued before the current B steps run.
struct SgmAdjArgs {
  SgmArgs s;
  const float* L;       // this path's L volume
  const float* gagg;    // dL/dagg
  const uint8_t* pvalid;
  float* gC;            // dL/dC, accumulated over paths
};

// FIRST: the adjoint is written to A.gC (0 at invalid pixels), else added to it.  With FIRST,
// A.gC may be A.L itself: step k writes pixel p_k after reading L at p_k's predecessor, and
// the loads ahead (lower k) read predecessors below the block being written, so every L value
// is read before it is overwritten.
template <int NV, bool FIRST, bool FULL, int B>
__device__ __forceinline__ void sgm_adjoint_scanline(const SgmAdjArgs& A, int line) {
  const SgmArgs& a = A.s;
  const int lane = threadIdx.x & 63;
  const float* __restrict__ ref = static_cast<const float*>(a.ref);
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

  float ga0[B][NV], lp0[B][NV], gc0[B][NV], ga1[B][NV], lp1[B][NV], gc1[B][NV];
  uint64_t vm0, vm1, sm0, sm1;   // bit i: step base-i valid / smooth
  // slot i of a block holds step base - i (steps below 0 re-read step 0)
  auto load_block = [&](int base, float (&ga)[B][NV], float (&lp)[B][NV], float (&gc)[B][NV],
                        uint64_t& vm, uint64_t& sm) {
    const int ks = max(base - lane, 0);
    const int q = p0 + ks * pstep;
    const float iq = ref[q];
    const float ip = ref[p0 + max(ks - 1, 0) * pstep];
    const bool v = A.pvalid[q] != 0;
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int i = 0; i < B; ++i) {
      const int k = max(base - i, 0);
      const long pq = (long)(p0 + k * pstep) * D;
      const long pp = (long)(p0 + max(k - 1, 0) * pstep) * D;
#pragma unroll
      for (int j = 0; j < NV; ++j) {
        ga[i][j] = A.gagg[pq + off[j]];
        lp[i][j] = A.L[pp + off[j]];
        if (!FIRST) gc[i][j] = A.gC[pq + off[j]];
      }
    }
    vm = __ballot(v);
    sm = __ballot(fabsf(iq - ip) < a.tau_f);
  };

  float carry[NV];
#pragma unroll
  for (int j = 0; j < NV; ++j) carry[j] = 0.0f;

  auto step = [&](int k, int i, const float (&ga)[B][NV], const float (&lp)[B][NV],
                  const float (&gc)[B][NV], uint64_t vm, uint64_t sm) {
    const long pq = (long)(p0 + k * pstep) * D;
    const bool qv = (vm >> i) & 1ull;
    float gL[NV];
#pragma unroll
    for (int j = 0; j < NV; ++j) {
      gL[j] = act[j] ? carry[j] + (qv ? ga[i][j] : 0.0f) : 0.0f;
      if (act[j]) {
        if (FIRST) A.gC[pq + off[j]] = qv ? gL[j] : 0.0f;
        else if (qv) A.gC[pq + off[j]] = gc[i][j] + gL[j];
      }
    }
    if (k == 0) return;
    // routing of L(q, d) = C'(q, d) + min(...) - m_p through the predecessor's values
    float Lp[NV];
#pragma unroll
    for (int j = 0; j < NV; ++j) Lp[j] = act[j] ? lp[i][j] : INF;
    float lm = Lp[0];
#pragma unroll
    for (int j = 1; j < NV; ++j) lm = umin(lm, Lp[j]);
    const float m = wave_min_nonneg(lm);
    const bool smooth = (sm >> i) & 1ull;
    const float P1 = smooth ? a.P1s : a.P1n;
    const float P2 = smooth ? a.P2s : a.P2n;
    const float lo = wave_shr1(Lp[NV - 1], INF);
    const float hi = wave_shl1(Lp[0], INF);
    const float c3 = m + P2;
    float x0[NV], xl[NV], xr[NV];
    float keep = 0.0f;
#pragma unroll
    for (int j = 0; j < NV; ++j) {
      const float left = j > 0 ? Lp[j - 1] : lo;
      const float right = j + 1 < NV ? Lp[j + 1] : hi;
      const float nb = umin(left, right) + P1;
      const bool ch0 = (Lp[j] <= nb) && (Lp[j] <= c3);
      const bool ch3 = !ch0 && (c3 < nb);
      const bool chl = !ch0 && !ch3 && (left <= right);
      x0[j] = ch0 ? gL[j] : 0.0f;
      xl[j] = chl ? gL[j] : 0.0f;                        // to L(p, d-1)
      xr[j] = (!ch0 && !ch3 && !chl) ? gL[j] : 0.0f;     // to L(p, d+1)
      keep += ch3 ? 0.0f : gL[j];
    }
    const float s3 = -wave_sum(keep);   // dL/dm_p: sum over the m_p+P2 branch minus all
    const float from_r = wave_shl1z(xl[0]);        // lane+1's plane 0 sends to our last plane
    const float from_l = wave_shr1z(xr[NV - 1]);   // lane-1's last plane sends to our plane 0
    // argmin_d L(p, d), lowest index
    int jm = NV;
#pragma unroll
    for (int j = NV - 1; j >= 0; --j)
      if (Lp[j] == m) jm = j;
    const uint64_t hit = __ballot(jm < NV);   // non-empty unless L holds NaN
    const int lstar = hit ? __builtin_ctzll(hit) : 64;
    const int jstar = __builtin_amdgcn_readlane(jm, lstar & 63);
#pragma unroll
    for (int j = 0; j < NV; ++j) {
      float c = x0[j];
      c += j + 1 < NV ? xl[j + 1] : from_r;
      c += j > 0 ? xr[j - 1] : from_l;
      if (lane == lstar && j == jstar) c += s3;
      carry[j] = act[j] ? c : 0.0f;
    }
  };

  int base = n - 1;
  load_block(base, ga0, lp0, gc0, vm0, sm0);
  while (base >= 0) {
    const int nb = base - B;
    if (nb >= 0) load_block(nb, ga1, lp1, gc1, vm1, sm1);
#pragma unroll
    for (int i = 0; i < B; ++i)
      if (base - i >= 0) step(base - i, i, ga0, lp0, gc0, vm0, sm0);
    base = nb;
    if (base < 0) break;
    const int nb2 = base - B;
    if (nb2 >= 0) load_block(nb2, ga0, lp0, gc0, vm0, sm0);
#pragma unroll
    for (int i = 0; i < B; ++i)
      if (base - i >= 0) step(base - i, i, ga1, lp1, gc1, vm1, sm1);
    base = nb2;
  }
}

// One path's reverse sweep, accumulated into A.gC (recompute mode: path after path).
template <int NV, bool FIRST, bool FULL, int B>
__global__ __launch_bounds__(256) void sgm_adjoint_kernel(const SgmAdjArgs A) {
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int line = blockIdx.x * 4 + wave;
  if (line >= A.s.nlines) return;
  sgm_adjoint_scanline<NV, FIRST, FULL, B>(A, line);
}

// Every path's reverse sweep in one launch (store mode): path r reads its L from
// base.L + r V and writes its adjoint over it (grad_cost_sum_kernel then sums them).
template <int NV, bool FULL, int B>
__global__ __launch_bounds__(256) void sgm_adjoint_paths_kernel(const SgmAdjArgs base,
                                                                int paths) {
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  int g = blockIdx.x * 4 + wave, path = 0, n = 0;
  for (; path < paths; ++path, g -= n) {
    n = sgm_lines(base.s.W, base.s.H, path_dx(path), path_dy(path));
    if (g < n) break;
  }
  if (path == paths) return;
  SgmAdjArgs A = base;
  A.s.dx = path_dx(path);
  A.s.dy = path_dy(path);
  A.s.nlines = n;
  float* vol = const_cast<float*>(base.L) + (size_t)path * ((size_t)base.s.W * base.s.H * base.s.D);
  A.L = vol;
  A.gC = vol;
  sgm_adjoint_scanline<NV, true, FULL, B>(A, g);
}

// Cost adjoint.  Tiles as in the forward (lanes = columns x0-1 .. x0+62, TY output rows, the
// planes split over the 4 waves), each tile taking the gradient of its own output pixels'
// terms.  For a plane d and source s it recomputes the centre samples of rows y0-1 .. y0+TY
// and the garbled samples of its output rows, the validity of each pixel (all nine samples
// valid), and the pixel weights w = dL/dC / (9 n) of the valid ones.  A centre sample at q is
// used by the A-terms of pixels q, q+(1,0), q+(1,1), q+(-1,-1), q+(1,-1) (against I(q)), by
// term 1 of q+(0,1) (against I(q+(0,2))) and term 2 of q-(0,1) (against I(q-(0,2))).
// Reference-intensity gradients of the tile's rows y0-1 .. y0+TY go to a per-tile partial,
// gathered per pixel afterwards; dL/d(H, t) sums go to a per-(tile, wave, source) partial.
struct CostAdjArgs {
  CostArgs c;
  const float* gC;
  const uint8_t* cnt;
  float* refpart;    // [tiles][TY + 6][64]
  float* posepart;   // [tiles * 4][S][12]
};

template <int TY, int DC>
__global__ __launch_bounds__(256) void cost_adjoint_kernel(const CostAdjArgs A) {
  __shared__ float wst[4][TY * DC * 64];
  __shared__ float gsum[4][(TY + 6) * 64];
  const CostArgs& a = A.c;
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int blk = xcd_block(blockIdx.x, gridDim.x);
  const int nstrips = (a.W + 61) / 62;
  const int strip = blk % nstrips;
  const int y0 = (blk / nstrips) * TY;
  const int W = a.W, H = a.H, D = a.D;
  const int x0 = strip * 62;
  const int u = x0 - 1 + lane;
  const int dper = (D + 3) / 4;
  const int dbeg = wave * dper, dend = min(D, dbeg + dper);
  const float Wm1 = (float)(W - 1), Hm1 = (float)(H - 1);
  const float INF = __uint_as_float(0x7F800000u);
  const float fu = (float)u;
  const float* refp = static_cast<const float*>(a.ref);
  float I[TY + 6], IR[TY + 1], GR[TY + 6], GRR[TY + 6];
#pragma unroll
  for (int k = 0; k < TY + 6; ++k) {
    I[k] = texel(refp, W, H, u, y0 - 3 + k);
    GR[k] = 0.0f;
    GRR[k] = 0.0f;
  }
#pragma unroll
  for (int k = 0; k <= TY; ++k) IR[k] = wave_shl1z(I[k + 2]);
  float* pp = A.posepart + ((size_t)blk * 4 + wave) * a.S * 12;
  if (dbeg >= dend && lane < 12)
    for (int s = 0; s < a.S; ++s) pp[s * 12 + lane] = 0.0f;

  for (int dc0 = dbeg; dc0 < dend; dc0 += DC) {
    const int dcn = min(DC, dend - dc0);
    // weights dL/dC / (9 n) of the tile's own pixels (lanes 1..62), 0 elsewhere
    for (int it = lane; it < TY * 64 * dcn; it += 64) {
      const int j = it % dcn, rest = it / dcn, ln = rest % 64, t = rest / 64;
      const int v = y0 + t, x = x0 - 1 + ln;
      float w = 0.0f;
      if (ln >= 1 && ln <= 62 && v < H && x < W) {
        const long e = ((long)v * W + x) * D + dc0 + j;
        const int n = A.cnt[e];
        w = n > 0 ? A.gC[e] / (9.0f * (float)n) : 0.0f;
      }
      wst[wave][(t * DC + j) * 64 + ln] = w;
    }
    __builtin_amdgcn_wave_barrier();
    for (int s = 0; s < a.S; ++s) {
      const float* hm = a.hm[s];
      const float h11 = hm[0], h12 = hm[1], h13 = hm[2], h21 = hm[3], h22 = hm[4],
                  h23 = hm[5], h31 = hm[6], h32 = hm[7], h33 = hm[8];
      const QuadTex q(a.quad[s], (uint32_t)W, (uint32_t)H);
      // per-lane sums: centre samples (X, Y, w and their products with the row), garbled
      // samples, translation
      float cX = 0, cXr = 0, cY = 0, cYr = 0, cW = 0, cWr = 0;
      float gS = 0, gSv = 0, g4X = 0, g4Y = 0, g4W = 0, g8X = 0, g8Y = 0, g8W = 0;
      float tX = 0, tY = 0, tW = 0;
      for (int d = dc0; d < dc0 + dcn; ++d) {
        const float idep = a.idep_min + (float)d * a.dep_sample;
        const float tx = hm[9] * idep, ty = hm[10] * idep, tz = hm[11] * idep;
        const float Xu = h11 * fu + h13 + tx;
        const float Yu = h21 * fu + h23 + ty;
        const float Zu = h31 * fu + h33 + tz;
        const float xu = h11 * fu + h13;
        const float g4x = xu + (h11 + tx), g4y = xu + (h21 + ty), g4z = xu + (h31 + tz);
        const float g8x = xu + (h11 - h12 + tx), g8y = xu + (h11 - h22 + ty),
                    g8z = xu + (h11 - h32 + tz);
        SampleGrad C[TY + 2], G4[TY], G8[TY];
#pragma unroll
        for (int k = 0; k < TY + 2; ++k) {
          const float fr = (float)(y0 - 1 + k);
          C[k] = warp_sample_grad(q, (uint32_t)W, Wm1, Hm1, fmaf(h12, fr, Xu), fmaf(h22, fr, Yu),
                                  fmaf(h32, fr, Zu));
        }
#pragma unroll
        for (int t = 0; t < TY; ++t) {
          const float fv = (float)(y0 + t);
          G4[t] = warp_sample_grad(q, (uint32_t)W, Wm1, Hm1, fmaf(h12, fv, g4x),
                                   fmaf(h12, fv, g4y), fmaf(h12, fv, g4z));
          G8[t] = warp_sample_grad(q, (uint32_t)W, Wm1, Hm1, fmaf(h12, fv, g8x),
                                   fmaf(h12, fv, g8y), fmaf(h12, fv, g8z));
        }
        // validity of each output pixel (the forward's nine-term sum is finite), weights
        float Wv[TY], WL[TY], WR[TY];
#pragma unroll
        for (int t = 0; t < TY; ++t) {
          const int k = t + 2;   // centre row index of row v + 1
          const float A0 = fabsf(I[k] - C[k - 2].s), A1 = fabsf(I[k + 1] - C[k - 1].s);
          const float A2 = fabsf(I[k + 2] - C[k].s);
          float T = A1 + wave_shr1z(A0);
          T += wave_shr1z(A1);
          T += wave_shl1z(A2);
          T += wave_shr1z(A2);
          T += fabsf(I[k + 2] - C[k - 2].s) + fabsf(I[k] - C[k].s);
          T += fabsf(IR[t + 1] - G4[t].s) + fabsf(IR[t] - G8[t].s);
          const float w = wst[wave][(t * DC + (d - dc0)) * 64 + lane];
          Wv[t] = T < INF ? w : 0.0f;
        }
#pragma unroll
        for (int t = 0; t < TY; ++t) {
          WL[t] = wave_shr1z(Wv[t]);
          WR[t] = wave_shl1z(Wv[t]);
        }
        auto wv = [&](int t) { return (t >= 0 && t < TY) ? Wv[t] : 0.0f; };
        auto wl = [&](int t) { return (t >= 0 && t < TY) ? WL[t] : 0.0f; };
        auto wr = [&](int t) { return (t >= 0 && t < TY) ? WR[t] : 0.0f; };
#pragma unroll
        for (int k = 0; k < TY + 2; ++k) {
          const float r = (float)(y0 - 1 + k);
          const float S = C[k].s;
          const float wA = wv(k - 1) + wr(k) + wr(k - 1) + wl(k - 2) + wr(k - 2);
          const float wUp = wv(k), wDn = wv(k - 2);
          const float sA = sgnf(I[k + 2] - S) * wA;
          const float sU = sgnf(I[k + 4] - S) * wUp;
          const float sD = sgnf(I[k] - S) * wDn;
          GR[k + 2] += sA;
          GR[k + 4] += sU;
          GR[k] += sD;
          const float g = -(sA + sU + sD);
          const float gx = g * C[k].a, gy = g * C[k].b, gw = g * C[k].c;
          cX += gx; cXr = fmaf(gx, r, cXr);
          cY += gy; cYr = fmaf(gy, r, cYr);
          cW += gw; cWr = fmaf(gw, r, cWr);
          tX = fmaf(gx, idep, tX); tY = fmaf(gy, idep, tY); tW = fmaf(gw, idep, tW);
        }
#pragma unroll
        for (int t = 0; t < TY; ++t) {
          const float v = (float)(y0 + t);
          const float s4 = sgnf(IR[t + 1] - G4[t].s) * Wv[t];   // term 4 against I(u+1, v)
          const float s8 = sgnf(IR[t] - G8[t].s) * Wv[t];       // term 8 against I(u+1, v-1)
          GRR[t + 3] += s4;
          GRR[t + 2] += s8;
          const float x4 = -s4 * G4[t].a, y4 = -s4 * G4[t].b, w4 = -s4 * G4[t].c;
          const float x8 = -s8 * G8[t].a, y8 = -s8 * G8[t].b, w8 = -s8 * G8[t].c;
          const float sum = (x4 + y4 + w4) + (x8 + y8 + w8);
          gS += sum; gSv = fmaf(sum, v, gSv);
          g4X += x4; g4Y += y4; g4W += w4;
          g8X += x8; g8Y += y8; g8W += w8;
          tX = fmaf(x4 + x8, idep, tX); tY = fmaf(y4 + y8, idep, tY); tW = fmaf(w4 + w8, idep, tW);
        }
      }
      // dL/dh11..h33, dL/dt1..t3 of this lane, then summed over the wave
      float g[12];
      g[0] = fmaf(fu, cX + gS, g4X + (g8X + g8Y + g8W));
      g[1] = cXr + gSv - g8X;
      g[2] = cX + gS;
      g[3] = fmaf(fu, cY, g4Y);
      g[4] = cYr - g8Y;
      g[5] = cY;
      g[6] = fmaf(fu, cW, g4W);
      g[7] = cWr - g8W;
      g[8] = cW;
      g[9] = tX; g[10] = tY; g[11] = tW;
#pragma unroll
      for (int i = 0; i < 12; ++i) {
        const float tot = wave_sum(g[i]);
        if (lane == 0) pp[s * 12 + i] = (dc0 == dbeg ? 0.0f : pp[s * 12 + i]) + tot;
      }
    }
    __builtin_amdgcn_wave_barrier();
  }
  // right-neighbour contributions land in the next lane; then sum the four waves
#pragma unroll
  for (int k = 0; k < TY + 6; ++k) {
    GR[k] += wave_shr1z(GRR[k]);
    gsum[wave][k * 64 + lane] = GR[k];
  }
  __syncthreads();
  float* out = A.refpart + (size_t)blk * (TY + 6) * 64;
  for (int e = threadIdx.x; e < (TY + 6) * 64; e += 256)
    out[e] = ((gsum[0][e] + gsum[1][e]) + gsum[2][e]) + gsum[3][e];
}

// dL/dI(x, y): the sum of the tile partials whose window (rows y0-3 .. y0+TY+2, lanes 0..63)
// holds (x, y), in a fixed order.
template <int TY>
__global__ __launch_bounds__(256) void ref_grad_gather_kernel(const float* __restrict__ refpart,
                                                              int W, int H,
                                                              float* __restrict__ gref) {
  const int x = blockIdx.x * 256 + threadIdx.x, y = blockIdx.y;
  if (x >= W) return;
  const int nstrips = (W + 61) / 62, nrows = (H + TY - 1) / TY;
  float acc = 0.0f;
  const int sa = (x + 1) / 62;   // lane x + 1 - 62 sa in [0, 61]
  const int jhi = min(nrows - 1, (y + 3) / TY);   // tile j covers rows j TY - 3 .. j TY + TY + 2
  const int jlo = max(0, jhi - (TY + 6) / TY - 1);
  for (int strip = max(sa - 1, 0); strip <= min(sa, nstrips - 1); ++strip) {
    const int ln = x + 1 - 62 * strip;
    if (ln < 0 || ln > 63) continue;
    for (int j = jlo; j <= jhi; ++j) {
      const int row = y - (j * TY - 3);
      if (row < 0 || row >= TY + 6) continue;
      acc += refpart[((size_t)(j * nstrips + strip) * (TY + 6) + row) * 64 + ln];
    }
  }
  gref[(long)y * W + x] = acc;
}

template <int NV, bool FULL>
pm_status launch_sgm_diff_path(const SgmArgs& a, bool first, bool agg, bool lout,
                               const float* cost, float* aggp, hipStream_t st) {
  const unsigned blocks = (unsigned)((a.nlines + 3) / 4);
  if (agg && lout) {
    if (first) sgm_diff_kernel<NV, true, FULL, true, true><<<blocks, 256, 0, st>>>(a, cost, aggp);
    else sgm_diff_kernel<NV, false, FULL, true, true><<<blocks, 256, 0, st>>>(a, cost, aggp);
  } else if (agg) {
    if (first) sgm_diff_kernel<NV, true, FULL, true, false><<<blocks, 256, 0, st>>>(a, cost, aggp);
    else sgm_diff_kernel<NV, false, FULL, true, false><<<blocks, 256, 0, st>>>(a, cost, aggp);
  } else {
    sgm_diff_kernel<NV, true, FULL, false, true><<<blocks, 256, 0, st>>>(a, cost, aggp);
  }
  return check_launch("sgm_diff_kernel");
}

pm_status sgm_diff_path(const SgmArgs& a, bool first, bool agg, bool lout, const float* cost,
                        float* aggp, hipStream_t st) {
  if (a.D == 64) return launch_sgm_diff_path<1, true>(a, first, agg, lout, cost, aggp, st);
  if (a.D < 64) return launch_sgm_diff_path<1, false>(a, first, agg, lout, cost, aggp, st);
  if (a.D == 128) return launch_sgm_diff_path<2, true>(a, first, agg, lout, cost, aggp, st);
  if (a.D < 128) return launch_sgm_diff_path<2, false>(a, first, agg, lout, cost, aggp, st);
  return launch_sgm_diff_path<4, false>(a, first, agg, lout, cost, aggp, st);
}

template <int NV, bool FULL>
pm_status launch_sgm_adjoint(const SgmAdjArgs& A, bool first, hipStream_t st) {
  const unsigned blocks = (unsigned)((A.s.nlines + 3) / 4);
  constexpr int B = 8 / NV;
  if (first) sgm_adjoint_kernel<NV, true, FULL, B><<<blocks, 256, 0, st>>>(A);
  else sgm_adjoint_kernel<NV, false, FULL, B><<<blocks, 256, 0, st>>>(A);
  return check_launch("sgm_adjoint_kernel");
}

template <int NV, bool FULL>
pm_status launch_sgm_adjoint_paths(const SgmAdjArgs& A, int paths, hipStream_t st) {
  int lines = 0;
  for (int path = 0; path < paths; ++path)
    lines += sgm_lines(A.s.W, A.s.H, kPathDx[path], kPathDy[path]);
  constexpr int B = 8 / NV;   // 16 / NV: twice the registers, half the waves, no faster
  sgm_adjoint_paths_kernel<NV, FULL, B><<<(unsigned)((lines + 3) / 4), 256, 0, st>>>(A, paths);
  return check_launch("sgm_adjoint_paths_kernel");
}

pm_status sgm_adjoint_paths(const SgmAdjArgs& A, int paths, hipStream_t st) {
  const int D = A.s.D;
  if (D == 64) return launch_sgm_adjoint_paths<1, true>(A, paths, st);
  if (D < 64) return launch_sgm_adjoint_paths<1, false>(A, paths, st);
  if (D == 128) return launch_sgm_adjoint_paths<2, true>(A, paths, st);
  if (D < 128) return launch_sgm_adjoint_paths<2, false>(A, paths, st);
  return launch_sgm_adjoint_paths<4, false>(A, paths, st);
}

template <int NV, bool FULL>
pm_status launch_sgm_diff_paths(const SgmArgs& a, int paths, const float* cost, float* lbuf,
                                hipStream_t st) {
  int lines = 0;
  for (int path = 0; path < paths; ++path)
    lines += sgm_lines(a.W, a.H, kPathDx[path], kPathDy[path]);
  sgm_diff_paths_kernel<NV, FULL><<<(unsigned)((lines + 3) / 4), 256, 0, st>>>(a, paths, cost,
                                                                               lbuf);
  return check_launch("sgm_diff_paths_kernel");
}

pm_status sgm_diff_paths(const SgmArgs& a, int paths, const float* cost, float* lbuf,
                         hipStream_t st) {
  if (a.D == 64) return launch_sgm_diff_paths<1, true>(a, paths, cost, lbuf, st);
  if (a.D < 64) return launch_sgm_diff_paths<1, false>(a, paths, cost, lbuf, st);
  if (a.D == 128) return launch_sgm_diff_paths<2, true>(a, paths, cost, lbuf, st);
  if (a.D < 128) return launch_sgm_diff_paths<2, false>(a, paths, cost, lbuf, st);
  return launch_sgm_diff_paths<4, false>(a, paths, cost, lbuf, st);
}

pm_status sgm_adjoint(const SgmAdjArgs& A, bool first, hipStream_t st) {
  const int D = A.s.D;
  if (D == 64) return launch_sgm_adjoint<1, true>(A, first, st);
  if (D < 64) return launch_sgm_adjoint<1, false>(A, first, st);
  if (D == 128) return launch_sgm_adjoint<2, true>(A, first, st);
  if (D < 128) return launch_sgm_adjoint<2, false>(A, first, st);
  return launch_sgm_adjoint<4, false>(A, first, st);
}

template <int PPL>
void launch_readout(const ReadoutArgs& r, hipStream_t st) {
  readout_kernel<PPL><<<kReadoutBlocks, 256, 0, st>>>(r);
}
template <int PPL>
void launch_readout_adjoint(const ReadoutArgs& r, float* gagg, hipStream_t st) {
  readout_adjoint_kernel<PPL><<<kReadoutBlocks, 256, 0, st>>>(r, gagg);
}

}  // namespace

}  // namespace pm

using namespace pm;

// ---- handle -----------------------------------------------------------------------------------
struct pm_diff {
  pm_shape shape;
  int S;
  pm_sgm_params sgm;
  pm_diff_params params;
  long N, V;
  int tiles_fwd, tiles_adj;
  // device buffers
  float* cost = nullptr;
  uint8_t* cnt = nullptr;
  float* agg = nullptr;        // aggregate; overwritten by dL/dagg in the backward
  float* L = nullptr;          // 8 path volumes (store) or one (recompute)
  float* gC = nullptr;
  uint8_t* pvalid = nullptr;
  double* part = nullptr;      // readout partials [kReadoutBlocks][2]
  double* sums = nullptr;      // [2]: sum |depth_hat - gt|, |M|
  float* refpart = nullptr;
  float* posepart = nullptr;
  // forward inputs, reused by the backward
  CostArgs c{};
  const float* gt = nullptr;
  float* depth = nullptr;
  bool have_forward = false;
  size_t bytes = 0;
};

namespace {

SgmArgs diff_sgm_args(const pm_diff* h, int path) {
  SgmArgs a{};
  a.W = h->shape.width; a.H = h->shape.height; a.D = h->shape.planes;
  a.dx = kPathDx[path]; a.dy = kPathDy[path];
  a.nlines = sgm_lines(a.W, a.H, a.dx, a.dy);
  a.ref = h->c.ref;
  a.P1n = h->sgm.pi1;
  a.P2n = h->sgm.pi2;
  a.P1s = h->sgm.pi1 * h->sgm.sgm_q1;
  a.P2s = h->sgm.pi2 * h->sgm.sgm_q2;
  a.tau_i = 0;
  a.tau_f = h->sgm.tau_so;
  a.lout = nullptr;
  return a;
}

ReadoutArgs diff_readout_args(const pm_diff* h) {
  ReadoutArgs r{};
  r.N = h->N; r.D = h->shape.planes;
  r.idep_min = h->c.idep_min; r.dep_sample = h->c.dep_sample;
  r.inv_ts = 1.0f / (h->params.cost_scale * h->params.temperature);
  r.agg = h->agg; r.cnt = h->cnt; r.gt = h->gt; r.depth = h->depth; r.pvalid = h->pvalid;
  r.part = h->part; r.sums = h->sums;
  return r;
}

template <typename F>
void dispatch_ppl(int D, F&& f) {
  const int ppl = (D + 15) / 16;
  if (ppl <= 2) f(std::integral_constant<int, 2>());
  else if (ppl <= 4) f(std::integral_constant<int, 4>());
  else if (ppl <= 8) f(std::integral_constant<int, 8>());
  else f(std::integral_constant<int, 16>());
}

// Runs every path of the SGM into the aggregate; store: also write path r's L to h->L + r V,
// all paths in one launch and the aggregate summed from their volumes afterwards (without
// store, path after path, accumulating the aggregate).
pm_status run_sgm_forward(pm_diff* h, bool store, hipStream_t st) {
  if (store) {
    pm_status s = sgm_diff_paths(diff_sgm_args(h, 0), h->sgm.paths, h->cost, h->L, st);
    if (s != PM_OK) return s;
    const unsigned blocks = (unsigned)((h->N * 16 + 255) / 256);
    const int D = h->shape.planes;
    if (D % 4 == 0) {
      const int d4 = D / 4;
      if (d4 <= 16)
        agg_sum4_kernel<1><<<blocks, 256, 0, st>>>(h->N, D, h->sgm.paths, h->cost, h->L, h->agg);
      else if (d4 <= 32)
        agg_sum4_kernel<2><<<blocks, 256, 0, st>>>(h->N, D, h->sgm.paths, h->cost, h->L, h->agg);
      else
        agg_sum4_kernel<4><<<blocks, 256, 0, st>>>(h->N, D, h->sgm.paths, h->cost, h->L, h->agg);
    } else {
      dispatch_ppl(D, [&](auto ppl) {
        agg_sum_kernel<decltype(ppl)::value><<<blocks, 256, 0, st>>>(
            h->N, D, h->sgm.paths, h->cost, h->L, h->agg);
      });
    }
    return check_launch("agg_sum_kernel");
  }
  for (int path = 0; path < h->sgm.paths; ++path) {
    SgmArgs a = diff_sgm_args(h, path);
    a.lout = store ? h->L + (size_t)path * h->V : nullptr;
    const pm_status s = sgm_diff_path(a, path == 0, true, store, h->cost, h->agg, st);
    if (s != PM_OK) return s;
  }
  return PM_OK;
}

}  // namespace

extern "C" pm_status pm_diff_create(pm_shape shape, int32_t num_sources, pm_sgm_params sgm,
                                    pm_diff_params params, pm_diff** out) {
  clear_error();
  if (!out) return fail(PM_ERR_INVALID_ARGUMENT, "pm_diff_create: null out");
  *out = nullptr;
  if (shape.width <= 0 || shape.height <= 0 || shape.width > 65535 || shape.height > 65535 ||
      (long)shape.width * shape.height >= (1L << 29))
    return fail(PM_ERR_INVALID_ARGUMENT, "pm_diff_create: bad image size");
  if (shape.planes < 2 || shape.planes > PM_MAX_PLANES)
    return fail(PM_ERR_INVALID_ARGUMENT, "pm_diff_create: planes must be in 2..256");
  if ((uint64_t)shape.width * shape.height * shape.planes * 4 >= (1ull << 32))
    return fail(PM_ERR_INVALID_ARGUMENT,
                "pm_diff_create: W H D must be below 2^30 (32-bit offsets)");
  if (num_sources < 1 || num_sources > PM_MAX_SOURCES)
    return fail(PM_ERR_INVALID_ARGUMENT, "pm_diff_create: num_sources must be in 1..16");
  if (sgm.paths != 4 && sgm.paths != 8)
    return fail(PM_ERR_INVALID_ARGUMENT, "pm_diff_create: paths must be 4 or 8");
  const float kn[5] = {sgm.pi1, sgm.pi2, sgm.tau_so, sgm.sgm_q1, sgm.sgm_q2};
  for (float v : kn)
    if (!(std::isfinite(v) && v >= 0.0f))
      return fail(PM_ERR_INVALID_ARGUMENT, "pm_diff_create: SGM knobs must be finite and >= 0");
  if (!(params.temperature > 0.0f && std::isfinite(params.temperature)) ||
      !(params.cost_scale > 0.0f && std::isfinite(params.cost_scale)))
    return fail(PM_ERR_INVALID_ARGUMENT,
                "pm_diff_create: temperature and cost_scale must be finite and > 0");
  pm_diff* h = new pm_diff();
  h->shape = shape; h->S = num_sources; h->sgm = sgm; h->params = params;
  h->N = (long)shape.width * shape.height;
  h->V = h->N * shape.planes;
  h->tiles_fwd = cost_tiles<kCostTy>(shape.width, shape.height);
  h->tiles_adj = cost_tiles<kAdjTy>(shape.width, shape.height);
  const size_t nL = params.recompute ? 1 : (size_t)sgm.paths;
  struct Req { void** p; size_t bytes; };
  const Req reqs[] = {
      {(void**)&h->cost, sizeof(float) * h->V},
      {(void**)&h->cnt, (size_t)h->V},
      {(void**)&h->agg, sizeof(float) * h->V},
      {(void**)&h->L, sizeof(float) * h->V * nL},
      {(void**)&h->gC, sizeof(float) * h->V},
      {(void**)&h->pvalid, (size_t)h->N},
      {(void**)&h->part, sizeof(double) * kReadoutBlocks * 2},
      {(void**)&h->sums, sizeof(double) * 2},
      {(void**)&h->refpart, sizeof(float) * (size_t)h->tiles_adj * (kAdjTy + 6) * 64},
      {(void**)&h->posepart, sizeof(float) * (size_t)h->tiles_adj * 4 * num_sources * 12},
  };
  for (const Req& r : reqs) {
    const hipError_t e = hipMalloc(r.p, r.bytes);
    if (e != hipSuccess) {
      pm_diff_destroy(h);
      return fail(PM_ERR_OUT_OF_MEMORY, std::string("pm_diff_create: ") + hipGetErrorString(e));
    }
    h->bytes += r.bytes;
  }
  *out = h;
  return PM_OK;
}

extern "C" void pm_diff_destroy(pm_diff* h) {
  if (!h) return;
  void* ps[] = {h->cost, h->cnt, h->agg, h->L, h->gC, h->pvalid, h->part, h->sums, h->refpart,
                h->posepart};
  for (void* p : ps)
    if (p) (void)hipFree(p);
  delete h;
}

extern "C" pm_status pm_diff_memory(const pm_diff* h, uint64_t* bytes) {
  clear_error();
  if (!h || !bytes) return fail(PM_ERR_INVALID_ARGUMENT, "pm_diff_memory: null pointer");
  *bytes = h->bytes;
  return PM_OK;
}

extern "C" pm_status pm_diff_download(const pm_diff* h, const char* name, void* host,
                                      size_t bytes) {
  clear_error();
  if (!h || !name || !host) return fail(PM_ERR_INVALID_ARGUMENT, "pm_diff_download: null pointer");
  const std::string n(name);
  const void* src = nullptr;
  size_t size = 0;
  if (n == "cost") { src = h->cost; size = sizeof(float) * h->V; }
  else if (n == "count") { src = h->cnt; size = (size_t)h->V; }
  else if (n == "agg") { src = h->agg; size = sizeof(float) * h->V; }
  else if (n == "grad_cost") { src = h->gC; size = sizeof(float) * h->V; }
  else if (n == "valid") { src = h->pvalid; size = (size_t)h->N; }
  else return fail(PM_ERR_INVALID_ARGUMENT, "pm_diff_download: unknown buffer " + n);
  if (bytes != size)
    return fail(PM_ERR_INVALID_ARGUMENT, "pm_diff_download: " + n + " holds " +
                                             std::to_string(size) + " bytes");
  const hipError_t e = hipMemcpy(host, src, size, hipMemcpyDeviceToHost);
  if (e != hipSuccess)
    return fail(PM_ERR_DEVICE, std::string("pm_diff_download: ") + hipGetErrorString(e));
  return PM_OK;
}

extern "C" pm_status pm_diff_forward(pm_diff* h, const float* ref,
                                     const uint64_t* const* src_quads, const pm_homography* homs,
                                     float idep_min, float dep_sample, const float* gt,
                                     float* depth, double* loss, pm_stream stream) {
  clear_error();
  if (!h || !ref || !src_quads || !homs || !gt || !depth || !loss)
    return fail(PM_ERR_INVALID_ARGUMENT, "pm_diff_forward: null pointer");
  if (!std::isfinite(idep_min) || !std::isfinite(dep_sample))
    return fail(PM_ERR_INVALID_ARGUMENT, "pm_diff_forward: non-finite plane spacing");
  for (int p = 0; p < h->shape.planes; ++p)
    if (!(idep_min + (float)p * dep_sample > 0.0f))
      return fail(PM_ERR_INVALID_ARGUMENT,
                  "pm_diff_forward: every plane's inverse depth must be > 0");
  hipStream_t st = (hipStream_t)stream;
  CostArgs a{};
  a.W = h->shape.width; a.H = h->shape.height; a.D = h->shape.planes; a.S = h->S;
  a.idep_min = idep_min; a.dep_sample = dep_sample; a.ref = ref;
  for (int s = 0; s < h->S; ++s) {
    if (!src_quads[s]) return fail(PM_ERR_INVALID_ARGUMENT, "pm_diff_forward: null source quad");
    a.quad[s] = src_quads[s];
    for (int k = 0; k < 9; ++k) a.hm[s][k] = homs[s].h[k];
    for (int k = 0; k < 3; ++k) a.hm[s][9 + k] = homs[s].t[k];
    for (int k = 0; k < 12; ++k)
      if (!std::isfinite(a.hm[s][k]))
        return fail(PM_ERR_INVALID_ARGUMENT, "pm_diff_forward: non-finite homography");
  }
  h->c = a; h->gt = gt; h->depth = depth; h->have_forward = false;
  cost_diff_kernel<<<h->tiles_fwd, 256, 0, st>>>(a, h->cost, h->cnt);
  pm_status s = check_launch("cost_diff_kernel");
  if (s != PM_OK) return s;
  s = run_sgm_forward(h, !h->params.recompute, st);
  if (s != PM_OK) return s;
  const ReadoutArgs r = diff_readout_args(h);
  dispatch_ppl(h->shape.planes, [&](auto ppl) { launch_readout<decltype(ppl)::value>(r, st); });
  reduce_columns_kernel<double><<<2, 256, 0, st>>>(h->part, kReadoutBlocks, h->sums, 2);
  finish_loss_kernel<<<1, 64, 0, st>>>(h->sums, loss);
  s = check_launch("readout_kernel");
  if (s != PM_OK) return s;
  h->have_forward = true;
  return PM_OK;
}

extern "C" pm_status pm_diff_backward(pm_diff* h, float grad_loss, float* grad_ref,
                                      double* grad_homs, pm_stream stream) {
  clear_error();
  if (!h || !grad_ref || !grad_homs)
    return fail(PM_ERR_INVALID_ARGUMENT, "pm_diff_backward: null pointer");
  if (!h->have_forward)
    return fail(PM_ERR_INVALID_STATE, "pm_diff_backward: no forward pass on this handle");
  if (!std::isfinite(grad_loss))
    return fail(PM_ERR_INVALID_ARGUMENT, "pm_diff_backward: non-finite grad_loss");
  hipStream_t st = (hipStream_t)stream;
  pm_status s;
  if (h->params.recompute) {   // the aggregate again
    s = run_sgm_forward(h, false, st);
    if (s != PM_OK) return s;
  }
  ReadoutArgs r = diff_readout_args(h);
  r.grad_loss = grad_loss;
  dispatch_ppl(h->shape.planes,
               [&](auto ppl) { launch_readout_adjoint<decltype(ppl)::value>(r, h->agg, st); });
  s = check_launch("readout_adjoint_kernel");
  if (s != PM_OK) return s;
  if (h->params.recompute) {
    // path after path: this path's L again, then its reverse sweep added into dL/dC
    for (int path = 0; path < h->sgm.paths; ++path) {
      SgmAdjArgs A{};
      A.s = diff_sgm_args(h, path);
      SgmArgs a = A.s;
      a.lout = h->L;
      s = sgm_diff_path(a, true, false, true, h->cost, nullptr, st);
      if (s != PM_OK) return s;
      A.L = h->L;
      A.gagg = h->agg; A.pvalid = h->pvalid; A.gC = h->gC;
      s = sgm_adjoint(A, path == 0, st);
      if (s != PM_OK) return s;
    }
  } else {
    // every path's reverse sweep at once, each over its own L volume, then their sum
    SgmAdjArgs A{};
    A.s = diff_sgm_args(h, 0);
    A.L = h->L;
    A.gagg = h->agg; A.pvalid = h->pvalid; A.gC = nullptr;
    s = sgm_adjoint_paths(A, h->sgm.paths, st);
    if (s != PM_OK) return s;
    const bool vec4 = h->V % 4 == 0;
    const long items = vec4 ? h->V / 4 : h->V;
    grad_cost_sum_kernel<<<(unsigned)std::min<long>((items + 255) / 256, 65536), 256, 0, st>>>(
        h->V, h->sgm.paths, vec4, h->L, h->gC);
    s = check_launch("grad_cost_sum_kernel");
    if (s != PM_OK) return s;
    // the aggregate now holds dL/dagg and the path volumes hold adjoints: another backward
    // needs another forward
    h->have_forward = false;
  }
  CostAdjArgs C{};
  C.c = h->c; C.gC = h->gC; C.cnt = h->cnt; C.refpart = h->refpart; C.posepart = h->posepart;
  cost_adjoint_kernel<kAdjTy, kAdjDc><<<h->tiles_adj, 256, 0, st>>>(C);
  s = check_launch("cost_adjoint_kernel");
  if (s != PM_OK) return s;
  dim3 g((h->shape.width + 255) / 256, h->shape.height);
  ref_grad_gather_kernel<kAdjTy><<<g, 256, 0, st>>>(h->refpart, h->shape.width, h->shape.height,
                                                    grad_ref);
  s = check_launch("ref_grad_gather_kernel");
  if (s != PM_OK) return s;
  // dL/d(H, t): the S * 12 columns of the per-(tile, wave) partials, summed in FP64
  reduce_columns_kernel<float><<<h->S * 12, 256, 0, st>>>(h->posepart, h->tiles_adj * 4,
                                                          grad_homs, h->S * 12);
  return check_launch("reduce_columns_kernel");
}

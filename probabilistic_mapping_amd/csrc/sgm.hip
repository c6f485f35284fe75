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
#include "sgm_scan.hpp"

#ifndef PM_SGM_PREFETCH
#define PM_SGM_PREFETCH 16
#endif

namespace pm {

// Steps per prefetch block: at most 16 cost values per lane per buffer, so the wider plane
// layouts (NV = 2, 4) keep the ping-pong buffers in registers instead of spilling.
template <int NV, int B = PM_SGM_PREFETCH>
constexpr int sgm_block() { return B * NV > 16 ? 16 / NV : B; }

template <int NV, typename CT, bool FIRST, bool FULL, int B>
__global__ __launch_bounds__(256) void sgm_path_kernel(const SgmArgs a,
                                                       const CT* __restrict__ cost,
                                                       float* __restrict__ agg) {
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int line = blockIdx.x * 4 + wave;
  if (line >= a.nlines) return;
  sgm_scanline<NV, CT, FIRST, FULL, B>(a, line, cost, agg);
}

// Every scanline direction of one keyframe in a single launch.  Global line g counts the lines
// of path 0, then path 1, ...; each path writes its own L volume (0 at invalid pixels) to
// lbuf + path * V, so no two waves touch the same output and all paths run at once.  The
// aggregate is the sum of those volumes in path order (combine_filter_kernel), which is the
// same sum, in the same order, as the path-after-path read-modify-write of sgm_path_kernel.
template <int NV, typename CT, bool FULL, int B>
__global__ __launch_bounds__(256) void sgm_paths_kernel(const SgmArgs base, int paths,
                                                        const CT* __restrict__ cost,
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
  sgm_scanline<NV, CT, true, FULL, B, uint8_t, false, true, true>(a, g, cost, nullptr);
}

// Compact form: every scanline direction of one keyframe in a single launch, each path
// writing its excess E = L - C' as 16-bit levels to ebuf + path * V (combine_excess_filter
// rebuilds the aggregate from the cost and these volumes).
template <int NV, typename CT, bool FULL, int B>
__global__ __launch_bounds__(256) void sgm_excess_kernel(const SgmArgs base, int paths,
                                                         const CT* __restrict__ cost,
                                                         uint16_t* __restrict__ ebuf) {
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
  a.eout = ebuf + (size_t)path * ((size_t)base.W * base.H * base.D);
  sgm_scanline<NV, CT, true, FULL, B, uint8_t, false, false, false, true>(a, g, cost, nullptr);
}

// Batched compact form: the work list of a class (same NV, FULL) is a sequence of segments,
// one per (keyframe, path), each of 4-scanline items; segment j is keyframe seg[2j], path
// seg[2j+1] and owns items first[j] .. first[j+1]-1.  Persistent blocks stride the list.
template <int NV, typename CT, bool FULL, int B>
__global__ __launch_bounds__(256) void sgm_excess_batch_kernel(const KfDesc* __restrict__ kfs,
                                                               const int* __restrict__ seg,
                                                               const int* __restrict__ first,
                                                               int nseg, int total) {
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  for (int i = blockIdx.x; i < total; i += gridDim.x) {
    const int j = owner_of(first, nseg, i);
    const KfDesc& k = kfs[seg[2 * j]];
    const int path = seg[2 * j + 1];
    SgmArgs a;
    a.W = k.c.W; a.H = k.c.H; a.D = k.c.D; a.dx = path_dx(path); a.dy = path_dy(path);
    a.nlines = sgm_lines(a.W, a.H, a.dx, a.dy);
    a.ref = k.c.ref;
    a.P1n = uniform(k.P1n); a.P2n = uniform(k.P2n); a.P1s = uniform(k.P1s);
    a.P2s = uniform(k.P2s); a.tau_i = __builtin_amdgcn_readfirstlane(k.tau_i);
    a.eout = k.ebuf + (size_t)path * a.W * a.H * a.D;
    a.eq_inv = uniform(k.einv);
    const int line = (i - first[j]) * 4 + wave;
    if (line < a.nlines)
      sgm_scanline<NV, CT, true, FULL, B, uint8_t, false, false, false, true>(
          a, line, static_cast<const CT*>(k.cost), nullptr);
  }
}

// Four scanlines per wave (sgm_rows_excess): compact form with 16-bit costs and D = 32, 64,
// 96 or 128.  Register blocks of B steps: PF 16 allows 32 cost dwords per buffer, PF 8 16.
template <int NV, int PF>
constexpr int rows_block() {
  constexpr int b = PF * 2 / (NV / 2);
  return b < 2 ? 2 : (b > 16 ? 16 : b);
}
bool rows_supported(pm_cost_type ct, int D) {
  return (ct == PM_COST_F16 || ct == PM_COST_BF16) && D % 32 == 0 && D >= 32 && D <= 128;
}
int rows_nv(int D) { return D / 16; }

// Single keyframe, every path in one launch: global group g (4 scanlines) counts the groups of
// path 0, then path 1, ...
template <int NV, typename CT, int B>
__global__ __launch_bounds__(256) void sgm_excess_rows_kernel(const SgmArgs base, int paths,
                                                              const CT* __restrict__ cost,
                                                              uint16_t* __restrict__ ebuf) {
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  int g = blockIdx.x * 4 + wave, path = 0, n = 0;
  for (; path < paths; ++path, g -= n) {
    n = (sgm_lines(base.W, base.H, path_dx(path), path_dy(path)) + 3) / 4;
    if (g < n) break;
  }
  if (path == paths) return;
  SgmArgs a = base;
  a.dx = path_dx(path);
  a.dy = path_dy(path);
  a.nlines = sgm_lines(a.W, a.H, a.dx, a.dy);
  a.eout = ebuf + (size_t)path * ((size_t)base.W * base.H * base.D);
  sgm_rows_excess<NV, CT, B>(a, g * 4, cost);
}

// Batched: segments (keyframe, path) as in sgm_excess_batch_kernel, of 16-scanline items.
template <int NV, typename CT, int B>
__global__ __launch_bounds__(256) void sgm_excess_rows_batch_kernel(
    const KfDesc* __restrict__ kfs, const int* __restrict__ seg, const int* __restrict__ first,
    int nseg, int total) {
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  for (int i = blockIdx.x; i < total; i += gridDim.x) {
    const int j = owner_of(first, nseg, i);
    const KfDesc& k = kfs[seg[2 * j]];
    const int path = seg[2 * j + 1];
    SgmArgs a;
    a.W = k.c.W; a.H = k.c.H; a.D = k.c.D; a.dx = path_dx(path); a.dy = path_dy(path);
    a.nlines = sgm_lines(a.W, a.H, a.dx, a.dy);
    a.ref = k.c.ref;
    a.P1n = uniform(k.P1n); a.P2n = uniform(k.P2n); a.P1s = uniform(k.P1s);
    a.P2s = uniform(k.P2s); a.tau_i = __builtin_amdgcn_readfirstlane(k.tau_i);
    a.eout = k.ebuf + (size_t)path * a.W * a.H * a.D;
    a.eq_inv = uniform(k.einv);
    const int line0 = (i - first[j]) * 16 + wave * 4;
    if (line0 < a.nlines) sgm_rows_excess<NV, CT, B>(a, line0, static_cast<const CT*>(k.cost));
  }
}

template <typename CT, int PF>
bool sgm_rows_batch_p(int nv, const KfDesc* kfs, const int* seg, const int* first, int nseg,
                      int total, int grid, hipStream_t st) {
  switch (nv) {
    case 2: sgm_excess_rows_batch_kernel<2, CT, rows_block<2, PF>()><<<grid, 256, 0, st>>>(kfs, seg, first, nseg, total); return true;
    case 4: sgm_excess_rows_batch_kernel<4, CT, rows_block<4, PF>()><<<grid, 256, 0, st>>>(kfs, seg, first, nseg, total); return true;
    case 6: sgm_excess_rows_batch_kernel<6, CT, rows_block<6, PF>()><<<grid, 256, 0, st>>>(kfs, seg, first, nseg, total); return true;
    case 8: sgm_excess_rows_batch_kernel<8, CT, rows_block<8, PF>()><<<grid, 256, 0, st>>>(kfs, seg, first, nseg, total); return true;
    default: return false;
  }
}

bool launch_sgm_excess_rows_batch(pm_cost_type ct, int prefetch, int nv, const KfDesc* kfs,
                                  const int* seg, const int* first, int nseg, int total,
                                  int grid, hipStream_t st) {
  if (prefetch != 8 && prefetch != 16) return false;
  switch (ct) {
    case PM_COST_F16:
      return prefetch == 8 ? sgm_rows_batch_p<__half, 8>(nv, kfs, seg, first, nseg, total, grid, st)
                           : sgm_rows_batch_p<__half, 16>(nv, kfs, seg, first, nseg, total, grid, st);
    case PM_COST_BF16:
      return prefetch == 8
                 ? sgm_rows_batch_p<__hip_bfloat16, 8>(nv, kfs, seg, first, nseg, total, grid, st)
                 : sgm_rows_batch_p<__hip_bfloat16, 16>(nv, kfs, seg, first, nseg, total, grid, st);
    default:
      return false;
  }
}

template <typename CT, int B>
bool sgm_excess_batch_b(int nv, bool full, const KfDesc* kfs, const int* seg, const int* first,
                        int nseg, int total, int grid, hipStream_t st) {
  if (nv == 1 && full) sgm_excess_batch_kernel<1, CT, true, B><<<grid, 256, 0, st>>>(kfs, seg, first, nseg, total);
  else if (nv == 1) sgm_excess_batch_kernel<1, CT, false, B><<<grid, 256, 0, st>>>(kfs, seg, first, nseg, total);
  else if (nv == 2 && full) sgm_excess_batch_kernel<2, CT, true, B><<<grid, 256, 0, st>>>(kfs, seg, first, nseg, total);
  else if (nv == 2) sgm_excess_batch_kernel<2, CT, false, B><<<grid, 256, 0, st>>>(kfs, seg, first, nseg, total);
  else if (nv == 4) sgm_excess_batch_kernel<4, CT, false, sgm_block<4, B>()><<<grid, 256, 0, st>>>(kfs, seg, first, nseg, total);
  else return false;
  return true;
}

template <typename CT>
bool sgm_excess_batch_c(int prefetch, int nv, bool full, const KfDesc* kfs, const int* seg,
                        const int* first, int nseg, int total, int grid, hipStream_t st) {
  if (prefetch == 8) return sgm_excess_batch_b<CT, 8>(nv, full, kfs, seg, first, nseg, total, grid, st);
  if (prefetch == 16) return sgm_excess_batch_b<CT, 16>(nv, full, kfs, seg, first, nseg, total, grid, st);
  return false;
}

bool launch_sgm_excess_batch(pm_cost_type ct, int prefetch, int nv, bool full,
                             const KfDesc* kfs, const int* seg, const int* first, int nseg,
                             int total, int grid, hipStream_t st) {
  switch (ct) {
    case PM_COST_F32: return sgm_excess_batch_c<float>(prefetch, nv, full, kfs, seg, first, nseg, total, grid, st);
    case PM_COST_F16: return sgm_excess_batch_c<__half>(prefetch, nv, full, kfs, seg, first, nseg, total, grid, st);
    case PM_COST_BF16: return sgm_excess_batch_c<__hip_bfloat16>(prefetch, nv, full, kfs, seg, first, nseg, total, grid, st);
    default: return false;
  }
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
    a.P1n = uniform(k.P1n); a.P2n = uniform(k.P2n); a.P1s = uniform(k.P1s);
    a.P2s = uniform(k.P2s); a.tau_i = __builtin_amdgcn_readfirstlane(k.tau_i);
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
  else if (nv == 4) sgm_batch_path_t<4, CT, false, sgm_block<4, B>()>(kfs, kf, first, n, total, grid, first_path, dx, dy, st);
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
  constexpr int B = sgm_block<NV>();
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

template <int NV, typename CT, bool FULL>
void launch_paths_concurrent(const SgmArgs& a, int paths, const CT* cost, float* lbuf,
                             hipStream_t st) {
  constexpr int B = sgm_block<NV>();
  long lines = 0;
  for (int path = 0; path < paths; ++path)
    lines += sgm_lines(a.W, a.H, kPathDx[path], kPathDy[path]);
  const unsigned blocks = (unsigned)((lines + 3) / 4);
  sgm_paths_kernel<NV, CT, FULL, B><<<blocks, 256, 0, st>>>(a, paths, cost, lbuf);
}

template <typename CT>
pm_status launch_sgm_concurrent(const SgmArgs& a, int paths, const CT* cost, float* lbuf,
                                hipStream_t st) {
  if (a.D == 64) launch_paths_concurrent<1, CT, true>(a, paths, cost, lbuf, st);
  else if (a.D == 128) launch_paths_concurrent<2, CT, true>(a, paths, cost, lbuf, st);
  else if (a.D < 64) launch_paths_concurrent<1, CT, false>(a, paths, cost, lbuf, st);
  else if (a.D < 128) launch_paths_concurrent<2, CT, false>(a, paths, cost, lbuf, st);
  else launch_paths_concurrent<4, CT, false>(a, paths, cost, lbuf, st);
  return check_launch("sgm_paths_kernel");
}

template <int NV, typename CT, bool FULL>
void launch_excess_t(const SgmArgs& a, int paths, const CT* cost, uint16_t* ebuf,
                     hipStream_t st) {
  constexpr int B = sgm_block<NV>();
  long lines = 0;
  for (int path = 0; path < paths; ++path)
    lines += sgm_lines(a.W, a.H, kPathDx[path], kPathDy[path]);
  const unsigned blocks = (unsigned)((lines + 3) / 4);
  sgm_excess_kernel<NV, CT, FULL, B><<<blocks, 256, 0, st>>>(a, paths, cost, ebuf);
}

template <int NV, typename CT>
void launch_excess_rows_t(const SgmArgs& a, int paths, const CT* cost, uint16_t* ebuf,
                          hipStream_t st) {
  long groups = 0;
  for (int path = 0; path < paths; ++path)
    groups += (sgm_lines(a.W, a.H, kPathDx[path], kPathDy[path]) + 3) / 4;
  const unsigned blocks = (unsigned)((groups + 3) / 4);
  sgm_excess_rows_kernel<NV, CT, rows_block<NV, PM_SGM_PREFETCH>()><<<blocks, 256, 0, st>>>(
      a, paths, cost, ebuf);
}

template <typename CT>
pm_status launch_sgm_excess(const SgmArgs& a, int paths, const CT* cost, uint16_t* ebuf,
                            hipStream_t st) {
  if constexpr (sizeof(CT) == 2) {
    if (a.D % 32 == 0 && a.D <= 128) {
      switch (a.D) {
        case 32: launch_excess_rows_t<2, CT>(a, paths, cost, ebuf, st); break;
        case 64: launch_excess_rows_t<4, CT>(a, paths, cost, ebuf, st); break;
        case 96: launch_excess_rows_t<6, CT>(a, paths, cost, ebuf, st); break;
        default: launch_excess_rows_t<8, CT>(a, paths, cost, ebuf, st); break;
      }
      return check_launch("sgm_excess_rows_kernel");
    }
  }
  if (a.D == 64) launch_excess_t<1, CT, true>(a, paths, cost, ebuf, st);
  else if (a.D == 128) launch_excess_t<2, CT, true>(a, paths, cost, ebuf, st);
  else if (a.D < 64) launch_excess_t<1, CT, false>(a, paths, cost, ebuf, st);
  else if (a.D < 128) launch_excess_t<2, CT, false>(a, paths, cost, ebuf, st);
  else launch_excess_t<4, CT, false>(a, paths, cost, ebuf, st);
  return check_launch("sgm_excess_kernel");
}

// Validates the common SGM arguments and fills the kernel arguments.
pm_status sgm_args(const std::string& w, pm_shape shape, const uint8_t* ref,
                   pm_sgm_params params, SgmArgs* out) {
  if (shape.width <= 0 || shape.height <= 0 || shape.width > 65535 || shape.height > 65535)
    return fail(PM_ERR_INVALID_ARGUMENT, w + ": bad image size");
  if (shape.planes <= 0 || shape.planes > PM_MAX_PLANES)
    return fail(PM_ERR_INVALID_ARGUMENT, w + ": planes must be in 1..256");
  if (params.paths != 4 && params.paths != 8)
    return fail(PM_ERR_INVALID_ARGUMENT, w + ": paths must be 4 or 8");
  if ((uint64_t)shape.width * shape.height * shape.planes * 4 >= (1ull << 32))
    return fail(PM_ERR_INVALID_ARGUMENT, w + ": W H D must be below 2^30 (32-bit offsets)");
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
  *out = a;
  return PM_OK;
}

pm_status sgm_entry(const char* who, pm_shape shape, const uint8_t* ref, const void* cost,
                    pm_cost_type cost_type, pm_sgm_params params, float* agg, hipStream_t st) {
  const std::string w(who);
  if (!ref || !cost || !agg) return fail(PM_ERR_INVALID_ARGUMENT, w + ": null pointer");
  SgmArgs a;
  const pm_status s = sgm_args(w, shape, ref, params, &a);
  if (s != PM_OK) return s;
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

extern "C" pm_status pm_sgm_workspace_size(pm_shape shape, int32_t paths, uint64_t* bytes) {
  clear_error();
  if (!bytes) return fail(PM_ERR_INVALID_ARGUMENT, "pm_sgm_workspace_size: null pointer");
  if (shape.width <= 0 || shape.height <= 0 || shape.planes <= 0 || (paths != 4 && paths != 8))
    return fail(PM_ERR_INVALID_ARGUMENT, "pm_sgm_workspace_size: bad shape or path count");
  *bytes = (uint64_t)paths * shape.width * shape.height * shape.planes * sizeof(float);
  return PM_OK;
}

extern "C" pm_status pm_sgm_filter_paths(pm_shape shape, const uint8_t* ref, const void* cost,
                                         pm_cost_type cost_type, pm_sgm_params params,
                                         float var_scale, float idep_min, float dep_sample,
                                         void* workspace, uint64_t workspace_bytes, float* agg,
                                         float* depth, pm_stream stream) {
  clear_error();
  const std::string w("pm_sgm_filter_paths");
  if (!ref || !cost || !workspace || !depth)
    return fail(PM_ERR_INVALID_ARGUMENT, w + ": null pointer");
  SgmArgs a;
  pm_status s = sgm_args(w, shape, ref, params, &a);
  if (s != PM_OK) return s;
  uint64_t need = 0;
  s = pm_sgm_workspace_size(shape, params.paths, &need);
  if (s != PM_OK) return s;
  if (workspace_bytes < need)
    return fail(PM_ERR_INVALID_ARGUMENT, w + ": workspace smaller than pm_sgm_workspace_size");
  if ((uintptr_t)workspace % 16 != 0)
    return fail(PM_ERR_INVALID_ARGUMENT, w + ": workspace must be 16-byte aligned");
  hipStream_t st = (hipStream_t)stream;
  float* lbuf = static_cast<float*>(workspace);
  switch (cost_type) {
    case PM_COST_F32: s = launch_sgm_concurrent<float>(a, params.paths, (const float*)cost, lbuf, st); break;
    case PM_COST_F16: s = launch_sgm_concurrent<__half>(a, params.paths, (const __half*)cost, lbuf, st); break;
    case PM_COST_BF16:
      s = launch_sgm_concurrent<__hip_bfloat16>(a, params.paths, (const __hip_bfloat16*)cost, lbuf, st);
      break;
    default: return fail(PM_ERR_INVALID_ARGUMENT, w + ": unknown cost type");
  }
  if (s != PM_OK) return s;
  launch_combine_filter(a.W, a.H, a.D, params.paths, lbuf, agg, var_scale, idep_min, dep_sample,
                        depth, st);
  return check_launch("combine_filter_kernel");
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

extern "C" pm_status pm_sgm_excess_workspace_size(pm_shape shape, int32_t paths,
                                                  uint64_t* bytes) {
  clear_error();
  if (!bytes) return fail(PM_ERR_INVALID_ARGUMENT, "pm_sgm_excess_workspace_size: null pointer");
  if (shape.width <= 0 || shape.height <= 0 || shape.planes <= 0 || (paths != 4 && paths != 8))
    return fail(PM_ERR_INVALID_ARGUMENT,
                "pm_sgm_excess_workspace_size: bad shape or path count");
  *bytes = (uint64_t)paths * shape.width * shape.height * shape.planes * sizeof(uint16_t);
  return PM_OK;
}

extern "C" pm_status pm_sgm_filter_excess(pm_shape shape, const uint8_t* ref, const void* cost,
                                          pm_cost_type cost_type, pm_sgm_params params,
                                          float var_scale, float idep_min, float dep_sample,
                                          void* workspace, uint64_t workspace_bytes, float* agg,
                                          float* depth, pm_stream stream) {
  clear_error();
  const std::string w("pm_sgm_filter_excess");
  if (!ref || !cost || !workspace || !depth)
    return fail(PM_ERR_INVALID_ARGUMENT, w + ": null pointer");
  SgmArgs a;
  pm_status s = sgm_args(w, shape, ref, params, &a);
  if (s != PM_OK) return s;
  uint64_t need = 0;
  s = pm_sgm_excess_workspace_size(shape, params.paths, &need);
  if (s != PM_OK) return s;
  if (workspace_bytes < need)
    return fail(PM_ERR_INVALID_ARGUMENT,
                w + ": workspace smaller than pm_sgm_excess_workspace_size");
  if ((uintptr_t)workspace % 16 != 0)
    return fail(PM_ERR_INVALID_ARGUMENT, w + ": workspace must be 16-byte aligned");
  const float step = excess_step(a.P2n, a.P2s);
  a.eq_inv = step > 0.0f ? 1.0f / step : 0.0f;
  hipStream_t st = (hipStream_t)stream;
  uint16_t* ebuf = static_cast<uint16_t*>(workspace);
  switch (cost_type) {
    case PM_COST_F32: s = launch_sgm_excess<float>(a, params.paths, (const float*)cost, ebuf, st); break;
    case PM_COST_F16: s = launch_sgm_excess<__half>(a, params.paths, (const __half*)cost, ebuf, st); break;
    case PM_COST_BF16:
      s = launch_sgm_excess<__hip_bfloat16>(a, params.paths, (const __hip_bfloat16*)cost, ebuf, st);
      break;
    default: return fail(PM_ERR_INVALID_ARGUMENT, w + ": unknown cost type");
  }
  if (s != PM_OK) return s;
  launch_combine_excess(cost_type, a.W, a.H, a.D, params.paths, cost, ebuf, step, agg,
                        var_scale, idep_min, dep_sample, depth, st);
  return check_launch("combine_excess_filter_kernel");
}

// The fusion step of outputFusionCPU with CASE >= 3: the reference's CPU rules, on the GPU.
//
// Reference: stereo_mapper.cpp:719-800 (the CASE >= 3 branch: measurement clamps, the pose
// composition, depthFuseCPU, output masks) and depthFuseCPU's four loops, :246-587:
// transformMapCPU (:246-326), holeFillingCPU (:328-388), depthPreditCPU (:390-431) and
// depthUpdateCPU (:439-545) with normpdfCPU (:434-437).  They differ from the GPU fusion
// (fusion.hip): propagation gate 0.5, a 7x7 hole fill, tau^2 = 64, initial variance 256,
// outlier density 1/16, beta += 1 on a missing measurement, the output gated on the
// posterior inlier ratio (> 0.51), and a table entry (y << 16) | x.
//
// The reference runs these loops serially in raster order, and two of them depend on that
// order; both are reproduced exactly:
//   - z-buffer (:312-325): the first arrival at a cell keeps it unless its transformed depth
//     exceeds a later arrival's ORIGINAL depth; the displaced arrival's mean is then set to
//     that original depth.  Here the arrivals are grouped by cell with a counting sort
//     (cpu_transform_kernel counts them, two scan kernels turn the counts into offsets,
//     cpu_scatter_kernel writes each source index into its cell's segment), and
//     cpu_zbuffer_kernel puts a cell's segment in raster order and folds it.
//   - hole filling (:361-388): sources are visited in raster order and fill the still-empty
//     cells of their 7x7 windows, so an empty cell ends with the entry of the first interior
//     source, in raster order, inside its own 7x7 window (cpu_update_kernel reads it from the
//     pre-fill table).
#include <cmath>

#include "common.hpp"

namespace pm {

namespace {

constexpr int kExpand = 3;        // holeFillingCPU's expand_sz
constexpr int kScanBlock = 1024;  // cells per block of the offset scan (256 threads x 4)
constexpr int kSmallCell = 32;    // arrivals a cell orders by repeated minimum; more: heap sort

struct CpuFuseArgs {
  int W, H;
  float fx, fy, cx, cy;
  double r[9], t[3];
};

// scratch layout (all [N] unless noted)
struct CpuFuseWork {
  int32_t* cnt;      // arrivals per cell (zeroed)
  int32_t* cur;      // scatter cursor per cell (zeroed; follows cnt)
  int32_t* off;      // exclusive prefix of cnt within its scan block
  int32_t* bsum;     // [blocks] exclusive prefix of the scan blocks' totals
  int32_t* list;     // source indices grouped by target cell
  int32_t* tgt;      // target cell of each source, -1 if it does not propagate
  float* zt;         // transformed depth of each propagating source
  float* dorig;      // its depth before the transform
  uint32_t* raw;     // pre-fill table
};

inline long scan_blocks(long N) { return (N + kScanBlock - 1) / kScanBlock; }

inline size_t cpu_fuse_work_bytes(long N) {
  return (size_t)N * (4 * 8 + 4) + (size_t)scan_blocks(N) * 4;
}

CpuFuseWork carve(void* base, long N) {
  char* p = static_cast<char*>(base);
  CpuFuseWork w;
  w.cnt = reinterpret_cast<int32_t*>(p); p += (size_t)N * 4;
  w.cur = reinterpret_cast<int32_t*>(p); p += (size_t)N * 4;
  w.off = reinterpret_cast<int32_t*>(p); p += (size_t)N * 4;
  w.list = reinterpret_cast<int32_t*>(p); p += (size_t)N * 4;
  w.tgt = reinterpret_cast<int32_t*>(p); p += (size_t)N * 4;
  w.zt = reinterpret_cast<float*>(p); p += (size_t)N * 4;
  w.dorig = reinterpret_cast<float*>(p); p += (size_t)N * 4;
  w.raw = reinterpret_cast<uint32_t*>(p); p += (size_t)N * 4;
  w.bsum = reinterpret_cast<int32_t*>(p);
  return w;
}

__global__ __launch_bounds__(256) void zero_i32_kernel(int32_t* p, long n) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0;
}

// transformMapCPU up to the z-buffer (:270-311)
__global__ __launch_bounds__(256) void cpu_transform_kernel(const CpuFuseArgs a,
                                                            const float* __restrict__ alpha,
                                                            const float* __restrict__ beta,
                                                            float* __restrict__ mu,
                                                            float* __restrict__ sig,
                                                            CpuFuseWork w) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x;
  const int y = blockIdx.y;
  if (x >= a.W) return;
  const long i = (long)y * a.W + x;
  w.tgt[i] = -1;
  const float al = alpha[i];
  if (al < kFloatEps) return;                                        // :275-278
  const float be = beta[i];
  if ((double)(al / (al + be)) < 0.5) return;                        // :279-282
  const float depth = mu[i];
  // p in float as the reference forms it, p2 = R p + t in double (:290-292)
  const double p0 = (double)(((float)x - a.cx) / a.fx * depth);
  const double p1 = (double)(((float)y - a.cy) / a.fy * depth);
  const double p2 = (double)depth;
  const double q0 = a.r[0] * p0 + a.r[1] * p1 + a.r[2] * p2 + a.t[0];
  const double q1 = a.r[3] * p0 + a.r[4] * p1 + a.r[5] * p2 + a.t[1];
  const double q2 = a.r[6] * p0 + a.r[7] * p1 + a.r[8] * p2 + a.t[2];
  if (q2 < kFloatEps) return;                                        // :293-299
  const int x2 = (int)(q0 / q2 * (double)a.fx + (double)a.cx + 0.5);   // :300-301 (truncation)
  const int y2 = (int)(q1 / q2 * (double)a.fy + (double)a.cy + 0.5);
  if (x2 < 0 || x2 >= a.W || y2 < 0 || y2 >= a.H) return;           // :302-308
  const float z2 = (float)q2;
  mu[i] = z2;                                                        // :310-311
  sig[i] = (float)((double)sig[i] + q2 * 0.05);
  const long cell = (long)y2 * a.W + x2;
  w.zt[i] = z2;
  w.dorig[i] = depth;
  w.tgt[i] = (int32_t)cell;
  atomicAdd(&w.cnt[cell], 1);
}

// exclusive scan of cnt within blocks of kScanBlock cells; the block totals go to bsum
__global__ __launch_bounds__(256) void cpu_scan_blocks_kernel(long N, CpuFuseWork w) {
  __shared__ int32_t part[256];
  const long base = (long)blockIdx.x * kScanBlock + threadIdx.x * 4;
  int32_t v[4];
  int32_t sum = 0;
  for (int k = 0; k < 4; ++k) {
    v[k] = base + k < N ? w.cnt[base + k] : 0;
    sum += v[k];
  }
  part[threadIdx.x] = sum;
  __syncthreads();
  for (int d = 1; d < 256; d <<= 1) {   // inclusive Hillis-Steele scan of the thread sums
    const int32_t t = threadIdx.x >= d ? part[threadIdx.x - d] : 0;
    __syncthreads();
    part[threadIdx.x] += t;
    __syncthreads();
  }
  int32_t run = part[threadIdx.x] - sum;
  for (int k = 0; k < 4; ++k) {
    if (base + k < N) w.off[base + k] = run;
    run += v[k];
  }
  if (threadIdx.x == 255) w.bsum[blockIdx.x] = part[255];
}

// exclusive scan of the block totals, in place, by one block looping over chunks of 256
__global__ __launch_bounds__(256) void cpu_scan_sums_kernel(long blocks, CpuFuseWork w) {
  __shared__ int32_t part[256];
  int32_t carry = 0;
  for (long c0 = 0; c0 < blocks; c0 += 256) {
    const long i = c0 + threadIdx.x;
    const int32_t v = i < blocks ? w.bsum[i] : 0;
    part[threadIdx.x] = v;
    __syncthreads();
    for (int d = 1; d < 256; d <<= 1) {
      const int32_t t = threadIdx.x >= d ? part[threadIdx.x - d] : 0;
      __syncthreads();
      part[threadIdx.x] += t;
      __syncthreads();
    }
    if (i < blocks) w.bsum[i] = carry + part[threadIdx.x] - v;
    carry += part[255];
    __syncthreads();
  }
}

__device__ __forceinline__ long cell_base(const CpuFuseWork& w, long c) {
  return (long)w.off[c] + (long)w.bsum[c / kScanBlock];
}

// each propagating source writes its index into its target cell's segment (in any order)
__global__ __launch_bounds__(256) void cpu_scatter_kernel(long N, CpuFuseWork w) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  const int32_t c = w.tgt[i];
  if (c < 0) return;
  const int32_t k = atomicAdd(&w.cur[c], 1);
  w.list[cell_base(w, c) + k] = (int32_t)i;
}

// in-place ascending heap sort of a cell's segment (cells with many arrivals)
__device__ void heap_sort(int32_t* a, long n) {
  auto sift = [&](long root, long end) {
    while (2 * root + 1 < end) {
      long child = 2 * root + 1;
      if (child + 1 < end && a[child] < a[child + 1]) ++child;
      if (!(a[root] < a[child])) return;
      const int32_t t = a[root]; a[root] = a[child]; a[child] = t;
      root = child;
    }
  };
  for (long s = n / 2 - 1; s >= 0; --s) sift(s, n);
  for (long end = n - 1; end > 0; --end) {
    const int32_t t = a[0]; a[0] = a[end]; a[end] = t;
    sift(0, end);
  }
}

// the z-buffer of transformMapCPU (:312-325), folded over a cell's arrivals in raster order
__global__ __launch_bounds__(256) void cpu_zbuffer_kernel(const CpuFuseArgs a,
                                                          float* __restrict__ mu,
                                                          CpuFuseWork w) {
  const long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const long N = (long)a.W * a.H;
  if (c >= N) return;
  const int n = w.cnt[c];
  const long base = n > 0 ? cell_base(w, c) : 0;
  uint32_t e = 0;
  long occ = -1;
  auto arrive = [&](long s) {
    if (e == 0) {                                   // empty, or held by pixel (0, 0)
      e = ((uint32_t)(s / a.W) << 16) | (uint32_t)(s % a.W);
      occ = s;
    } else if (!(w.zt[occ] <= w.dorig[s])) {
      mu[occ] = w.dorig[s];                         // :322
      e = ((uint32_t)(s / a.W) << 16) | (uint32_t)(s % a.W);
      occ = s;
    }
  };
  // arrival order = raster order of the sources
  if (n <= kSmallCell) {
    int32_t prev = -1;
    for (int k = 0; k < n; ++k) {                   // the next larger index, n times
      int32_t next = 0x7fffffff;
      for (int m = 0; m < n; ++m) {
        const int32_t v = w.list[base + m];
        if (v > prev && v < next) next = v;
      }
      arrive(next);
      prev = next;
    }
  } else {
    heap_sort(w.list + base, n);
    for (int k = 0; k < n; ++k) arrive(w.list[base + k]);
  }
  w.raw[c] = e;
}

// normpdfCPU (:434-437): the exponential in float, times the square root in double, rounded
// to float on return
__device__ __forceinline__ float normpdf_cpu(float x, float m, float sigma_sq) {
  const float ex = expf(-(x - m) * (x - m) / (2.0f * sigma_sq));
  return (float)((double)ex * sqrt(2.0 * M_PI * (double)sigma_sq));
}

// holeFillingCPU + depthPreditCPU + depthUpdateCPU for one cell, with the measurement clamps
// and output masks of outputFusionCPU (:742-775)
__global__ __launch_bounds__(256) void cpu_update_kernel(
    const CpuFuseArgs a, const uint32_t* __restrict__ raw, int32_t* __restrict__ table,
    const float* __restrict__ pa, const float* __restrict__ pb, const float* __restrict__ pm_,
    const float* __restrict__ ps, float* __restrict__ ca, float* __restrict__ cb,
    float* __restrict__ cm, float* __restrict__ cs, float* __restrict__ depth) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x;
  const int y = blockIdx.y;
  const int W = a.W, H = a.H;
  if (x >= W) return;
  const long j = (long)y * W + x;
  uint32_t e = raw[j];
  if (e == 0) {
    for (int dy = -kExpand; dy <= kExpand && e == 0; ++dy) {
      const int sy = y + dy;
      if (sy < kExpand || sy >= H - kExpand) continue;
      for (int dx = -kExpand; dx <= kExpand; ++dx) {
        const int sx = x + dx;
        if (sx < kExpand || sx >= W - kExpand) continue;
        const uint32_t v = raw[(long)sy * W + sx];
        if (v != 0) { e = v; break; }
      }
    }
  }
  if (table) table[j] = (int32_t)e;
  // depthPreditCPU: no source -> alpha 0 only (the rest keeps stale values)
  float al = 0.0f, bv = 0.0f, mv = 0.0f, sv = 0.0f;
  if (e == 0) {
    ca[j] = 0.0f;
  } else {
    const long s = (long)(e >> 16) * W + (long)(e & 0xFFFFu);
    al = pa[s]; bv = pb[s]; mv = pm_[s]; sv = ps[s];
    ca[j] = al; cb[j] = bv; cm[j] = mv; cs[j] = sv;
  }
  // the measurement: the filtered depth with the clamps of :742-757
  float d = depth[j];
  const bool masked = d > 100.0f;
  if (masked) d = 0.0f;
  if (d > 10.0f) d = 0.0f;
  const double thr = (double)0.51f;
  float out;
  if (al < kFloatEps) {                                              // :467-494
    out = 0.0f;
    if (d < kFloatEps) {
      cm[j] = d;
    } else {
      ca[j] = 10.0f; cb[j] = 10.0f; cm[j] = d; cs[j] = 256.0f;
    }
  } else if (d < kFloatEps) {                                        // :496-518
    bv = bv + 1.0f;
    cb[j] = bv;
    out = ((double)(al / (al + bv)) > thr) ? mv : 0.0f;
  } else {                                                           // :521-545
    const float unc = 64.0f;
    const float aa = al, bb = bv, miu = mv, sq = sv;
    const float new_sq = unc * sq / (unc + sq);
    const float new_miu = (d * sq + miu * unc) / (unc + sq);
    float c1 = (aa / (aa + bb)) * normpdf_cpu(d, miu, unc + sq);
    float c2 = (bb / (aa + bb)) * 1.0f / 16.0f;
    const float nc = c1 + c2;
    c1 = c1 / nc;
    c2 = c2 / nc;
    const float f = c1 * ((aa + 1.0f) / (aa + bb + 1.0f)) + c2 * (aa / (aa + bb + 1.0f));
    const float ee = c1 * (((aa + 1.0f) * (aa + 2.0f)) / ((aa + bb + 1.0f) * (aa + bb + 2.0f))) +
                     c2 * (aa * (aa + 1.0f) / ((aa + bb + 1.0f) * (aa + bb + 2.0f)));
    const float mu_p = c1 * new_miu + c2 * miu;
    const float sig_p = c1 * (new_sq + new_miu * new_miu) + c2 * (sq + miu * miu) - mu_p * mu_p;
    const float a_p = (ee - f) / (f - ee / f);
    const float b_p = a_p * (1.0f - f) / f;
    ca[j] = a_p; cb[j] = b_p; cm[j] = mu_p; cs[j] = sig_p;
    out = ((double)(a_p / (a_p + b_p)) > thr) ? mu_p : 0.0f;
  }
  if (masked) out = 0.0f;                                            // :765-775
  if (out < 0.1f) out = 0.0f;
  depth[j] = out;
}

}  // namespace

}  // namespace pm

using namespace pm;

extern "C" pm_status pm_depth_fuse_cpu_rules_workspace_size(int32_t width, int32_t height,
                                                            uint64_t* bytes) {
  clear_error();
  if (!bytes || width <= 0 || height <= 0 || width > 65535 || height > 65535)
    return fail(PM_ERR_INVALID_ARGUMENT, "pm_depth_fuse_cpu_rules_workspace_size: bad argument");
  if ((long)width * height > 0x7fffffffL)
    return fail(PM_ERR_INVALID_ARGUMENT, "pm_depth_fuse_cpu_rules_workspace_size: image too large");
  *bytes = cpu_fuse_work_bytes((long)width * height);
  return PM_OK;
}

extern "C" pm_status pm_depth_fuse_cpu_rules(int32_t width, int32_t height, pm_camera cam,
                                             const double r_p2c[9], const double t_p2c[3],
                                             float* depth, int32_t* table, void* work,
                                             uint64_t work_bytes, pm_fusion_state pre,
                                             pm_fusion_state cur, pm_stream stream) {
  clear_error();
  if (!r_p2c || !t_p2c || !depth || !work || !pre.alpha || !pre.beta || !pre.mu || !pre.sig ||
      !cur.alpha || !cur.beta || !cur.mu || !cur.sig)
    return fail(PM_ERR_INVALID_ARGUMENT, "pm_depth_fuse_cpu_rules: null pointer");
  if (width <= 0 || height <= 0 || width > 65535 || height > 65535)
    return fail(PM_ERR_INVALID_ARGUMENT, "pm_depth_fuse_cpu_rules: bad image size");
  const long N = (long)width * height;
  if (N > 0x7fffffffL)   // cells and sources are int32 indices
    return fail(PM_ERR_INVALID_ARGUMENT, "pm_depth_fuse_cpu_rules: image too large");
  if (work_bytes < cpu_fuse_work_bytes(N))
    return fail(PM_ERR_INVALID_ARGUMENT, "pm_depth_fuse_cpu_rules: workspace too small");
  if (pre.alpha == cur.alpha || pre.beta == cur.beta || pre.mu == cur.mu || pre.sig == cur.sig)
    return fail(PM_ERR_INVALID_ARGUMENT, "pm_depth_fuse_cpu_rules: pre and cur must not alias");
  CpuFuseArgs a{};
  a.W = width; a.H = height;
  a.fx = cam.fx; a.fy = cam.fy; a.cx = cam.cx; a.cy = cam.cy;
  if (!(std::isfinite(a.fx) && std::isfinite(a.fy) && std::isfinite(a.cx) &&
        std::isfinite(a.cy) && a.fx != 0.0f && a.fy != 0.0f))
    return fail(PM_ERR_INVALID_ARGUMENT, "pm_depth_fuse_cpu_rules: bad intrinsics");
  for (int k = 0; k < 9; ++k) {
    a.r[k] = r_p2c[k];
    if (!std::isfinite(a.r[k])) return fail(PM_ERR_INVALID_ARGUMENT, "pm_depth_fuse_cpu_rules: bad pose");
  }
  for (int k = 0; k < 3; ++k) {
    a.t[k] = t_p2c[k];
    if (!std::isfinite(a.t[k])) return fail(PM_ERR_INVALID_ARGUMENT, "pm_depth_fuse_cpu_rules: bad pose");
  }
  hipStream_t st = (hipStream_t)stream;
  const CpuFuseWork w = carve(work, N);
  const unsigned lin = (unsigned)((N + 255) / 256);
  const long blocks = scan_blocks(N);
  zero_i32_kernel<<<(unsigned)((2 * N + 255) / 256), 256, 0, st>>>(w.cnt, 2 * N);   // cnt, cur
  pm_status s = check_launch("zero_i32_kernel");
  if (s != PM_OK) return s;
  const dim3 grid((width + 255) / 256, height);
  cpu_transform_kernel<<<grid, 256, 0, st>>>(a, pre.alpha, pre.beta, pre.mu, pre.sig, w);
  s = check_launch("cpu_transform_kernel");
  if (s != PM_OK) return s;
  cpu_scan_blocks_kernel<<<(unsigned)blocks, 256, 0, st>>>(N, w);
  s = check_launch("cpu_scan_blocks_kernel");
  if (s != PM_OK) return s;
  cpu_scan_sums_kernel<<<1, 256, 0, st>>>(blocks, w);
  s = check_launch("cpu_scan_sums_kernel");
  if (s != PM_OK) return s;
  cpu_scatter_kernel<<<lin, 256, 0, st>>>(N, w);
  s = check_launch("cpu_scatter_kernel");
  if (s != PM_OK) return s;
  cpu_zbuffer_kernel<<<lin, 256, 0, st>>>(a, pre.mu, w);
  s = check_launch("cpu_zbuffer_kernel");
  if (s != PM_OK) return s;
  cpu_update_kernel<<<grid, 256, 0, st>>>(a, w.raw, table, pre.alpha, pre.beta, pre.mu, pre.sig,
                                          cur.alpha, cur.beta, cur.mu, cur.sig, depth);
  return check_launch("cpu_update_kernel");
}

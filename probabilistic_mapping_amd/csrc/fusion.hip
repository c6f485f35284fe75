// Gaussian x Beta depth fusion (one keyframe step).
//
// Reference: depth_fuse and its four kernels, motion_stereo/depth_fusion.cu:8-280:
// transformMap (:8-78) forward-warps the previous filter state through a z-buffer built with
// a bounded atomicCAS loop, hole_filling (:80-105) dilates the propagation table in place,
// depth_predict (:107-137) gathers the predicted state and depthUpdate (:139-225) applies
// the Vogiatzis-Hernandez moment-matched update.  Each reference kernel ends in a device-wide
// synchronisation, and both table kernels are racy.
//
// Here the step is two kernels and deterministic:
//   fuse_transform_kernel : gate + transform + in-place mu/sig update of the previous state,
//       then a 64-bit atomicMin on (transformed depth bits << 32 | source index): the nearest
//       transformed depth wins, ties go to the lowest source index.  Source pixel 0 is the
//       table's empty sentinel and never inserts (depth_fusion.cu:54, 61, 93, 128).
//   fuse_update_kernel : resolves the table entry, fills a hole from the pre-fill z-buffer
//       (lowest-linear-index interior 3x3 neighbour with an entry), gathers the predicted
//       state and runs the update, writing the fused depth.
// Both rules are admissible outcomes of the reference's races.
//
// The z-buffer must be empty (all bits set) when the transform starts.  The update kernel
// cannot empty its own entries behind it, since the hole fill reads the neighbours' entries
// from other threads.  So a caller that runs steps back to back gives each step a second
// buffer, zbuf_next, which the transform kernel empties pixel by pixel for the next step (the
// two buffers alternate); without one, the step empties its z-buffer first (one more launch).
#include <mutex>
#include <string>

#include "common.hpp"

namespace pm {

constexpr uint64_t kEmpty = ~0ull;

struct FuseArgs {
  int W, H;
  float fx, fy, cx, cy;
  float r[9], t[3];
};

__global__ __launch_bounds__(256) void fill_u64_kernel(uint64_t* p, long n, uint64_t v) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

__global__ __launch_bounds__(256) void fill_f32_kernel(float* p, long n, float v) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

__global__ __launch_bounds__(256) void fuse_transform_kernel(const FuseArgs a,
                                                             const float* __restrict__ alpha,
                                                             const float* __restrict__ beta,
                                                             float* __restrict__ mu,
                                                             float* __restrict__ sig,
                                                             uint64_t* __restrict__ zbuf,
                                                             uint64_t* __restrict__ znext) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x;
  const int y = blockIdx.y;
  if (x >= a.W) return;
  const long i = (long)y * a.W + x;
  if (znext) znext[i] = kEmpty;   // the next step's z-buffer, emptied for it
  const float al = alpha[i];
  if (al < kFloatEps) return;                                   // :26
  const float be = beta[i];
  if ((double)(al / (al + be)) < 0.4) return;                   // :30
  const float depth = mu[i];
  const float px = ((float)x - a.cx) / a.fx * depth;            // :34-36
  const float py = ((float)y - a.cy) / a.fy * depth;
  const float pz = depth;
  const float px2 = a.r[0] * px + a.r[1] * py + a.r[2] * pz + a.t[0];   // :37-39
  const float py2 = a.r[3] * px + a.r[4] * py + a.r[5] * pz + a.t[1];
  const float pz2 = a.r[6] * px + a.r[7] * py + a.r[8] * pz + a.t[2];
  if (pz2 < kFloatEps) return;                                  // :40
  const int x2 = (int)((double)(px2 / pz2 * a.fx + a.cx) + 0.5);   // :43-44 (truncation)
  const int y2 = (int)((double)(py2 / pz2 * a.fy + a.cy) + 0.5);
  if (x2 < 0 || x2 >= a.W || y2 < 0 || y2 >= a.H) return;     // :45
  mu[i] = pz2;                                                  // :48-49
  sig[i] = (float)((double)sig[i] + (double)pz2 * 0.05);
  if (i == 0) return;
  const uint64_t key = ((uint64_t)__float_as_uint(pz2) << 32) | (uint64_t)(uint32_t)i;
  atomicMin((unsigned long long*)(zbuf + (long)y2 * a.W + x2), (unsigned long long)key);
}

__device__ __forceinline__ int32_t zbuf_entry(const uint64_t* zbuf, long j) {
  const uint64_t k = zbuf[j];
  return k == kEmpty ? 0 : (int32_t)(uint32_t)(k & 0xFFFFFFFFull);
}

// normpdf, depth_fusion.cu:3-6
__device__ __forceinline__ float normpdf(float x, float mu, float sigma_sq) {
  return expf(-(x - mu) * (x - mu) / (2.0f * sigma_sq)) *
         rsqrtf((float)(2.0 * M_PI * (double)sigma_sq));
}

__global__ __launch_bounds__(256) void fuse_update_kernel(
    const FuseArgs a, const uint64_t* __restrict__ zbuf, int32_t* __restrict__ table,
    const float* __restrict__ pa, const float* __restrict__ pb, const float* __restrict__ pm_,
    const float* __restrict__ ps, float* __restrict__ ca, float* __restrict__ cb,
    float* __restrict__ cm, float* __restrict__ cs, const float* depth, float* fused) {
  // depth and fused may be the same buffer (in place, as tmpDepthMap): each thread reads its
  // own pixel of depth before writing that pixel of fused
  const int x = blockIdx.x * blockDim.x + threadIdx.x;
  const int y = blockIdx.y;
  const int W = a.W, H = a.H;
  if (x >= W) return;
  const long j = (long)y * W + x;
  // table entry, with hole filling from the pre-fill z-buffer (:87-104)
  int32_t src = zbuf_entry(zbuf, j);
  if (src == 0) {
    for (int dy = -1; dy <= 1 && src == 0; ++dy)
      for (int dx = -1; dx <= 1 && src == 0; ++dx) {
        const int nx = x + dx, ny = y + dy;
        if (nx <= 0 || nx >= W - 1 || ny <= 0 || ny >= H - 1) continue;
        src = zbuf_entry(zbuf, (long)ny * W + nx);
      }
  }
  if (table) table[j] = src;
  // depth_predict (:123-136); with no source only alpha is written (beta, mu, sig keep
  // stale values, as in the reference; they are not read while alpha < FLOAT_EPS)
  float al = -1.0f, bv = 0.0f, mv = 0.0f, sv = 0.0f;
  if (src == 0) {
    ca[j] = al;
  } else {
    al = pa[src]; bv = pb[src]; mv = pm_[src]; sv = ps[src];
    ca[j] = al; cb[j] = bv; cm[j] = mv; cs[j] = sv;
  }
  // depthUpdate (:150-224)
  const float dest = depth[j];
  const float unc = (float)((double)(dest * dest) * 0.01);
  if (al < kFloatEps) {
    if (dest < kFloatEps) {
      fused[j] = 0.0f;
    } else {
      ca[j] = 10.0f; cb[j] = 10.0f; cm[j] = dest; cs[j] = 2500.0f;
      fused[j] = dest;   // the init branch leaves the measurement as the output (:161-174)
    }
    return;
  }
  const float aa = al, bb = bv;
  if (dest < kFloatEps) {
    fused[j] = ((double)(aa / (aa + bb)) > 0.6) ? mv : 0.0f;
    return;
  }
  const float miu = mv, sq = sv;
  const float new_sq = unc * sq / (unc + sq);
  const float new_miu = (dest * sq + miu * unc) / (unc + sq);
  float c1 = (aa / (aa + bb)) * normpdf(dest, miu, unc + sq);
  float c2 = (bb / (aa + bb)) * 1.0f / 50.0f;
  const float nc = c1 + c2;
  c1 = c1 / nc;
  c2 = c2 / nc;
  const float f = c1 * ((aa + 1.0f) / (aa + bb + 1.0f)) + c2 * (aa / (aa + bb + 1.0f));
  const float e = c1 * (((aa + 1.0f) * (aa + 2.0f)) / ((aa + bb + 1.0f) * (aa + bb + 2.0f))) +
                  c2 * (aa * (aa + 1.0f) / ((aa + bb + 1.0f) * (aa + bb + 2.0f)));
  const float mu_p = c1 * new_miu + c2 * miu;
  const float sig_p = c1 * (new_sq + new_miu * new_miu) + c2 * (sq + miu * miu) - mu_p * mu_p;
  const float a_p = (e - f) / (f - e / f);
  const float b_p = a_p * (1.0f - f) / f;
  ca[j] = a_p; cb[j] = b_p; cm[j] = mu_p; cs[j] = sig_p;
  fused[j] = ((double)(aa / (aa + bb)) > 0.6) ? mu_p : 0.0f;
}

}  // namespace pm

using namespace pm;

extern "C" pm_status pm_fill_f32(float* dst, size_t n, float value, pm_stream stream) {
  clear_error();
  if (!dst && n) return fail(PM_ERR_INVALID_ARGUMENT, "pm_fill_f32: null pointer");
  if (n == 0) return PM_OK;
  fill_f32_kernel<<<(unsigned)((n + 255) / 256), 256, 0, (hipStream_t)stream>>>(dst, (long)n,
                                                                                 value);
  return check_launch("fill_f32_kernel");
}

// The checks and launches of one fusion step; zbuf_next == nullptr empties zbuf first.
static pm_status depth_fuse_impl(const char* who, int32_t width, int32_t height, pm_camera cam,
                                 const float* r_p2c, const float* t_p2c, const float* depth,
                                 float* fused, uint64_t* zbuf, uint64_t* zbuf_next,
                                 int32_t* table, pm_fusion_state pre, pm_fusion_state cur,
                                 hipStream_t st) {
  const std::string w(who);
  if (!r_p2c || !t_p2c || !depth || !fused || !zbuf || !pre.alpha || !pre.beta || !pre.mu ||
      !pre.sig || !cur.alpha || !cur.beta || !cur.mu || !cur.sig)
    return fail(PM_ERR_INVALID_ARGUMENT, w + ": null pointer");
  if (width <= 0 || height <= 0 || width > 65535 || height > 65535)
    return fail(PM_ERR_INVALID_ARGUMENT, w + ": bad image size");
  if (pre.alpha == cur.alpha || pre.beta == cur.beta || pre.mu == cur.mu || pre.sig == cur.sig)
    return fail(PM_ERR_INVALID_ARGUMENT, w + ": pre and cur state must not alias");
  if (zbuf_next == zbuf)
    return fail(PM_ERR_INVALID_ARGUMENT, w + ": zbuf_next must be another buffer than zbuf");
  FuseArgs a{};
  a.W = width; a.H = height;
  a.fx = cam.fx; a.fy = cam.fy; a.cx = cam.cx; a.cy = cam.cy;
  for (int k = 0; k < 9; ++k) a.r[k] = r_p2c[k];
  for (int k = 0; k < 3; ++k) a.t[k] = t_p2c[k];
  if (!(std::isfinite(a.fx) && std::isfinite(a.fy) && std::isfinite(a.cx) &&
        std::isfinite(a.cy) && a.fx != 0.0f && a.fy != 0.0f))
    return fail(PM_ERR_INVALID_ARGUMENT, w + ": bad intrinsics");
  for (int k = 0; k < 9; ++k)
    if (!std::isfinite(a.r[k])) return fail(PM_ERR_INVALID_ARGUMENT, w + ": bad pose");
  for (int k = 0; k < 3; ++k)
    if (!std::isfinite(a.t[k])) return fail(PM_ERR_INVALID_ARGUMENT, w + ": bad pose");
  const long N = (long)width * height;
  pm_status s;
  if (!zbuf_next) {
    fill_u64_kernel<<<(unsigned)((N + 255) / 256), 256, 0, st>>>(zbuf, N, kEmpty);
    s = check_launch("fill_u64_kernel");
    if (s != PM_OK) return s;
  }
  dim3 grid((width + 255) / 256, height);
  fuse_transform_kernel<<<grid, 256, 0, st>>>(a, pre.alpha, pre.beta, pre.mu, pre.sig, zbuf,
                                              zbuf_next);
  s = check_launch("fuse_transform_kernel");
  if (s != PM_OK) return s;
  fuse_update_kernel<<<grid, 256, 0, st>>>(a, zbuf, table, pre.alpha, pre.beta, pre.mu,
                                            pre.sig, cur.alpha, cur.beta, cur.mu, cur.sig,
                                            depth, fused);
  return check_launch("fuse_update_kernel");
}

extern "C" pm_status pm_depth_fuse(int32_t width, int32_t height, pm_camera cam,
                                   const float r_p2c[9], const float t_p2c[3],
                                   const float* depth, float* fused, uint64_t* zbuf,
                                   uint64_t* zbuf_next, int32_t* table, pm_fusion_state pre,
                                   pm_fusion_state cur, pm_stream stream) {
  clear_error();
  return depth_fuse_impl("pm_depth_fuse", width, height, cam, r_p2c, t_p2c, depth, fused, zbuf,
                         zbuf_next, table, pre, cur, (hipStream_t)stream);
}

// ---- allocation-free variant ---------------------------------------------------------------
// The z-buffer of pm_depth_fuse_noalloc: a file-scope device array of PM_FUSE_NOALLOC_MAX_PIXELS
// entries (16 MiB), which every GPU the library runs on holds until the process ends.  In host
// code `scratch` names a host-side stand-in, not device memory: its device address is taken
// once per device with hipGetSymbolAddress and only that address is used.
static __device__ __align__(256) uint64_t scratch[PM_FUSE_NOALLOC_MAX_PIXELS];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
uint64_t* g_scratch_ptr[kMaxDevices] = {};

// The current device's address of `scratch`, looked up on the first call on that device.
pm_status scratch_address(uint64_t** out) {
  int dev = -1;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess || dev < 0 || dev >= kMaxDevices)
    return fail(PM_ERR_DEVICE, std::string("pm_depth_fuse_noalloc: hipGetDevice: ") +
                                   hipGetErrorString(e));
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return fail(PM_ERR_DEVICE, std::string("pm_depth_fuse_noalloc: hipGetSymbolAddress: ") +
                                     hipGetErrorString(e));
    g_scratch_ptr[dev] = static_cast<uint64_t*>(p);
  }
  *out = g_scratch_ptr[dev];
  return PM_OK;
}
}  // namespace

extern "C" pm_status pm_depth_fuse_noalloc(int32_t width, int32_t height, pm_camera cam,
                                           const float r_p2c[9], const float t_p2c[3],
                                           const float* depth, float* fused, int32_t* table,
                                           pm_fusion_state pre, pm_fusion_state cur,
                                           pm_stream stream) {
  clear_error();
  if (width <= 0 || height <= 0 || (long)width * height > PM_FUSE_NOALLOC_MAX_PIXELS)
    return fail(PM_ERR_INVALID_ARGUMENT,
                "pm_depth_fuse_noalloc: width * height must be in 1.." +
                    std::to_string(PM_FUSE_NOALLOC_MAX_PIXELS));
  uint64_t* zbuf = nullptr;
  const pm_status s = scratch_address(&zbuf);
  if (s != PM_OK) return s;
  // zbuf_next = nullptr: the step empties the scratch z-buffer first, with a fill kernel
  return depth_fuse_impl("pm_depth_fuse_noalloc", width, height, cam, r_p2c, t_p2c, depth,
                         fused, zbuf, nullptr, table, pre, cur, (hipStream_t)stream);
}

extern "C" pm_status pm_zbuf_reset(uint64_t* zbuf, size_t n, pm_stream stream) {
  clear_error();
  if (!zbuf && n) return fail(PM_ERR_INVALID_ARGUMENT, "pm_zbuf_reset: null pointer");
  if (n == 0) return PM_OK;
  fill_u64_kernel<<<(unsigned)((n + 255) / 256), 256, 0, (hipStream_t)stream>>>(zbuf, (long)n,
                                                                                 kEmpty);
  return check_launch("fill_u64_kernel");
}

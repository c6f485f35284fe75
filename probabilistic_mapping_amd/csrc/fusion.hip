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

// transformMap for one pixel i = (x, y) of the previous state (depth_fusion.cu:26-49), given
// that pixel's alpha, beta, mu, sigma^2: gate, transform, rewrite mu[i] and sig[i] in place,
// insert into the z-buffer.
__device__ __forceinline__ void transform_pixel(const FuseArgs& a, int x, int y, long i, float al,
                                                float be, float depth, float sg,
                                                float* __restrict__ mu, float* __restrict__ sig,
                                                uint64_t* __restrict__ zbuf) {
  if (al < kFloatEps) return;                                   // :26
  if ((double)(al / (al + be)) < 0.4) return;                   // :30
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
  sig[i] = (float)((double)sg + (double)pz2 * 0.05);
  if (i == 0) return;
  const uint64_t key = ((uint64_t)__float_as_uint(pz2) << 32) | (uint64_t)(uint32_t)i;
  atomicMin((unsigned long long*)(zbuf + (long)y2 * a.W + x2), (unsigned long long)key);
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
  if (al < kFloatEps) return;
  transform_pixel(a, x, y, i, al, beta[i], mu[i], sig[i], mu, sig, zbuf);
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

// The fusion update of pixel j = (x, y) with measurement dest: table entry, prediction and
// depthUpdate, writing the table, the current state and the fused depth.  Returns the state
// pixel j holds afterwards in (fa, fb, fm, fs); where alpha ends below FLOAT_EPS only fa is
// meaningful (the other fields keep stale values there, as in the reference).
__device__ __forceinline__ void update_pixel(
    const FuseArgs& a, int x, int y, long j, float dest, const uint64_t* __restrict__ zbuf,
    int32_t* __restrict__ table, const float* __restrict__ pa, const float* __restrict__ pb,
    const float* __restrict__ pm_, const float* __restrict__ ps, float* __restrict__ ca,
    float* __restrict__ cb, float* __restrict__ cm, float* __restrict__ cs, float* fused,
    float& fa, float& fb, float& fm, float& fs) {
  const int W = a.W, H = a.H;
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
  fa = al; fb = bv; fm = mv; fs = sv;
  // depthUpdate (:150-224)
  const float unc = (float)((double)(dest * dest) * 0.01);
  if (al < kFloatEps) {
    if (dest < kFloatEps) {
      fused[j] = 0.0f;
    } else {
      ca[j] = 10.0f; cb[j] = 10.0f; cm[j] = dest; cs[j] = 2500.0f;
      fa = 10.0f; fb = 10.0f; fm = dest; fs = 2500.0f;
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
  fa = a_p; fb = b_p; fm = mu_p; fs = sig_p;
  fused[j] = ((double)(aa / (aa + bb)) > 0.6) ? mu_p : 0.0f;
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
  if (x >= a.W) return;
  const long j = (long)y * a.W + x;
  const float dest = depth[j];
  float fa, fb, fm, fs;
  update_pixel(a, x, y, j, dest, zbuf, table, pa, pb, pm_, ps, ca, cb, cm, cs, fused, fa, fb, fm,
               fs);
}

// One step of a fusion chain in one launch: (the 3x3 median of this step's depth ->) the
// update of this step -> the forward warp of the state it produced into the next step's
// z-buffer.  The update of pixel j writes the state of pixel j, and the next step's
// transformMap of pixel j reads exactly that state, so the two fuse per pixel without a
// device-wide barrier; the warp rewrites cm/cs in place as fuse_transform_kernel rewrites the
// previous state.  MEDIAN: dest is median3x3_at(depth) (also written to dep_out if given);
// NEXT: warp with the next step's pose `an` into znext, and empty zclear (the z-buffer of the
// step after next; this step's own z-buffer is still being read by neighbouring threads).
template <bool MEDIAN, bool NEXT>
__global__ __launch_bounds__(256) void fuse_step_kernel(
    const FuseArgs a, const FuseArgs an, const uint64_t* __restrict__ zbuf,
    int32_t* __restrict__ table, const float* __restrict__ pa, const float* __restrict__ pb,
    const float* __restrict__ pm_, const float* __restrict__ ps, float* __restrict__ ca,
    float* __restrict__ cb, float* __restrict__ cm, float* __restrict__ cs,
    const float* depth, float* __restrict__ dep_out, float* fused, uint64_t* __restrict__ znext,
    uint64_t* __restrict__ zclear) {
  // without the median, depth and fused may be the same buffer, as in fuse_update_kernel
  const int x = blockIdx.x * blockDim.x + threadIdx.x;
  const int y = blockIdx.y;
  if (x >= a.W) return;
  const long j = (long)y * a.W + x;
  if (NEXT && zclear) zclear[j] = kEmpty;
  const float dest = MEDIAN ? median3x3_at(depth, a.W, a.H, x, y) : depth[j];
  if (MEDIAN && dep_out) dep_out[j] = dest;
  float fa, fb, fm, fs;
  update_pixel(a, x, y, j, dest, zbuf, table, pa, pb, pm_, ps, ca, cb, cm, cs, fused, fa, fb, fm,
               fs);
  if (NEXT) transform_pixel(an, x, y, j, fa, fb, fm, fs, cm, cs, znext);
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

// The kernel arguments of a step with camera cam and previous->current pose (r, t), checked.
static pm_status fuse_args(const std::string& w, int32_t width, int32_t height, pm_camera cam,
                           const float* r, const float* t, FuseArgs& a) {
  if (width <= 0 || height <= 0 || width > 65535 || height > 65535)
    return fail(PM_ERR_INVALID_ARGUMENT, w + ": bad image size");
  a = FuseArgs{};
  a.W = width; a.H = height;
  a.fx = cam.fx; a.fy = cam.fy; a.cx = cam.cx; a.cy = cam.cy;
  if (!(std::isfinite(a.fx) && std::isfinite(a.fy) && std::isfinite(a.cx) &&
        std::isfinite(a.cy) && a.fx != 0.0f && a.fy != 0.0f))
    return fail(PM_ERR_INVALID_ARGUMENT, w + ": bad intrinsics");
  if (!r || !t) return PM_OK;   // no pose: a step that only updates
  for (int k = 0; k < 9; ++k) a.r[k] = r[k];
  for (int k = 0; k < 3; ++k) a.t[k] = t[k];
  for (int k = 0; k < 9; ++k)
    if (!std::isfinite(a.r[k])) return fail(PM_ERR_INVALID_ARGUMENT, w + ": bad pose");
  for (int k = 0; k < 3; ++k)
    if (!std::isfinite(a.t[k])) return fail(PM_ERR_INVALID_ARGUMENT, w + ": bad pose");
  return PM_OK;
}

static bool state_complete(const pm_fusion_state& s) {
  return s.alpha && s.beta && s.mu && s.sig;
}
static bool states_alias(const pm_fusion_state& p, const pm_fusion_state& c) {
  return p.alpha == c.alpha || p.beta == c.beta || p.mu == c.mu || p.sig == c.sig;
}

// The checks and launches of one fusion step; zbuf_next == nullptr empties zbuf first.
static pm_status depth_fuse_impl(const char* who, int32_t width, int32_t height, pm_camera cam,
                                 const float* r_p2c, const float* t_p2c, const float* depth,
                                 float* fused, uint64_t* zbuf, uint64_t* zbuf_next,
                                 int32_t* table, pm_fusion_state pre, pm_fusion_state cur,
                                 hipStream_t st) {
  const std::string w(who);
  if (!r_p2c || !t_p2c || !depth || !fused || !zbuf || !state_complete(pre) ||
      !state_complete(cur))
    return fail(PM_ERR_INVALID_ARGUMENT, w + ": null pointer");
  FuseArgs a;
  pm_status s = fuse_args(w, width, height, cam, r_p2c, t_p2c, a);
  if (s != PM_OK) return s;
  if (states_alias(pre, cur))
    return fail(PM_ERR_INVALID_ARGUMENT, w + ": pre and cur state must not alias");
  if (zbuf_next == zbuf)
    return fail(PM_ERR_INVALID_ARGUMENT, w + ": zbuf_next must be another buffer than zbuf");
  const long N = (long)width * height;
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

extern "C" pm_status pm_depth_fuse_warp(int32_t width, int32_t height, pm_camera cam,
                                        const float r_p2c[9], const float t_p2c[3],
                                        pm_fusion_state state, uint64_t* zbuf,
                                        uint64_t* zbuf_clear, pm_stream stream) {
  clear_error();
  const std::string w = "pm_depth_fuse_warp";
  if (!r_p2c || !t_p2c || !zbuf || !state_complete(state))
    return fail(PM_ERR_INVALID_ARGUMENT, w + ": null pointer");
  if (zbuf_clear == zbuf)
    return fail(PM_ERR_INVALID_ARGUMENT, w + ": zbuf_clear must be another buffer than zbuf");
  FuseArgs a;
  const pm_status s = fuse_args(w, width, height, cam, r_p2c, t_p2c, a);
  if (s != PM_OK) return s;
  dim3 grid((width + 255) / 256, height);
  fuse_transform_kernel<<<grid, 256, 0, (hipStream_t)stream>>>(a, state.alpha, state.beta,
                                                               state.mu, state.sig, zbuf,
                                                               zbuf_clear);
  return check_launch("fuse_transform_kernel");
}

extern "C" pm_status pm_depth_fuse_step(int32_t width, int32_t height, pm_camera cam,
                                        const float* depth, int32_t median, float* depth_out,
                                        float* fused, const uint64_t* zbuf, int32_t* table,
                                        pm_fusion_state pre, pm_fusion_state cur,
                                        const float* r_next, const float* t_next,
                                        uint64_t* zbuf_next, uint64_t* zbuf_clear,
                                        pm_stream stream) {
  clear_error();
  const std::string w = "pm_depth_fuse_step";
  if (!depth || !fused || !zbuf || !state_complete(pre) || !state_complete(cur))
    return fail(PM_ERR_INVALID_ARGUMENT, w + ": null pointer");
  if (states_alias(pre, cur))
    return fail(PM_ERR_INVALID_ARGUMENT, w + ": pre and cur state must not alias");
  if (median && (depth == fused || depth == depth_out || depth_out == fused))
    return fail(PM_ERR_INVALID_ARGUMENT,
                w + ": with the median, depth, depth_out and fused must not alias");
  if (!median && depth_out)
    return fail(PM_ERR_INVALID_ARGUMENT, w + ": depth_out is written only with the median");
  const bool next = r_next || t_next;
  if (next && !(r_next && t_next && zbuf_next))
    return fail(PM_ERR_INVALID_ARGUMENT, w + ": the next step needs r_next, t_next, zbuf_next");
  if (!next && (zbuf_next || zbuf_clear))
    return fail(PM_ERR_INVALID_ARGUMENT, w + ": zbuf_next / zbuf_clear without a next pose");
  if (next && (zbuf_next == zbuf || zbuf_clear == zbuf || zbuf_clear == zbuf_next))
    return fail(PM_ERR_INVALID_ARGUMENT, w + ": zbuf, zbuf_next, zbuf_clear must differ");
  FuseArgs a, an;
  pm_status s = fuse_args(w, width, height, cam, nullptr, nullptr, a);
  if (s != PM_OK) return s;
  if (next) {
    s = fuse_args(w, width, height, cam, r_next, t_next, an);
    if (s != PM_OK) return s;
  } else {
    an = a;
  }
  dim3 grid((width + 255) / 256, height);
  hipStream_t st = (hipStream_t)stream;
#define PM_FUSE_STEP(MED, NX)                                                                 \
  fuse_step_kernel<MED, NX><<<grid, 256, 0, st>>>(a, an, zbuf, table, pre.alpha, pre.beta,     \
                                                  pre.mu, pre.sig, cur.alpha, cur.beta,        \
                                                  cur.mu, cur.sig, depth, depth_out, fused,    \
                                                  zbuf_next, zbuf_clear)
  if (median && next) PM_FUSE_STEP(true, true);
  else if (median) PM_FUSE_STEP(true, false);
  else if (next) PM_FUSE_STEP(false, true);
  else PM_FUSE_STEP(false, false);
#undef PM_FUSE_STEP
  return check_launch("fuse_step_kernel");
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

// Shared device helpers and host-side error plumbing for the gfx950 depth pipeline.
#pragma once

#include <hip/hip_bf16.h>
#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>

#include "pm_capi.h"

namespace pm {

constexpr float kFloatEps = 1e-5f;   // FLOAT_EPS, motion_stereo/parameters.h:6
constexpr float kDepInf = 1000.0f;   // DEP_INF, parameters.h:28
constexpr float kCostInf = 0.0f;     // COST_INF, parameters.h:29
constexpr int kWave = 64;
// Dword 3 of a raw buffer resource on gfx9 (32-bit data format; no swizzle, no index stride).
constexpr int kBufferDword3 = 0x00020000;

// Scanline directions in aggregation order: 0-3 are the reference's paths
// (sgm.cu:119-131), 4-7 the diagonal extension.
constexpr int kPathDx[8] = {1, -1, 0, 0, 1, -1, 1, -1};
constexpr int kPathDy[8] = {0, 0, 1, -1, 1, -1, -1, 1};
// The same tables packed two bits per path (value + 1), for device code indexing them with a
// run-time path (a local array would live in scratch memory).
__host__ __device__ inline int path_dx(int p) { return ((0x2252 >> (2 * p)) & 3) - 1; }
__host__ __device__ inline int path_dy(int p) { return ((0x8225 >> (2 * p)) & 3) - 1; }

// ---- per-keyframe stage arguments -------------------------------------------------------
// Plane-sweep cost inputs of one keyframe (passed by value to the single-keyframe kernels,
// read from a device table by the batched ones).
// Sources a TY-row cost tile can hold the per-plane warp records of (16 floats per source in
// the plane's TY x 64-float output slot).
__host__ __device__ constexpr int cost_tile_max_sources(int TY) { return TY * 64 / 16; }

struct CostArgs {
  int W, H, D, S;
  float idep_min, dep_sample;
  const void* ref;                        // uint8 [H][W], or float [H][W] (differentiable form)
  const uint64_t* quad[PM_MAX_SOURCES];   // difference texels, see pm_make_quad
  float hm[PM_MAX_SOURCES][12];   // h11..h33, t1..t3
};

// SGM inputs of one keyframe and one path direction.
struct SgmArgs {
  int W, H, D, dx, dy;
  int nlines;
  const void* ref;  // uint8 [H][W], or float [H][W] (differentiable form)
  float P1n, P2n;   // penalties where |I(p) - I(q)| >= tau_so
  float P1s, P2s;   // penalties scaled by q1, q2 where |I(p) - I(q)| < tau_so
  int tau_i;        // integer form of the test: |dI| < tau_so  <=>  |dI| < ceil(tau_so)
  float tau_f;      // tau_so itself, for float reference intensities
  float* lout;      // per-path L volume [H][W][D] (differentiable form), else unused
  uint16_t* eout;   // per-path excess volume [H][W][D] (compact form), else unused
  float eq_inv;     // excess quantisation: q = rint(E * eq_inv), E in [0, max(P2n, P2s)]
};

// Compact per-path storage of the SGM excess E = L - C' (0 <= E <= P2 on every step): 16-bit
// fixed point with step max(P2n, P2s) / kExcessMaxQ.
constexpr uint32_t kExcessMaxQ = 0xFFFEu;
// The level of an excess e >= 0: e / step rounded to the nearest integer (ties to even), at
// most kExcessMaxQ, in the low 16 bits of the result (the high bits hold 0x4B00).  One fma puts
// e * inv + 2^23 in a float whose unit in the last place is 1, so its mantissa holds the rounded
// quotient (exact for quotients below 2^23); the clamp works on that float's bit pattern.
__device__ __forceinline__ uint32_t excess_level_bits(float e, float inv) {
  return min(0x4B000000u | kExcessMaxQ, __float_as_uint(fmaf(e, inv, 8388608.0f)));
}
__host__ __device__ inline float excess_step(float P2n, float P2s) {
  const float m = P2n > P2s ? P2n : P2s;
  return m > 0.0f ? m / (float)kExcessMaxQ : 0.0f;
}

// One keyframe of a batch (device table entry).
struct KfDesc {
  CostArgs c;
  void* cost;        // [H][W][D] of the batch's cost type
  float* agg;        // [H][W][D]
  float* depth;      // [H][W]
  float P1n, P2n, P1s, P2s;
  int tau_i, paths;
  float var_scale;
  uint16_t* ebuf;    // compact form: paths x [H][W][D] excess volumes; sweep form: 3 volumes
  float estep;       // excess_step(P2n, P2s)
  float einv;        // 1 / estep (0 when estep is 0)
  float* sweep;      // sweep form: two halves of 3 x [W][D] row-boundary path states
};

// Sweep form of the SGM pass (batched): row blocks of kSweepRows rows per launch; a workgroup
// of kSweepWaves waves owns kSweepBand columns and recomputes kSweepRows - 1 halo columns on
// each side at a block's first row, shrinking by one per row (the diagonal paths' reach).
constexpr int kSweepRows = 16;
constexpr int kSweepBand = 64;
constexpr int kSweepWaves = 16;
constexpr int kSweepCols = (kSweepBand + 2 * (kSweepRows - 1) + kSweepWaves - 1) / kSweepWaves;
__host__ __device__ inline int sweep_bands(int W) { return (W + kSweepBand - 1) / kSweepBand; }
__host__ __device__ inline int sweep_blocks(int H) { return (H + kSweepRows - 1) / kSweepRows; }

__host__ __device__ inline int sgm_lines(int W, int H, int dx, int dy) {
  return (dx != 0 && dy != 0) ? W + H - 1 : (dx != 0 ? H : W);
}

// Index of the keyframe owning work item b: the largest k with first[k] <= b
// (first[] ascending, first[0] = 0; nkf <= 64).
__device__ __forceinline__ int owner_of(const int* __restrict__ first, int nkf, int b) {
  int lo = 0, hi = nkf - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (first[mid] <= b) lo = mid; else hi = mid - 1;
  }
  return lo;
}

// ---- batched launches (defined with their kernels; used by the batch host code) --------
int cost_items(int W, int H, int tile_rows);
bool launch_cost_batch(pm_cost_type ct, int tile_rows, int stage_planes, const KfDesc* kfs,
                       const int* first, const int* npg, int nkf, const int* xr, int grid,
                       hipStream_t st);
bool launch_sgm_batch_path(pm_cost_type ct, int prefetch, int nv, bool full, const KfDesc* kfs,
                           const int* kf, const int* first, int n, int total, int grid,
                           bool first_path, int dx, int dy, hipStream_t st);
int filter_lanes(int D);
int filter_items(int W, int H, int D);
void launch_filter_batch(int LP, const KfDesc* kfs, const int* kf, const int* first, int n,
                         int total, int grid, hipStream_t st);
// Sum of the per-path L volumes (path order) -> agg (skipped if NULL) -> filter -> depth.
void launch_combine_filter(int W, int H, int D, int paths, const float* lbuf, float* agg,
                           float var_scale, float idep_min, float dep_sample, float* depth,
                           hipStream_t st);

// Compact form: agg = P C + step sum_r q_r (0 at invalid pixels) -> agg (skipped if NULL)
// -> filter -> depth.
void launch_combine_excess(pm_cost_type ct, int W, int H, int D, int paths, const void* cost,
                           const uint16_t* ebuf, float step, float* agg, float var_scale,
                           float idep_min, float dep_sample, float* depth, hipStream_t st);
void launch_combine_excess_batch(pm_cost_type ct, int LP, const KfDesc* kfs, const int* kf,
                                 const int* first, int n, int total, hipStream_t st);
// Sweep form: row block rb of the downward (up = false) or upward sweep over the keyframes
// kf[j] of one lane class, keyframe kf[j] owning workgroups first[j] .. first[j+1]-1.
bool launch_sgm_sweep_batch(pm_cost_type ct, int nv, bool full, bool up, const KfDesc* kfs,
                            const int* kf, const int* first, int n, int total, int rb,
                            hipStream_t st);
// Compact form, four scanlines per wave: 16-bit costs and D = 32, 64, 96 or 128 (nv = D / 16).
bool rows_supported(pm_cost_type ct, int D);
int rows_nv(int D);
bool launch_sgm_excess_rows_batch(pm_cost_type ct, int prefetch, int nv, const KfDesc* kfs,
                                  const int* seg, const int* first, int nseg, int total,
                                  int grid, hipStream_t st);
bool launch_sgm_excess_batch(pm_cost_type ct, int prefetch, int nv, bool full,
                             const KfDesc* kfs, const int* seg, const int* first, int nseg,
                             int total, int grid, hipStream_t st);

// ---- host error plumbing -------------------------------------------------------------
void set_error(const std::string& msg);
void clear_error();
pm_status fail(pm_status code, const std::string& msg);
// Checks the last launch on this thread; returns PM_OK or PM_ERR_LAUNCH with a message.
pm_status check_launch(const char* what);

// FP64 plane-sweep homography H = K R_r^T R_l K^-1, t = K R_r^T (T_l - T_r)
// (stereo_mapper.cpp:79-80).  Returns false for a singular K or non-finite input.
bool homography_fp64(const double K[9], const double R_l[9], const double T_l[3],
                     const double R_r[9], const double T_r[3], double H[9], double t[3]);

// Global-memory views of device pointers that arrive as generic pointers (e.g. read from a
// descriptor table): loads and stores through them use the scalar-base + 32-bit vector offset
// global instructions instead of flat ones with 64-bit vector addresses.
template <typename T>
using gptr = T __attribute__((address_space(1)))*;
template <typename T>
__device__ __forceinline__ gptr<T> to_global(T* p) {
  return (gptr<T>)p;
}

typedef uint32_t u32x2 __attribute__((ext_vector_type(2)));
typedef uint32_t u32x3 __attribute__((ext_vector_type(3)));
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));

// A device array read and written through a raw buffer resource: the base and size live in
// scalar registers, an access is a 32-bit per-lane byte offset plus a 32-bit wave-uniform one
// (no per-lane 64-bit address arithmetic), and the hardware range check turns an access past
// `bytes` into a zero read / dropped write.  Arrays must be smaller than 4 GiB.
struct Buf {
  __amdgpu_buffer_rsrc_t r;
  __device__ __forceinline__ Buf(const void* p, uint32_t bytes)
      : r(__builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(p), 0, (int)bytes,
                                            kBufferDword3)) {}
  __device__ __forceinline__ float f32(uint32_t voff, uint32_t soff) const {
    return __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(r, (int)voff, (int)soff, 0));
  }
  __device__ __forceinline__ uint32_t u16(uint32_t voff, uint32_t soff) const {
    return __builtin_amdgcn_raw_buffer_load_b16(r, (int)voff, (int)soff, 0);
  }
  __device__ __forceinline__ void put_f32(float v, uint32_t voff, uint32_t soff) const {
    __builtin_amdgcn_raw_buffer_store_b32(__float_as_uint(v), r, (int)voff, (int)soff, 0);
  }
  // non-temporal store (cache policy "nt"), for data that is read once, much later
  __device__ __forceinline__ void put_f32_nt(float v, uint32_t voff, uint32_t soff) const {
    __builtin_amdgcn_raw_buffer_store_b32(__float_as_uint(v), r, (int)voff, (int)soff, 2);
  }
  __device__ __forceinline__ void put_u16(uint32_t v, uint32_t voff, uint32_t soff) const {
    __builtin_amdgcn_raw_buffer_store_b16((unsigned short)v, r, (int)voff, (int)soff, 0);
  }
  __device__ __forceinline__ void put_u16_nt(uint32_t v, uint32_t voff, uint32_t soff) const {
    __builtin_amdgcn_raw_buffer_store_b16((unsigned short)v, r, (int)voff, (int)soff, 2);
  }
  // 8-byte stores (voff + soff a multiple of 8); AUX 2 = non-temporal
  template <int AUX>
  __device__ __forceinline__ void put_b64(u32x2 v, uint32_t voff, uint32_t soff) const {
    __builtin_amdgcn_raw_buffer_store_b64(v, r, (int)voff, (int)soff, AUX);
  }
  // Stores of more than 8 bytes.  A VALU write to such a store's data registers less than two
  // wait states after the store can replace data the store has not read yet; compiled from the
  // store builtins, the scheduler placed such writes directly behind the store, and a few
  // dwords of the last lanes of a row came out wrong.  These stores are therefore written as
  // asm with their own wait states: two after the store for its data, and five before it,
  // since its scalar offset may come from a v_readfirstlane just ahead of the asm, and a
  // vector memory instruction reading an SGPR that a VALU instruction wrote needs five (the
  // compiler inserts neither inside an asm statement).  AUX 2 = non-temporal.  soff is
  // wave-uniform, as for every Buf access.
  template <int AUX>
  __device__ __forceinline__ void put_b128(u32x4 v, uint32_t voff, uint32_t soff) const {
    static_assert(AUX == 0 || AUX == 2, "default or non-temporal policy");
    if constexpr (AUX == 2)
      asm volatile("s_nop 4\n\tbuffer_store_dwordx4 %0, %1, %2, %3 offen nt\n\ts_nop 1"
                   :: "v"(v), "v"(voff), "s"(r), "s"(__builtin_amdgcn_readfirstlane(soff))
                   : "memory");
    else
      asm volatile("s_nop 4\n\tbuffer_store_dwordx4 %0, %1, %2, %3 offen\n\ts_nop 1"
                   :: "v"(v), "v"(voff), "s"(r), "s"(__builtin_amdgcn_readfirstlane(soff))
                   : "memory");
  }
  template <int AUX>
  __device__ __forceinline__ void put_b96(u32x3 v, uint32_t voff, uint32_t soff) const {
    static_assert(AUX == 0 || AUX == 2, "default or non-temporal policy");
    if constexpr (AUX == 2)
      asm volatile("s_nop 4\n\tbuffer_store_dwordx3 %0, %1, %2, %3 offen nt\n\ts_nop 1"
                   :: "v"(v), "v"(voff), "s"(r), "s"(__builtin_amdgcn_readfirstlane(soff))
                   : "memory");
    else
      asm volatile("s_nop 4\n\tbuffer_store_dwordx3 %0, %1, %2, %3 offen\n\ts_nop 1"
                   :: "v"(v), "v"(voff), "s"(r), "s"(__builtin_amdgcn_readfirstlane(soff))
                   : "memory");
  }
  template <int AUX>
  __device__ __forceinline__ void put_b32(uint32_t v, uint32_t voff, uint32_t soff) const {
    __builtin_amdgcn_raw_buffer_store_b32(v, r, (int)voff, (int)soff, AUX);
  }
};
// A cost element through a Buf (byte offsets).
template <typename CT>
__device__ __forceinline__ float buf_cost(const Buf& b, uint32_t voff, uint32_t soff);
template <>
__device__ __forceinline__ float buf_cost<float>(const Buf& b, uint32_t voff, uint32_t soff) {
  return b.f32(voff, soff);
}
template <>
__device__ __forceinline__ float buf_cost<__half>(const Buf& b, uint32_t voff, uint32_t soff) {
  return __half2float(__ushort_as_half((unsigned short)b.u16(voff, soff)));
}
template <>
__device__ __forceinline__ float buf_cost<__hip_bfloat16>(const Buf& b, uint32_t voff,
                                                         uint32_t soff) {
  return __uint_as_float(b.u16(voff, soff) << 16);
}

// A wave-uniform float read from memory, moved to a scalar register.
__device__ __forceinline__ float uniform(float v) {
  return __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(v)));
}

// ---- depth maps -------------------------------------------------------------------------
__device__ __forceinline__ bool valid_depth(float z) { return z > 0.0f && z < kDepInf; }

// The 3x3 median filter at pixel (x, y): an invalid centre is kept as it is; else the lower
// median of the valid depths of the window, all in registers.  Absent or invalid neighbours
// become +inf, an odd-even transposition network (9 rounds, which sorts any 9 values) orders
// the window, and the median is selected from the sorted values by the valid count.  Depths
// and +inf are positive, so their bit patterns order like unsigned integers and the
// compare-exchanges are integer min / max.
__device__ __forceinline__ float median3x3_at(const float* __restrict__ in, int W, int H, int x,
                                              int y) {
  const float z = in[(long)y * W + x];
  if (!valid_depth(z)) return z;
  constexpr uint32_t kAbsent = 0x7F800000u;   // +inf
  uint32_t v[9];
  int n = 0;
#pragma unroll
  for (int j = -1; j <= 1; ++j)
#pragma unroll
    for (int i = -1; i <= 1; ++i) {
      const int xx = x + i, yy = y + j;
      const bool inside = xx >= 0 && yy >= 0 && xx < W && yy < H;
      const float w = inside ? in[(long)yy * W + xx] : 0.0f;
      const bool ok = inside && valid_depth(w);
      v[(j + 1) * 3 + (i + 1)] = ok ? __float_as_uint(w) : kAbsent;
      n += ok ? 1 : 0;
    }
#pragma unroll
  for (int round = 0; round < 9; ++round)
#pragma unroll
    for (int a = round & 1; a + 1 < 9; a += 2) {
      const uint32_t lo = min(v[a], v[a + 1]), hi = max(v[a], v[a + 1]);
      v[a] = lo;
      v[a + 1] = hi;
    }
  const int k = (n - 1) / 2;   // 0..4
  uint32_t med = v[0];
#pragma unroll
  for (int i = 1; i <= 4; ++i) med = k == i ? v[i] : med;
  return __uint_as_float(med);
}

// ---- cost storage ---------------------------------------------------------------------
template <typename CT>
__device__ __forceinline__ float load_cost(gptr<const CT> p);
template <>
__device__ __forceinline__ float load_cost<float>(gptr<const float> p) { return *p; }
template <>
__device__ __forceinline__ float load_cost<__half>(gptr<const __half> p) {
  return __half2float(__ushort_as_half(*(gptr<const unsigned short>)p));
}
template <>
__device__ __forceinline__ float load_cost<__hip_bfloat16>(gptr<const __hip_bfloat16> p) {
  return __uint_as_float((uint32_t)(*(gptr<const unsigned short>)p) << 16);
}

template <typename CT>
__device__ __forceinline__ float load_cost(const CT* p);
template <>
__device__ __forceinline__ float load_cost<float>(const float* p) { return *p; }
template <>
__device__ __forceinline__ float load_cost<__half>(const __half* p) { return __half2float(*p); }
template <>
__device__ __forceinline__ float load_cost<__hip_bfloat16>(const __hip_bfloat16* p) {
  return __bfloat162float(*p);
}

template <typename CT>
__device__ __forceinline__ CT store_cost(float v);
template <>
__device__ __forceinline__ float store_cost<float>(float v) { return v; }
template <>
__device__ __forceinline__ __half store_cost<__half>(float v) { return __float2half_rn(v); }
template <>
__device__ __forceinline__ __hip_bfloat16 store_cost<__hip_bfloat16>(float v) {
  return __float2bfloat16(v);
}

// ---- cross-lane helpers (wave64, DPP) ---------------------------------------------------
// Values handled here are non-negative floats (or +inf), whose IEEE bit patterns order like
// unsigned integers; integer min avoids the NaN-canonicalising moves fminf needs.
__device__ __forceinline__ float umin(float a, float b) {
  return __uint_as_float(min(__float_as_uint(a), __float_as_uint(b)));
}

template <int CTRL>
__device__ __forceinline__ uint32_t dpp_u32(uint32_t old, uint32_t src) {
  return (uint32_t)__builtin_amdgcn_update_dpp((int)old, (int)src, CTRL, 0xF, 0xF, false);
}

// lane i receives lane i-1's value; lane 0 receives `fill`
__device__ __forceinline__ float wave_shr1(float v, float fill) {
  return __uint_as_float(dpp_u32<0x138>(__float_as_uint(fill), __float_as_uint(v)));
}
// lane i receives lane i+1's value; lane 63 receives `fill`
__device__ __forceinline__ float wave_shl1(float v, float fill) {
  return __uint_as_float(dpp_u32<0x130>(__float_as_uint(fill), __float_as_uint(v)));
}

// Lane shifts whose edge lane receives 0; the compiler folds these into the consuming VALU
// instruction as a DPP source operand.
__device__ __forceinline__ float wave_shr1z(float v) {
  return __uint_as_float((uint32_t)__builtin_amdgcn_update_dpp(0, (int)__float_as_uint(v), 0x138,
                                                               0xF, 0xF, true));
}
__device__ __forceinline__ float wave_shl1z(float v) {
  return __uint_as_float((uint32_t)__builtin_amdgcn_update_dpp(0, (int)__float_as_uint(v), 0x130,
                                                               0xF, 0xF, true));
}

// Minimum over the 64 lanes of a non-negative float, wave-uniform (a scalar register).
// Inside each row of 16 lanes: four DPP stages (quad_perm [1,0,3,2], [2,3,0,1], row_ror:4,
// row_ror:8) that fold into v_min_u32_dpp (every source lane exists, so bound_ctrl changes
// nothing).  Across rows: row_bcast:15 (rows 1 and 3 take the previous row's minimum) and
// row_bcast:31 (rows 2 and 3 take row 1's), leaving the minimum of all 64 lanes in lane 63,
// which a readlane moves to a scalar register.  The broadcasts' DPP "old" value is 0xFFFFFFFF,
// the identity of the unsigned minimum, so they too fold into v_min_u32_dpp: seven VALU
// instructions.
template <int CTRL>
__device__ __forceinline__ uint32_t umin_dpp(uint32_t u) {
  return min(u, (uint32_t)__builtin_amdgcn_mov_dpp((int)u, CTRL, 0xF, 0xF, true));
}
template <int CTRL, int ROWS>
__device__ __forceinline__ uint32_t umin_bcast(uint32_t u) {
  return min(u, (uint32_t)__builtin_amdgcn_update_dpp(-1, (int)u, CTRL, ROWS, 0xF, false));
}
__device__ __forceinline__ float wave_min_nonneg(float v) {
  uint32_t u = __float_as_uint(v);
  u = umin_dpp<0xB1>(u);
  u = umin_dpp<0x4E>(u);
  u = umin_dpp<0x124>(u);
  u = umin_dpp<0x128>(u);
  u = umin_bcast<0x142, 0xA>(u);   // row_bcast:15 into rows 1 and 3
  u = umin_bcast<0x143, 0xC>(u);   // row_bcast:31 into rows 2 and 3
  return __uint_as_float(__builtin_amdgcn_readlane(u, 63));
}

// min(lane i-1's `lo`, lane i+1's `hi`) for the neighbouring planes of the SGM recurrence; a
// missing neighbour (lane 0's left, lane 63's right) is 0xFFFFFFFF, above the bit pattern of
// every non-negative float, so it never wins.
__device__ __forceinline__ uint32_t shr1_max(float lo) {
  return (uint32_t)__builtin_amdgcn_update_dpp(-1, (int)__float_as_uint(lo), 0x138, 0xF, 0xF,
                                               false);
}
// shr1_max carried across calls: `prev` is the previous result.  Lane 0 has no source lane and
// is not written, so it keeps the 0xFFFFFFFF it was initialised with, and the per-call move of
// the fill value into the destination disappears.
__device__ __forceinline__ uint32_t shr1_max_carry(uint32_t prev, float lo) {
  return (uint32_t)__builtin_amdgcn_update_dpp((int)prev, (int)__float_as_uint(lo), 0x138, 0xF,
                                               0xF, false);
}
__device__ __forceinline__ uint32_t shl1_max(float hi) {
  return (uint32_t)__builtin_amdgcn_update_dpp(-1, (int)__float_as_uint(hi), 0x130, 0xF, 0xF,
                                               false);
}


// Workgroups are dispatched round-robin over the 8 XCDs (block i runs on XCD i % 8).  This
// remaps the launch index so that each XCD receives one contiguous range of logical blocks,
// and neighbouring tiles share that XCD's L2.  A bijection on [0, nblocks).
__device__ __forceinline__ int xcd_block(int b, int nblocks) {
  constexpr int kXcd = 8;
  const int per = nblocks / kXcd, rem = nblocks % kXcd;
  const int xcd = b % kXcd, idx = b / kXcd;
  return xcd < rem ? xcd * (per + 1) + idx : rem * (per + 1) + (xcd - rem) * per + idx;
}

// Reference intensity at an integer position, 0 outside the image (border addressing).
template <typename RT>
__device__ __forceinline__ float texel(const RT* img, int W, int H, int x, int y) {
  return (x < 0 || y < 0 || x >= W || y >= H) ? 0.0f : (float)img[(long)y * W + x];
}

}  // namespace pm

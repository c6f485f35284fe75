// Winner-take-all with rejection rules and parabola subpixel fit, and the optional 3x3 median.
//
// Reference: filterCostKernel + filter_cost, motion_stereo/filter_cost.cu:3-70.  The reference
// launches one 64-thread block per pixel and reduces (value, index) pairs through shared memory
// with log2(D) barriers.  Here a pixel's D aggregate values are read by a group of LP lanes,
// four consecutive planes per lane (16-byte loads), so a wave reads 64/LP whole pixels as one
// contiguous stream; the argmin is reduced across the group with lane shuffles and ties go to
// the lowest plane index.
#include "common.hpp"

namespace pm {

// Filter of the pixels of one wave; load(pc, d0, v) fills v[0..3] with planes d0 .. d0+3 of
// pixel pc (+inf past D).
template <int LP, typename Load>
__device__ __forceinline__ void filter_pixels_impl(long N, int D, Load&& load, float var_scale,
                                                   float idep_min, float dep_sample,
                                                   float* __restrict__ depth, long wave) {
  const int lane = threadIdx.x & 63;
  const int grp = lane / LP, k = lane % LP;
  const long p = wave * (64 / LP) + grp;
  const long pc = p < N ? p : N - 1;   // out-of-range groups re-read the last pixel
  const int d0 = 4 * k;
  float v[4];
  load(pc, p < N, d0, v);
  float best = v[0];
  int mi = d0;
#pragma unroll
  for (int j = 1; j < 4; ++j)
    if (v[j] < best) { best = v[j]; mi = d0 + j; }
#pragma unroll
  for (int off = LP / 2; off >= 1; off >>= 1) {
    const float ov = __shfl_xor(best, off, 64);
    const int oi = __shfl_xor(mi, off, 64);
    if (ov < best || (ov == best && oi < mi)) { best = ov; mi = oi; }
  }
  // neighbours of the minimum, fetched from the lanes that hold them
  const int dm = mi > 0 ? mi - 1 : 0, dp = mi + 1 < D ? mi + 1 : D - 1;
  const float sm = (dm & 3) == 0 ? v[0] : (dm & 3) == 1 ? v[1] : (dm & 3) == 2 ? v[2] : v[3];
  const float sp = (dp & 3) == 0 ? v[0] : (dp & 3) == 1 ? v[1] : (dp & 3) == 2 ? v[2] : v[3];
  const float cm = __shfl(sm, grp * LP + dm / 4, 64);
  const float cp = __shfl(sp, grp * LP + dp / 4, 64);
  float out;
  if ((double)best < 1e-5 && (double)best > -1e-5) {          // rule 1, filter_cost.cu:37-38
    out = kDepInf;
  } else if (mi == 0 || mi == D - 1) {                        // rule 2, :39-40
    out = 0.0f;
  } else if (cm + cp < 2.0f * best * var_scale) {             // rule 3, :41-42
    out = 0.0f;
  } else {                                                    // parabola, :45-51
    const float A = cm - 2.0f * best + cp;
    const float Bq = -cm + cp;
    const float s = (float)mi - Bq / (2.0f * A);
    out = 1.0f / (idep_min + s * dep_sample);
  }
  if (k == 0 && p < N) depth[p] = out;
}

template <int LP>
__device__ __forceinline__ void filter_pixels(long N, int D, bool vec4,
                                              const float* __restrict__ agg, float var_scale,
                                              float idep_min, float dep_sample,
                                              float* __restrict__ depth, long wave) {
  const float INF = __uint_as_float(0x7F800000u);
  auto load = [&](long pc, bool, int d0, float (&v)[4]) {
    const float* c = agg + pc * D;
    if (vec4 && d0 + 3 < D) {
      const float4 t = *reinterpret_cast<const float4*>(c + d0);
      v[0] = t.x; v[1] = t.y; v[2] = t.z; v[3] = t.w;
    } else {
#pragma unroll
      for (int j = 0; j < 4; ++j) v[j] = d0 + j < D ? c[d0 + j] : INF;
    }
  };
  filter_pixels_impl<LP>(N, D, load, var_scale, idep_min, dep_sample, depth, wave);
}

// Aggregate = sum of the per-path volumes lbuf[r] (r = 0 .. paths-1, each [H][W][D], 0 at
// invalid pixels), added in path order exactly as the sequential read-modify-write passes add
// them (sgm.cu:101-108), stored to agg, then filtered as above.
template <int LP>
__global__ __launch_bounds__(256) void combine_filter_kernel(long N, int D, int paths, bool vec4,
                                                             const float* __restrict__ lbuf,
                                                             long V, float* __restrict__ agg,
                                                             float var_scale, float idep_min,
                                                             float dep_sample,
                                                             float* __restrict__ depth) {
  const float INF = __uint_as_float(0x7F800000u);
  const long wave = (long)blockIdx.x * 4 + (threadIdx.x >> 6);
  auto load = [&](long pc, bool own, int d0, float (&v)[4]) {
    const long e = pc * D + d0;
    if (vec4 && d0 + 3 < D) {
      // the path volumes are read exactly once: non-temporal loads.  For 8 and 4 paths all
      // loads are issued before the sum (a loop over a run-time count waits for each load
      // before issuing the next); the sum keeps the path order.
      typedef float vf4 __attribute__((ext_vector_type(4)));
      auto ld = [&](int r) {
        return __builtin_nontemporal_load(reinterpret_cast<const vf4*>(lbuf + (long)r * V + e));
      };
      vf4 u[8];
      if (paths == 8) {
#pragma unroll
        for (int r = 0; r < 8; ++r) u[r] = ld(r);
      } else if (paths == 4) {
#pragma unroll
        for (int r = 0; r < 4; ++r) u[r] = ld(r);
      } else {
        u[0] = ld(0);
      }
      float4 t = make_float4(u[0].x, u[0].y, u[0].z, u[0].w);
      if (paths == 8 || paths == 4) {
#pragma unroll
        for (int r = 1; r < 8; ++r) {
          if (r < paths) { t.x += u[r].x; t.y += u[r].y; t.z += u[r].z; t.w += u[r].w; }
        }
      } else {
        for (int r = 1; r < paths; ++r) {
          const vf4 w = ld(r);
          t.x += w.x; t.y += w.y; t.z += w.z; t.w += w.w;
        }
      }
      if (own && agg) *reinterpret_cast<float4*>(agg + e) = t;
      v[0] = t.x; v[1] = t.y; v[2] = t.z; v[3] = t.w;
    } else {
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        if (d0 + j < D) {
          float a = lbuf[e + j];
          for (int r = 1; r < paths; ++r) a += lbuf[(long)r * V + e + j];
          if (own && agg) agg[e + j] = a;
          v[j] = a;
        } else {
          v[j] = INF;
        }
      }
    }
  };
  filter_pixels_impl<LP>(N, D, load, var_scale, idep_min, dep_sample, depth, wave);
}

void launch_combine_filter(int W, int H, int D, int paths, const float* lbuf, float* agg,
                           float var_scale, float idep_min, float dep_sample, float* depth,
                           hipStream_t st) {
  const long N = (long)W * H;
  const long V = N * D;
  const bool vec4 = (D % 4 == 0) && (agg == nullptr || (uintptr_t)agg % 16 == 0);
  const int LP = filter_lanes(D);
  const unsigned blocks = (unsigned)filter_items(W, H, D);
  switch (LP) {
    case 4: combine_filter_kernel<4><<<blocks, 256, 0, st>>>(N, D, paths, vec4, lbuf, V, agg, var_scale, idep_min, dep_sample, depth); break;
    case 8: combine_filter_kernel<8><<<blocks, 256, 0, st>>>(N, D, paths, vec4, lbuf, V, agg, var_scale, idep_min, dep_sample, depth); break;
    case 16: combine_filter_kernel<16><<<blocks, 256, 0, st>>>(N, D, paths, vec4, lbuf, V, agg, var_scale, idep_min, dep_sample, depth); break;
    case 32: combine_filter_kernel<32><<<blocks, 256, 0, st>>>(N, D, paths, vec4, lbuf, V, agg, var_scale, idep_min, dep_sample, depth); break;
    default: combine_filter_kernel<64><<<blocks, 256, 0, st>>>(N, D, paths, vec4, lbuf, V, agg, var_scale, idep_min, dep_sample, depth); break;
  }
}

// Four consecutive planes (d0 % 4 == 0, 4-element aligned) of a cost volume, as floats.
template <typename CT>
__device__ __forceinline__ void load_cost4(const CT* p, float (&v)[4]);
template <>
__device__ __forceinline__ void load_cost4<float>(const float* p, float (&v)[4]) {
  const float4 t = *reinterpret_cast<const float4*>(p);
  v[0] = t.x; v[1] = t.y; v[2] = t.z; v[3] = t.w;
}
template <>
__device__ __forceinline__ void load_cost4<__half>(const __half* p, float (&v)[4]) {
  // the combine is the cost volume's last reader: non-temporal
  typedef uint32_t vu2 __attribute__((ext_vector_type(2)));
  const vu2 t = __builtin_nontemporal_load(reinterpret_cast<const vu2*>(p));
  v[0] = __half2float(__ushort_as_half((unsigned short)(t.x & 0xFFFFu)));
  v[1] = __half2float(__ushort_as_half((unsigned short)(t.x >> 16)));
  v[2] = __half2float(__ushort_as_half((unsigned short)(t.y & 0xFFFFu)));
  v[3] = __half2float(__ushort_as_half((unsigned short)(t.y >> 16)));
}
template <>
__device__ __forceinline__ void load_cost4<__hip_bfloat16>(const __hip_bfloat16* p,
                                                           float (&v)[4]) {
  const uint2 t = *reinterpret_cast<const uint2*>(p);
  v[0] = __uint_as_float(t.x << 16);
  v[1] = __uint_as_float(t.x & 0xFFFF0000u);
  v[2] = __uint_as_float(t.y << 16);
  v[3] = __uint_as_float(t.y & 0xFFFF0000u);
}

// Compact form: the aggregate is rebuilt from the cost and the per-path excess volumes,
//   agg(p,d) = P C(p,d) + step * sum_r q_r(p,d)   (sum_r L_r = sum_r (C' + E_r), sgm.cu:98).
// The sum of the 16-bit levels is exact in 32-bit integers, so the only rounding beyond the
// levels themselves is one fma.  Returns whether any of the lane's planes has a negative cost
// (the pixel is then invalid and its aggregate 0, sgm.cu:33-38, 101-104).
template <typename CT>
__device__ __forceinline__ bool excess_agg4(const CT* __restrict__ cost,
                                            const uint16_t* __restrict__ ebuf, long V, long e,
                                            int D, int d0, int paths, bool vec4, float step,
                                            float (&v)[4]) {
  const float INF = __uint_as_float(0x7F800000u);
  bool neg = false;
  if (vec4 && d0 + 3 < D) {
    float c[4];
    load_cost4<CT>(cost + e, c);
    uint32_t s[4] = {0u, 0u, 0u, 0u};
    // the excess volumes are read exactly once: non-temporal loads.  For 8 and 4 paths all
    // loads are issued before the first is used (a loop over a run-time count waits for each
    // one before issuing the next); the integer sum does not depend on the order.  Each path
    // volume is read through a buffer resource of its own (base and size in scalar registers),
    // so a lane's only address is its 32-bit byte offset e * 2 (e < V < 2^30), shared by all
    // paths, instead of a 64-bit address per path.
    auto add = [&](u32x2 u) {
      const uint32_t q[4] = {u.x & 0xFFFFu, u.x >> 16, u.y & 0xFFFFu, u.y >> 16};
#pragma unroll
      for (int j = 0; j < 4; ++j) s[j] += q[j];
    };
    const uint32_t eb = (uint32_t)e * 2u;
    auto ld = [&](int r) {
      const Buf b(ebuf + (long)r * V, (uint32_t)V * 2u);
      return __builtin_amdgcn_raw_buffer_load_b64(b.r, (int)eb, 0, 2);   // non-temporal
    };
    if (paths == 8) {
      u32x2 u[8];
#pragma unroll
      for (int r = 0; r < 8; ++r) u[r] = ld(r);
#pragma unroll
      for (int r = 0; r < 8; ++r) add(u[r]);
    } else if (paths == 4) {
      u32x2 u[4];
#pragma unroll
      for (int r = 0; r < 4; ++r) u[r] = ld(r);
#pragma unroll
      for (int r = 0; r < 4; ++r) add(u[r]);
    } else {
      for (int r = 0; r < paths; ++r) add(ld(r));
    }
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      neg = neg || c[j] < 0.0f;
      v[j] = fmaf((float)s[j], step, (float)paths * c[j]);
    }
  } else {
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      if (d0 + j < D) {
        uint32_t sum = 0;
        for (int r = 0; r < paths; ++r) sum += ebuf[(long)r * V + e + j];
        const float c = load_cost<CT>(cost + e + j);
        neg = neg || c < 0.0f;
        v[j] = fmaf((float)sum, step, (float)paths * c);
      } else {
        v[j] = INF;
      }
    }
  }
  return neg;
}

template <int LP, typename CT>
__device__ __forceinline__ void combine_excess_pixels(long N, int D, int paths, bool vec4,
                                                      const CT* __restrict__ cost,
                                                      const uint16_t* __restrict__ ebuf,
                                                      float step, float* __restrict__ agg,
                                                      float var_scale, float idep_min,
                                                      float dep_sample,
                                                      float* __restrict__ depth, long wave) {
  const long V = N * D;
  auto load = [&](long pc, bool own, int d0, float (&v)[4]) {
    const long e = pc * D + d0;
    int bad = excess_agg4<CT>(cost, ebuf, V, e, D, d0, paths, vec4, step, v) ? 1 : 0;
    // invalid pixel: a negative cost on any plane of the LP lanes holding it
#pragma unroll
    for (int off = LP / 2; off >= 1; off >>= 1) bad |= __shfl_xor(bad, off, 64);
    if (bad) {
#pragma unroll
      for (int j = 0; j < 4; ++j) v[j] = d0 + j < D ? kCostInf : v[j];
    }
    if (own && agg) {
      if (vec4 && d0 + 3 < D) {
        *reinterpret_cast<float4*>(agg + e) = make_float4(v[0], v[1], v[2], v[3]);
      } else {
#pragma unroll
        for (int j = 0; j < 4; ++j)
          if (d0 + j < D) agg[e + j] = v[j];
      }
    }
  };
  filter_pixels_impl<LP>(N, D, load, var_scale, idep_min, dep_sample, depth, wave);
}

template <int LP, typename CT>
__global__ __launch_bounds__(256) void combine_excess_filter_kernel(
    long N, int D, int paths, bool vec4, const CT* __restrict__ cost,
    const uint16_t* __restrict__ ebuf, float step, float* __restrict__ agg, float var_scale,
    float idep_min, float dep_sample, float* __restrict__ depth) {
  const long wave = (long)blockIdx.x * 4 + (threadIdx.x >> 6);
  combine_excess_pixels<LP, CT>(N, D, paths, vec4, cost, ebuf, step, agg, var_scale, idep_min,
                                dep_sample, depth, wave);
}

template <typename CT>
void launch_combine_excess_t(int W, int H, int D, int paths, const CT* cost,
                             const uint16_t* ebuf, float step, float* agg, float var_scale,
                             float idep_min, float dep_sample, float* depth, hipStream_t st) {
  const long N = (long)W * H;
  const bool vec4 = (D % 4 == 0) && (uintptr_t)cost % 16 == 0 && (uintptr_t)ebuf % 8 == 0 &&
                    (agg == nullptr || (uintptr_t)agg % 16 == 0);
  const int LP = filter_lanes(D);
  const unsigned blocks = (unsigned)filter_items(W, H, D);
  switch (LP) {
    case 4: combine_excess_filter_kernel<4, CT><<<blocks, 256, 0, st>>>(N, D, paths, vec4, cost, ebuf, step, agg, var_scale, idep_min, dep_sample, depth); break;
    case 8: combine_excess_filter_kernel<8, CT><<<blocks, 256, 0, st>>>(N, D, paths, vec4, cost, ebuf, step, agg, var_scale, idep_min, dep_sample, depth); break;
    case 16: combine_excess_filter_kernel<16, CT><<<blocks, 256, 0, st>>>(N, D, paths, vec4, cost, ebuf, step, agg, var_scale, idep_min, dep_sample, depth); break;
    case 32: combine_excess_filter_kernel<32, CT><<<blocks, 256, 0, st>>>(N, D, paths, vec4, cost, ebuf, step, agg, var_scale, idep_min, dep_sample, depth); break;
    default: combine_excess_filter_kernel<64, CT><<<blocks, 256, 0, st>>>(N, D, paths, vec4, cost, ebuf, step, agg, var_scale, idep_min, dep_sample, depth); break;
  }
}

void launch_combine_excess(pm_cost_type ct, int W, int H, int D, int paths, const void* cost,
                           const uint16_t* ebuf, float step, float* agg, float var_scale,
                           float idep_min, float dep_sample, float* depth, hipStream_t st) {
  switch (ct) {
    case PM_COST_F32:
      launch_combine_excess_t<float>(W, H, D, paths, (const float*)cost, ebuf, step, agg, var_scale, idep_min, dep_sample, depth, st);
      break;
    case PM_COST_F16:
      launch_combine_excess_t<__half>(W, H, D, paths, (const __half*)cost, ebuf, step, agg, var_scale, idep_min, dep_sample, depth, st);
      break;
    default:
      launch_combine_excess_t<__hip_bfloat16>(W, H, D, paths, (const __hip_bfloat16*)cost, ebuf, step, agg, var_scale, idep_min, dep_sample, depth, st);
      break;
  }
}

// Batched compact form: keyframes kf[j] of one LP class own items first[j] .. first[j+1]-1.
template <int LP, typename CT>
__global__ __launch_bounds__(256) void combine_excess_batch_kernel(const KfDesc* __restrict__ kfs,
                                                                   const int* __restrict__ kf,
                                                                   const int* __restrict__ first,
                                                                   int n, int total) {
  for (int i = blockIdx.x; i < total; i += gridDim.x) {
    const int j = owner_of(first, n, i);
    const KfDesc& k = kfs[kf[j]];
    const long N = (long)k.c.W * k.c.H;
    const int D = k.c.D;
    const bool vec4 = (D % 4 == 0) && ((uintptr_t)k.cost % 16 == 0) &&
                      ((uintptr_t)k.ebuf % 8 == 0) && ((uintptr_t)k.agg % 16 == 0);
    const long wave = (long)(i - first[j]) * 4 + (threadIdx.x >> 6);
    combine_excess_pixels<LP, CT>(N, D, k.paths, vec4, static_cast<const CT*>(k.cost), k.ebuf,
                                  k.estep, k.agg, k.var_scale, k.c.idep_min, k.c.dep_sample,
                                  k.depth, wave);
  }
}

template <typename CT>
void launch_combine_excess_batch_t(int LP, const KfDesc* kfs, const int* kf, const int* first,
                                   int n, int total, hipStream_t st) {
  switch (LP) {
    case 4: combine_excess_batch_kernel<4, CT><<<total, 256, 0, st>>>(kfs, kf, first, n, total); break;
    case 8: combine_excess_batch_kernel<8, CT><<<total, 256, 0, st>>>(kfs, kf, first, n, total); break;
    case 16: combine_excess_batch_kernel<16, CT><<<total, 256, 0, st>>>(kfs, kf, first, n, total); break;
    case 32: combine_excess_batch_kernel<32, CT><<<total, 256, 0, st>>>(kfs, kf, first, n, total); break;
    default: combine_excess_batch_kernel<64, CT><<<total, 256, 0, st>>>(kfs, kf, first, n, total); break;
  }
}

void launch_combine_excess_batch(pm_cost_type ct, int LP, const KfDesc* kfs, const int* kf,
                                 const int* first, int n, int total, hipStream_t st) {
  switch (ct) {
    case PM_COST_F32: launch_combine_excess_batch_t<float>(LP, kfs, kf, first, n, total, st); break;
    case PM_COST_F16: launch_combine_excess_batch_t<__half>(LP, kfs, kf, first, n, total, st); break;
    default: launch_combine_excess_batch_t<__hip_bfloat16>(LP, kfs, kf, first, n, total, st); break;
  }
}

template <int LP>
__global__ __launch_bounds__(256) void filter_kernel(long N, int D, bool vec4,
                                                     const float* __restrict__ agg,
                                                     float var_scale, float idep_min,
                                                     float dep_sample, float* __restrict__ depth) {
  const long wave = (long)blockIdx.x * 4 + (threadIdx.x >> 6);
  filter_pixels<LP>(N, D, vec4, agg, var_scale, idep_min, dep_sample, depth, wave);
}

// Batched form: keyframes kf[j] of one LP class own items first[j] .. first[j+1]-1, an item
// being 4 waves of 64/LP pixels each.
template <int LP>
__global__ __launch_bounds__(256) void filter_batch_kernel(const KfDesc* __restrict__ kfs,
                                                           const int* __restrict__ kf,
                                                           const int* __restrict__ first, int n,
                                                           int total) {
  for (int i = blockIdx.x; i < total; i += gridDim.x) {
    const int j = owner_of(first, n, i);
    const KfDesc& k = kfs[kf[j]];
    const long N = (long)k.c.W * k.c.H;
    const int D = k.c.D;
    const bool vec4 = (D % 4 == 0) && ((uintptr_t)k.agg % 16 == 0);
    const long wave = (long)(i - first[j]) * 4 + (threadIdx.x >> 6);
    filter_pixels<LP>(N, D, vec4, k.agg, k.var_scale, k.c.idep_min, k.c.dep_sample, k.depth,
                      wave);
  }
}

int filter_lanes(int D) { return D <= 16 ? 4 : D <= 32 ? 8 : D <= 64 ? 16 : D <= 128 ? 32 : 64; }

int filter_items(int W, int H, int D) {
  const long N = (long)W * H;
  const int LP = filter_lanes(D);
  const long waves = (N + 64 / LP - 1) / (64 / LP);
  return (int)((waves + 3) / 4);
}

void launch_filter_batch(int LP, const KfDesc* kfs, const int* kf, const int* first, int n,
                         int total, int grid, hipStream_t st) {
  switch (LP) {
    case 4: filter_batch_kernel<4><<<grid, 256, 0, st>>>(kfs, kf, first, n, total); break;
    case 8: filter_batch_kernel<8><<<grid, 256, 0, st>>>(kfs, kf, first, n, total); break;
    case 16: filter_batch_kernel<16><<<grid, 256, 0, st>>>(kfs, kf, first, n, total); break;
    case 32: filter_batch_kernel<32><<<grid, 256, 0, st>>>(kfs, kf, first, n, total); break;
    default: filter_batch_kernel<64><<<grid, 256, 0, st>>>(kfs, kf, first, n, total); break;
  }
}

// The optional 3x3 median over valid depths (median3x3_at, in common.hpp), one thread per
// pixel.
__global__ __launch_bounds__(256) void median3x3_kernel(const float* __restrict__ in,
                                                        float* __restrict__ out, int W, int H) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x;
  const int y = blockIdx.y;
  if (x >= W) return;
  out[(long)y * W + x] = median3x3_at(in, W, H, x, y);
}

// CASE 0 / 1 winner-take-all (stereo_mapper.cpp:637-660, 681-704): the running minimum
// starts at 100000 (a double) with index 0, so a pixel whose costs are all >= 100000 keeps
// index 0 (the reference leaves the index uninitialised there).
template <typename CT>
__global__ __launch_bounds__(256) void wta_kernel(long N, int D, const CT* __restrict__ cost,
                                                  float idep_min, float dep_sample,
                                                  float* __restrict__ depth) {
  const long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= N) return;
  const CT* c = cost + p * D;
  double best = 100000.0;
  int mi = 0;
  for (int d = 0; d < D; ++d) {
    const double v = (double)load_cost<CT>(c + d);
    if (v < best) { best = v; mi = d; }
  }
  float out;
  if (best < 1e-5) out = 10000.0f;
  else if (mi == 0 || mi == D - 1) out = 0.0f;
  else out = 1.0f / (idep_min + (float)mi * dep_sample);
  depth[p] = out;
}

}  // namespace pm

using namespace pm;

extern "C" pm_status pm_wta(pm_shape shape, const void* cost, pm_cost_type cost_type,
                            float idep_min, float dep_sample, float* depth, pm_stream stream) {
  clear_error();
  if (!cost || !depth) return fail(PM_ERR_INVALID_ARGUMENT, "pm_wta: null pointer");
  if (shape.width <= 0 || shape.height <= 0 || shape.width > 65535 || shape.height > 65535)
    return fail(PM_ERR_INVALID_ARGUMENT, "pm_wta: bad image size");
  if (shape.planes <= 0 || shape.planes > PM_MAX_PLANES)
    return fail(PM_ERR_INVALID_ARGUMENT, "pm_wta: planes must be in 1..256");
  const long N = (long)shape.width * shape.height;
  const unsigned blocks = (unsigned)((N + 255) / 256);
  hipStream_t st = (hipStream_t)stream;
  switch (cost_type) {
    case PM_COST_F32:
      wta_kernel<float><<<blocks, 256, 0, st>>>(N, shape.planes, (const float*)cost, idep_min,
                                                dep_sample, depth);
      break;
    case PM_COST_F16:
      wta_kernel<__half><<<blocks, 256, 0, st>>>(N, shape.planes, (const __half*)cost,
                                                 idep_min, dep_sample, depth);
      break;
    case PM_COST_BF16:
      wta_kernel<__hip_bfloat16><<<blocks, 256, 0, st>>>(
          N, shape.planes, (const __hip_bfloat16*)cost, idep_min, dep_sample, depth);
      break;
    default:
      return fail(PM_ERR_INVALID_ARGUMENT, "pm_wta: unknown cost type");
  }
  return check_launch("wta_kernel");
}

extern "C" pm_status pm_filter_cost(pm_shape shape, const float* agg, float var_scale,
                                    float idep_min, float dep_sample, float* depth,
                                    pm_stream stream) {
  clear_error();
  if (!agg || !depth) return fail(PM_ERR_INVALID_ARGUMENT, "pm_filter_cost: null pointer");
  if (shape.width <= 0 || shape.height <= 0 || shape.width > 65535 || shape.height > 65535)
    return fail(PM_ERR_INVALID_ARGUMENT, "pm_filter_cost: bad image size");
  if (shape.planes <= 0 || shape.planes > PM_MAX_PLANES)
    return fail(PM_ERR_INVALID_ARGUMENT, "pm_filter_cost: planes must be in 1..256");
  const long N = (long)shape.width * shape.height;
  const bool vec4 = (shape.planes % 4 == 0) && ((uintptr_t)agg % 16 == 0);
  const int D = shape.planes;
  const int LP = filter_lanes(D);
  const unsigned blocks = (unsigned)filter_items(shape.width, shape.height, D);
  hipStream_t st = (hipStream_t)stream;
  switch (LP) {
    case 4: filter_kernel<4><<<blocks, 256, 0, st>>>(N, D, vec4, agg, var_scale, idep_min, dep_sample, depth); break;
    case 8: filter_kernel<8><<<blocks, 256, 0, st>>>(N, D, vec4, agg, var_scale, idep_min, dep_sample, depth); break;
    case 16: filter_kernel<16><<<blocks, 256, 0, st>>>(N, D, vec4, agg, var_scale, idep_min, dep_sample, depth); break;
    case 32: filter_kernel<32><<<blocks, 256, 0, st>>>(N, D, vec4, agg, var_scale, idep_min, dep_sample, depth); break;
    default: filter_kernel<64><<<blocks, 256, 0, st>>>(N, D, vec4, agg, var_scale, idep_min, dep_sample, depth); break;
  }
  return check_launch("filter_kernel");
}

extern "C" pm_status pm_median3x3(const float* in, float* out, int32_t width, int32_t height,
                                  pm_stream stream) {
  clear_error();
  if (!in || !out) return fail(PM_ERR_INVALID_ARGUMENT, "pm_median3x3: null pointer");
  if (in == out) return fail(PM_ERR_INVALID_ARGUMENT, "pm_median3x3: in and out alias");
  if (width <= 0 || height <= 0 || width > 65535 || height > 65535)
    return fail(PM_ERR_INVALID_ARGUMENT, "pm_median3x3: bad image size");
  dim3 grid((width + 255) / 256, height);
  median3x3_kernel<<<grid, 256, 0, (hipStream_t)stream>>>(in, out, width, height);
  return check_launch("median3x3_kernel");
}

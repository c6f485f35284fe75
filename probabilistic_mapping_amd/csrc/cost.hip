// Plane-sweep matching cost, fused over all measurement frames of a keyframe.
//
// Reference: ADCalcCostKernel + ad_calc_cost, motion_stereo/calc_cost.cu:11-253, called once
// per measurement frame by StereoMapper::update (stereo_mapper.cpp:73-91) on a volume reset
// to -1 (stereo_mapper.cpp:69-70).  The reference keeps a running mean in the cost volume and
// a second float volume of counts, re-reading and re-writing both per source (16 B per element
// per source).  Here one launch visits every source while the per-(pixel, plane) sum and count
// stay in registers, and the volume is written once.
//
// Mapping: see cost_tile_kernel -- lanes are image columns, each pixel's centre warp is
// computed once per (plane, source) and shared with the neighbours whose reference terms use
// it, so a (pixel, plane, source) costs about 3.25 warps instead of 9.
//
// Bilinear sampling reads the "quad" texture of a measurement frame: 8 bytes per texel holding
// the bilinear patch from (x, y) to (x+1, y+1) as four FP16 numbers
//   a = I(x,y),  bx = I(x+1,y) - a,  by = I(x,y+1) - a,  bxy = I(x+1,y+1) - I(x+1,y) - by
// (border 0).  They are integers in [-510, 510], exact in FP16, and the sample
//   a + ax bx + ay (by + ax bxy)
// is three FMAs with FP16 operands widened in the instruction (v_fma_mix_f32).
#include "common.hpp"

#ifndef PM_COST_TILE_ROWS
#define PM_COST_TILE_ROWS 4
#endif
#ifndef PM_COST_STAGE_PLANES
#define PM_COST_STAGE_PLANES 8
#endif

namespace pm {


__global__ __launch_bounds__(256) void make_quad_kernel(const uint8_t* __restrict__ img,
                                                        uint64_t* __restrict__ quad, int W,
                                                        int H) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x;
  const int y = blockIdx.y;
  if (x >= W) return;
  const int i00 = img[(long)y * W + x];
  const int i10 = x + 1 < W ? img[(long)y * W + x + 1] : 0;
  const int i01 = y + 1 < H ? img[(long)(y + 1) * W + x] : 0;
  const int i11 = (x + 1 < W && y + 1 < H) ? img[(long)(y + 1) * W + x + 1] : 0;
  const __half2 lo = __floats2half2_rn((float)i00, (float)(i10 - i00));
  const __half2 hi = __floats2half2_rn((float)(i01 - i00), (float)(i11 - i10 - i01 + i00));
  uint2 t;
  t.x = *reinterpret_cast<const uint32_t*>(&lo);
  t.y = *reinterpret_cast<const uint32_t*>(&hi);
  reinterpret_cast<uint2*>(quad)[(long)y * W + x] = t;
}

// Bilinear sample of the quad texture at (U, V) with the warp's validity (calc_cost.cu:85-93);
// returns +inf for an invalid warp.  Validity is "the clamp to the image changes nothing" and
// w >= 0, so a NaN warp (0/0) is invalid; an invalid warp still reads an in-bounds texel (the
// clamped position), so every gather of a row can be in flight at once.  The byte offset fits
// 32 bits (W H <= 2^29 is checked on the host), which lets the gather use the scalar base
// address + 32-bit vector offset form.
__device__ __forceinline__ float warp_sample(const uint64_t* __restrict__ quad, uint32_t W,
                                             float Wm1, float Hm1, float X, float Y, float w) {
  const float rw = __builtin_amdgcn_rcpf(w);
  const float U = X * rw;
  const float V = Y * rw;
  const float Uc = __builtin_amdgcn_fmed3f(U, 0.0f, Wm1);
  const float Vc = __builtin_amdgcn_fmed3f(V, 0.0f, Hm1);
  const bool valid = (w >= 0.0f) & (Uc == U) & (Vc == V);
  // non-negative, so truncation is floor; W, y <= 65535 fit the 24-bit multiply
  const uint32_t off = __umul24((uint32_t)Vc, W) * 8u + (uint32_t)Uc * 8u;
  const uint2 t = *reinterpret_cast<const uint2*>(reinterpret_cast<const char*>(quad) + off);
  const float ax = __builtin_amdgcn_fractf(Uc), ay = __builtin_amdgcn_fractf(Vc);
  const __half2 lo = *reinterpret_cast<const __half2*>(&t.x);
  const __half2 hi = *reinterpret_cast<const __half2*>(&t.y);
  const float top = fmaf(ax, __high2float(lo), __low2float(lo));
  const float dv = fmaf(ax, __high2float(hi), __low2float(hi));
  return valid ? fmaf(ay, dv, top) : __uint_as_float(0x7F800000u);
}

// Tile kernel.  A wave covers a strip of 64 columns u = x0-1 .. x0+62 (62 output columns and
// a one-column halo on each side) and TY output rows, for a quarter of the planes.  For each
// (plane, source) it walks the rows y0-1 .. y0+TY once, warping every pixel's own centre
// H_d (u, r, 1) a single time.  The reference's nine warps of pixel (u, v) are the centre
// warps of its 3x3 neighbours, except two garbled ones (calc_cost.cu:50-51, 70-71):
//   term 0 (u,v), 3 (u-1,v), 5 (u-1,v-1), 6 (u+1,v+1), 7 (u-1,v+1): |I(q) - S(q)| at q
//   term 1: warp (u,v-1) against I(u,v+1);  term 2: warp (u,v+1) against I(u,v-1)
//   term 4: (x+h11, x+h21, x+h31) against I(u+1,v)
//   term 8: (x+h11-h12, x+h11-h22, x+h11-h32) against I(u+1,v-1),  x = h11 u + h12 v + h13
// so each centre sample yields |I(q)-S|, |I(q+(0,2))-S| and |I(q-(0,2))-S|, kept in a 3-row
// register window; column neighbours come from DPP lane shifts.  An invalid warp is +inf, so
// a source contributes iff the 9-term sum is finite (calc_cost.cu:89-91).
// Output staging: a wave finishes DC planes for its TY x 62 pixels, parks them in its own LDS
// tile [TY][DC][64] (lanes contiguous, conflict-free), then writes each pixel's DC contiguous
// planes with consecutive lanes, so every store instruction writes whole aligned segments
// instead of 64 isolated 4-byte words 4*D bytes apart.
template <typename CT>
__device__ __forceinline__ void store_quad(CT* dst, float a, float b, float c, float d);
template <>
__device__ __forceinline__ void store_quad<float>(float* dst, float a, float b, float c,
                                                  float d) {
  *reinterpret_cast<float4*>(dst) = make_float4(a, b, c, d);
}
template <>
__device__ __forceinline__ void store_quad<__half>(__half* dst, float a, float b, float c,
                                                   float d) {
  const __half2 lo = __floats2half2_rn(a, b), hi = __floats2half2_rn(c, d);
  uint2 u;
  u.x = *reinterpret_cast<const uint32_t*>(&lo);
  u.y = *reinterpret_cast<const uint32_t*>(&hi);
  *reinterpret_cast<uint2*>(dst) = u;
}
template <>
__device__ __forceinline__ void store_quad<__hip_bfloat16>(__hip_bfloat16* dst, float a,
                                                           float b, float c, float d) {
  dst[0] = __float2bfloat16(a); dst[1] = __float2bfloat16(b);
  dst[2] = __float2bfloat16(c); dst[3] = __float2bfloat16(d);
}

template <typename CT, int TY, int DC>
__device__ __forceinline__ void cost_tile(const CostArgs& a, int blk, CT* __restrict__ cost,
                                          float* __restrict__ tile) {
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int nstrips = (a.W + 61) / 62;
  const int strip = blk % nstrips;
  const int y0 = (blk / nstrips) * TY;
  const int W = a.W, H = a.H, D = a.D;
  const int u = strip * 62 - 1 + lane;
  const int dper = (D + 3) / 4;
  const int dbeg = wave * dper, dend = min(D, dbeg + dper);
  const float Wm1 = (float)(W - 1), Hm1 = (float)(H - 1);
  const float INF = __uint_as_float(0x7F800000u);
  const float fu = (float)u;
  // reference intensities of this column for rows y0-3 .. y0+TY+2 (0 outside the image), and
  // of the right-hand neighbour column for the garbled terms' rows y0-1 .. y0+TY-1
  float I[TY + 6], IR[TY + 1];
#pragma unroll
  for (int k = 0; k < TY + 6; ++k) I[k] = texel_u8(a.ref, W, H, u, y0 - 3 + k);
#pragma unroll
  for (int k = 0; k <= TY; ++k) IR[k] = wave_shl1z(I[k + 2]);

  // vector stores need D and the chunk start to be multiples of 4 (16-byte / 8-byte aligned)
  const bool vec = (D % 4 == 0) && (dper % 4 == 0);
  for (int dc0 = dbeg; dc0 < dend; dc0 += DC) {
  const int dcn = min(DC, dend - dc0);
  for (int d = dc0; d < dc0 + dcn; ++d) {
    const float idep = a.idep_min + (float)d * a.dep_sample;
    float acc[TY], cnt[TY];
#pragma unroll
    for (int t = 0; t < TY; ++t) { acc[t] = 0.0f; cnt[t] = 0.0f; }
    for (int s = 0; s < a.S; ++s) {
      const float* hm = a.hm[s];
      const float h11 = hm[0], h12 = hm[1], h13 = hm[2], h21 = hm[3], h22 = hm[4],
                  h23 = hm[5], h31 = hm[6], h32 = hm[7], h33 = hm[8];
      const float tx = hm[9] * idep, ty = hm[10] * idep, tz = hm[11] * idep;
      const uint64_t* q = a.quad[s];
      // column parts of H (u, r, 1) + t idep
      const float Xu = h11 * fu + h13 + tx;
      const float Yu = h21 * fu + h23 + ty;
      const float Zu = h31 * fu + h33 + tz;
      // garbled warps, offsets from x = h11 u + h12 v + h13 (calc_cost.cu:49-51, 69-71)
      const float xu = h11 * fu + h13;
      const float g4x = xu + (h11 + tx), g4y = xu + (h21 + ty), g4z = xu + (h31 + tz);
      const float g8x = xu + (h11 - h12 + tx), g8y = xu + (h11 - h22 + ty),
                  g8z = xu + (h11 - h32 + tz);
      // window over centre rows r-2, r-1: |I-S| (A, row r-1 only), its left-neighbour copy (B),
      // up terms (Up)
      float A1 = 0.0f, B0 = 0.0f, B1 = 0.0f, Up0 = 0.0f, Up1 = 0.0f;
#pragma unroll
      for (int k = 0; k < TY + 2; ++k) {   // centre row r = y0 - 1 + k
        const float fr = (float)(y0 - 1 + k);
        const float S = warp_sample(q, (uint32_t)W, Wm1, Hm1, fmaf(h12, fr, Xu),
                                    fmaf(h22, fr, Yu), fmaf(h32, fr, Zu));
        const float A2 = fabsf(I[k + 2] - S);     // |I(u,r) - S(u,r)|
        const float Up2 = fabsf(I[k + 4] - S);    // |I(u,r+2) - S(u,r)|
        const float B2 = wave_shr1z(A2);          // the same at (u-1, r)
        if (k >= 2) {
          const int t = k - 2;                    // output row v = r - 1
          const float fv = (float)(y0 + t);
          const float Dn = fabsf(I[k] - S);       // |I(u,r-2) - S(u,r)| = term 2 of row v
          const float G4 = fabsf(IR[t + 1] - warp_sample(q, (uint32_t)W, Wm1, Hm1,
                                                         fmaf(h12, fv, g4x), fmaf(h12, fv, g4y),
                                                         fmaf(h12, fv, g4z)));
          const float G8 = fabsf(IR[t] - warp_sample(q, (uint32_t)W, Wm1, Hm1,
                                                     fmaf(h12, fv, g8x), fmaf(h12, fv, g8y),
                                                     fmaf(h12, fv, g8z)));
          float T = A1 + B0;                      // terms 0, 5
          T += B1;                                // term 3
          T += wave_shl1z(A2);                    // term 6
          T += B2;                                // term 7
          T += Up0 + Dn;                          // terms 1, 2
          T += G4 + G8;                           // terms 4, 8
          const bool ok = T < INF;
          acc[t] += ok ? T : 0.0f;
          cnt[t] += ok ? 1.0f : 0.0f;
        }
        A1 = A2;
        B0 = B1; B1 = B2;
        Up0 = Up1; Up1 = Up2;
      }
    }
#pragma unroll
    for (int t = 0; t < TY; ++t)   // mean over sources of (sum of 9 terms)/9, or -1
      tile[(t * DC + (d - dc0)) * 64 + lane] = cnt[t] > 0.0f ? acc[t] / (9.0f * cnt[t]) : -1.0f;
  }
  __builtin_amdgcn_wave_barrier();
  const int x0 = strip * 62;
  if (vec) {
    // items: (row t, pixel 0..61, quad of 4 planes); consecutive lanes cover one pixel's planes
    const int qp = dcn / 4, items = TY * 62 * qp;
    for (int it = lane; it < items; it += 64) {
      const int q = it % qp, rest = it / qp, px = rest % 62, t = rest / 62;
      const int v = y0 + t, x = x0 + px;
      if (v < H && x < W) {
        const float* src = tile + (t * DC + 4 * q) * 64 + px + 1;
        store_quad<CT>(cost + ((long)v * W + x) * D + dc0 + 4 * q, src[0], src[64], src[128],
                       src[192]);
      }
    }
  } else {
    const int items = TY * 62 * dcn;
    for (int it = lane; it < items; it += 64) {
      const int j = it % dcn, rest = it / dcn, px = rest % 62, t = rest / 62;
      const int v = y0 + t, x = x0 + px;
      if (v < H && x < W)
        cost[((long)v * W + x) * D + dc0 + j] = store_cost<CT>(tile[(t * DC + j) * 64 + px + 1]);
    }
  }
  __builtin_amdgcn_wave_barrier();
  }
}

// Tiles of a W x H keyframe: 62-column strips x TY-row blocks.
template <int TY>
__host__ __device__ inline int cost_tiles(int W, int H) {
  return ((W + 61) / 62) * ((H + TY - 1) / TY);
}

template <typename CT, int TY, int DC>
__global__ __launch_bounds__(256) void cost_tile_kernel(const CostArgs a, CT* __restrict__ cost) {
  __shared__ float stage[4][TY * DC * 64];
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  cost_tile<CT, TY, DC>(a, xcd_block(blockIdx.x, gridDim.x), cost, stage[wave]);
}

// Batched form: the tiles of every keyframe of a batch form one work list (keyframe k owns
// items first[k] .. first[k+1]-1).  Persistent blocks stride through it; with a grid that is a
// multiple of 8, block i always runs on XCD i % 8, and xcd_block gives each XCD a contiguous
// range of the list.
template <typename CT, int TY, int DC>
__global__ __launch_bounds__(256) void cost_batch_kernel(const KfDesc* __restrict__ kfs,
                                                         const int* __restrict__ first, int nkf,
                                                         int total) {
  __shared__ float stage[4][TY * DC * 64];
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  for (int i = blockIdx.x; i < total; i += gridDim.x) {
    const int item = xcd_block(i, total);
    const int k = owner_of(first, nkf, item);
    cost_tile<CT, TY, DC>(kfs[k].c, item - first[k], static_cast<CT*>(kfs[k].cost),
                          stage[wave]);
  }
}

int cost_items(int W, int H, int tile_rows) {
  return tile_rows == 2 ? cost_tiles<2>(W, H) : cost_tiles<4>(W, H);
}

template <typename CT>
bool launch_cost_batch_t(int tile_rows, int stage_planes, const KfDesc* kfs, const int* first,
                         int nkf, int total, int grid, hipStream_t st) {
  if (tile_rows == 2 && stage_planes == 8)
    cost_batch_kernel<CT, 2, 8><<<grid, 256, 0, st>>>(kfs, first, nkf, total);
  else if (tile_rows == 2 && stage_planes == 16)
    cost_batch_kernel<CT, 2, 16><<<grid, 256, 0, st>>>(kfs, first, nkf, total);
  else if (tile_rows == 4 && stage_planes == 8)
    cost_batch_kernel<CT, 4, 8><<<grid, 256, 0, st>>>(kfs, first, nkf, total);
  else if (tile_rows == 4 && stage_planes == 16)
    cost_batch_kernel<CT, 4, 16><<<grid, 256, 0, st>>>(kfs, first, nkf, total);
  else
    return false;
  return true;
}

bool launch_cost_batch(pm_cost_type ct, int tile_rows, int stage_planes, const KfDesc* kfs,
                       const int* first, int nkf, int total, int grid, hipStream_t st) {
  switch (ct) {
    case PM_COST_F32:
      return launch_cost_batch_t<float>(tile_rows, stage_planes, kfs, first, nkf, total, grid, st);
    case PM_COST_F16:
      return launch_cost_batch_t<__half>(tile_rows, stage_planes, kfs, first, nkf, total, grid, st);
    case PM_COST_BF16:
      return launch_cost_batch_t<__hip_bfloat16>(tile_rows, stage_planes, kfs, first, nkf, total,
                                                 grid, st);
    default:
      return false;
  }
}

}  // namespace pm

using namespace pm;

extern "C" pm_status pm_make_quad(const uint8_t* image, uint64_t* quad, int32_t width,
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
                                  const uint64_t* const* src_quads, const pm_homography* homs,
                                  float idep_min, float dep_sample, void* cost,
                                  pm_cost_type cost_type, pm_stream stream) {
  clear_error();
  if (!ref || !cost) return fail(PM_ERR_INVALID_ARGUMENT, "pm_calc_cost: null pointer");
  if (shape.width <= 0 || shape.height <= 0 || shape.width > 65535 || shape.height > 65535 ||
      (long)shape.width * shape.height > (1L << 29))
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
  constexpr int TY = PM_COST_TILE_ROWS;
  constexpr int DC = PM_COST_STAGE_PLANES;
  const unsigned blocks = (unsigned)cost_tiles<TY>(a.W, a.H);
  hipStream_t st = (hipStream_t)stream;
  switch (cost_type) {
    case PM_COST_F32:
      cost_tile_kernel<float, TY, DC><<<blocks, 256, 0, st>>>(a, (float*)cost);
      break;
    case PM_COST_F16:
      cost_tile_kernel<__half, TY, DC><<<blocks, 256, 0, st>>>(a, (__half*)cost);
      break;
    case PM_COST_BF16:
      cost_tile_kernel<__hip_bfloat16, TY, DC><<<blocks, 256, 0, st>>>(a, (__hip_bfloat16*)cost);
      break;
    default:
      return fail(PM_ERR_INVALID_ARGUMENT, "pm_calc_cost: unknown cost type");
  }
  return check_launch("cost_tile_kernel");
}

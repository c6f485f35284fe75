// Plane-sweep cost tile body shared by the cost kernels (plain, batched, differentiable) and
// the cost adjoint.  Reference: ADCalcCostKernel, motion_stereo/calc_cost.cu:11-210.
#pragma once

#include <type_traits>

#include "common.hpp"

namespace pm {

// A measurement frame's quad texture ([H][W] texels of 8 bytes) read through a buffer resource:
// wave-uniform base and size in scalar registers, a 32-bit byte offset per lane, and the
// hardware range check (a read past W H 8 bytes returns 0 instead of touching memory).  W H < 2^29
// is checked on the host, so the byte size fits the 32-bit record count.
struct QuadTex {
  __amdgpu_buffer_rsrc_t r;
  __device__ __forceinline__ QuadTex(const uint64_t* q, uint32_t W, uint32_t H)
      : r(__builtin_amdgcn_make_buffer_rsrc(const_cast<uint64_t*>(q), 0, (int)(W * H * 8u),
                                            kBufferDword3)) {}
  __device__ __forceinline__ uint2 at(uint32_t byte_off) const {
    const auto v = __builtin_amdgcn_raw_buffer_load_b64(r, (int)byte_off, 0, 0);
    return make_uint2(v[0], v[1]);
  }
};

// Bilinear sample of the quad texture at (U, V) with the warp's validity (calc_cost.cu:85-93);
// returns +inf for an invalid warp.  Validity is "the clamp to the image changes nothing" and
// w >= 0, so a NaN warp (0/0) is invalid; an invalid warp still reads an in-bounds texel (the
// clamped position), so every gather of a row can be in flight at once.
__device__ __forceinline__ float warp_sample(const QuadTex& quad, uint32_t W,
                                             float Wm1, float Hm1, float X, float Y, float w) {
  const float rw = __builtin_amdgcn_rcpf(w);
  const float U = X * rw;
  const float V = Y * rw;
  const float Uc = __builtin_amdgcn_fmed3f(U, 0.0f, Wm1);
  const float Vc = __builtin_amdgcn_fmed3f(V, 0.0f, Hm1);
  const bool valid = (w >= 0.0f) & (Uc == U) & (Vc == V);
  // non-negative, so truncation is floor; W, y <= 65535 fit the 24-bit multiply
  const uint32_t off = __umul24((uint32_t)Vc, W) * 8u + (uint32_t)Uc * 8u;
  const uint2 t = quad.at(off);
  const float ax = __builtin_amdgcn_fractf(Uc), ay = __builtin_amdgcn_fractf(Vc);
  const __half2 lo = *reinterpret_cast<const __half2*>(&t.x);
  const __half2 hi = *reinterpret_cast<const __half2*>(&t.y);
  const float top = fmaf(ax, __high2float(lo), __low2float(lo));
  const float dv = fmaf(ax, __high2float(hi), __low2float(hi));
  return valid ? fmaf(ay, dv, top) : __uint_as_float(0x7F800000u);
}

// warp_sample for a warp known to be valid and inside the image (see tile_inside_mask): the
// same arithmetic without the clamp and the validity test, so it returns the same value.
__device__ __forceinline__ float warp_sample_inside(const QuadTex& quad, uint32_t W,
                                                    float X, float Y, float w) {
  const float rw = __builtin_amdgcn_rcpf(w);
  const float U = X * rw;
  const float V = Y * rw;
  const uint32_t off = __umul24((uint32_t)V, W) * 8u + (uint32_t)U * 8u;
  const uint2 t = quad.at(off);
  const float ax = __builtin_amdgcn_fractf(U), ay = __builtin_amdgcn_fractf(V);
  const __half2 lo = *reinterpret_cast<const __half2*>(&t.x);
  const __half2 hi = *reinterpret_cast<const __half2*>(&t.y);
  const float top = fmaf(ax, __high2float(lo), __low2float(lo));
  const float dv = fmaf(ax, __high2float(hi), __low2float(hi));
  return fmaf(ay, dv, top);
}

// Conservative whole-tile validity, per source.  Every warp a cost_tile wave evaluates for one
// (plane, source) is a linear-fractional map of the tile's pixel coordinates: the centre warps
// H (u, r, 1) + t idep over the rectangle u in [u_lo, u_hi], r in [r_lo, r_hi], and the two
// garbled warps (x + c_X, x + c_Y, x + c_Z) over x = h11 u + h12 v + h13 in [x_min, x_max].
// Where the denominator w is positive at the corners it is positive on the whole box, and U, V
// take their extremes at the corners; so if every corner has w >= kInsideW and U, V inside the
// image by kInsideMargin pixels, every warp of the tile is valid with room to spare for FP32
// rounding (< 4e-3 px at w >= kInsideW for |coordinates| < 2^16).  Then the validity tests can
// be skipped: warp_sample_inside returns what warp_sample would.
// Lane l tests corner l & 7 of source l >> 3 (sources 0..7); tile_inside_mask returns a
// wave-uniform mask whose bit s says source s passed for the plane at idep.
constexpr float kInsideW = 0.25f;
constexpr float kInsideMargin = 1.0f / 32.0f;

struct InsideCorner {
  float X0, Y0, Z0;   // the corner's warp at idep = 0
  float tx, ty, tz;   // its idep slope (t of the lane's source)
  bool live;          // lane tests a real source
};

template <typename CA>
__device__ __forceinline__ InsideCorner tile_inside_setup(const CA& a, float u_lo, float u_hi,
                                                          float r_lo, float r_hi, float v_lo,
                                                          float v_hi) {
  const int lane = threadIdx.x & 63;
  const int s = lane >> 3, k = lane & 7;
  InsideCorner c;
  c.live = s < a.S && s < 8;
  const int ss = c.live ? s : 0;
  const float* hm = a.hm[ss];
  const float h11 = hm[0], h12 = hm[1], h13 = hm[2], h21 = hm[3], h22 = hm[4], h23 = hm[5],
              h31 = hm[6], h32 = hm[7], h33 = hm[8];
  c.tx = hm[9]; c.ty = hm[10]; c.tz = hm[11];
  if (k < 4) {
    const float u = (k & 1) ? u_hi : u_lo, r = (k & 2) ? r_hi : r_lo;
    c.X0 = h11 * u + h12 * r + h13;
    c.Y0 = h21 * u + h22 * r + h23;
    c.Z0 = h31 * u + h32 * r + h33;
  } else {
    const float x00 = h11 * u_lo + h12 * v_lo + h13, x10 = h11 * u_hi + h12 * v_lo + h13;
    const float x01 = h11 * u_lo + h12 * v_hi + h13, x11 = h11 * u_hi + h12 * v_hi + h13;
    const float xmin = fminf(fminf(x00, x10), fminf(x01, x11));
    const float xmax = fmaxf(fmaxf(x00, x10), fmaxf(x01, x11));
    const float x = (k & 1) ? xmax : xmin;
    if (k < 6) {   // term 4: (x + h11, x + h21, x + h31)
      c.X0 = x + h11; c.Y0 = x + h21; c.Z0 = x + h31;
    } else {       // term 8: (x + h11 - h12, x + h11 - h22, x + h11 - h32)
      c.X0 = x + (h11 - h12); c.Y0 = x + (h11 - h22); c.Z0 = x + (h11 - h32);
    }
  }
  return c;
}

__device__ __forceinline__ uint32_t tile_inside_mask(const InsideCorner& c, float idep,
                                                     float Wm1, float Hm1) {
  const float X = fmaf(c.tx, idep, c.X0), Y = fmaf(c.ty, idep, c.Y0);
  const float Z = fmaf(c.tz, idep, c.Z0);
  const float U = X / Z, V = Y / Z;
  const bool ok = !c.live || (Z >= kInsideW && U >= kInsideMargin && U <= Wm1 - kInsideMargin &&
                              V >= kInsideMargin && V <= Hm1 - kInsideMargin);
  const uint64_t bad = __ballot(!ok);
  uint32_t m = 0;
#pragma unroll
  for (int s = 0; s < 8; ++s) m |= ((bad >> (8 * s)) & 0xFFull) == 0 ? (1u << s) : 0u;
  return m;
}

// warp_sample with the sample's derivatives with respect to the warped point (X, Y, w):
// S = B(X/w, Y/w), dS/dX = a, dS/dY = b, dS/dw = c.  The value is computed exactly as
// warp_sample computes it.  Inside a texel cell the bilinear slopes are
// dB/dU = bx + ay bxy and dB/dV = by + ax bxy (the cell of floor(U), floor(V)); an invalid warp
// returns +inf and zero derivatives.
struct SampleGrad {
  float s, a, b, c;
};
__device__ __forceinline__ SampleGrad warp_sample_grad(const QuadTex& quad,
                                                       uint32_t W, float Wm1, float Hm1,
                                                       float X, float Y, float w) {
  const float rw = __builtin_amdgcn_rcpf(w);
  const float U = X * rw;
  const float V = Y * rw;
  const float Uc = __builtin_amdgcn_fmed3f(U, 0.0f, Wm1);
  const float Vc = __builtin_amdgcn_fmed3f(V, 0.0f, Hm1);
  const bool valid = (w >= 0.0f) & (Uc == U) & (Vc == V);
  const uint32_t off = __umul24((uint32_t)Vc, W) * 8u + (uint32_t)Uc * 8u;
  const uint2 t = quad.at(off);
  const float ax = __builtin_amdgcn_fractf(Uc), ay = __builtin_amdgcn_fractf(Vc);
  const __half2 lo = *reinterpret_cast<const __half2*>(&t.x);
  const __half2 hi = *reinterpret_cast<const __half2*>(&t.y);
  const float top = fmaf(ax, __high2float(lo), __low2float(lo));
  const float dv = fmaf(ax, __high2float(hi), __low2float(hi));
  const float dU = fmaf(ay, __high2float(hi), __high2float(lo));
  SampleGrad g;
  g.s = valid ? fmaf(ay, dv, top) : __uint_as_float(0x7F800000u);
  const float a = dU * rw, b = dv * rw;
  g.a = valid ? a : 0.0f;
  g.b = valid ? b : 0.0f;
  g.c = valid ? -fmaf(a, Uc, b * Vc) : 0.0f;
  return g;
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

// RT: reference element type (uint8_t, or float for the differentiable form).  CNT: also write
// the number of sources that saw each (pixel, plane) to cnt (uint8 [H][W][D]), staged in ctile.
// Work item blk is tile blk / npg, plane group blk % npg: the tile's planes split over npg
// blocks of 4 waves (npg > 1 gives a single keyframe's launch enough blocks to fill the chip).
template <typename CT, int TY, int DC, typename RT = uint8_t, bool CNT = false>
__device__ __forceinline__ void cost_tile(const CostArgs& a, int blk, CT* __restrict__ cost,
                                          float* __restrict__ tile,
                                          uint8_t* __restrict__ cnt_out = nullptr,
                                          uint8_t* __restrict__ ctile = nullptr, int npg = 1) {
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int pg = blk % npg;
  blk /= npg;
  const int nstrips = (a.W + 61) / 62;
  const int strip = blk % nstrips;
  const int y0 = (blk / nstrips) * TY;
  const int W = a.W, H = a.H, D = a.D;
  const int u = strip * 62 - 1 + lane;
  const int dper = (D + 4 * npg - 1) / (4 * npg);
  const int dbeg = (pg * 4 + wave) * dper, dend = min(D, dbeg + dper);
  const float Wm1 = (float)(W - 1), Hm1 = (float)(H - 1);
  const float INF = __uint_as_float(0x7F800000u);
  const float fu = (float)u;
  // reference intensities of this column for rows y0-3 .. y0+TY+2 (0 outside the image), and
  // of the right-hand neighbour column for the garbled terms' rows y0-1 .. y0+TY-1
  float I[TY + 6], IR[TY + 1];
#pragma unroll
  for (int k = 0; k < TY + 6; ++k)
    I[k] = texel(static_cast<const RT*>(a.ref), W, H, u, y0 - 3 + k);
#pragma unroll
  for (int k = 0; k <= TY; ++k) IR[k] = wave_shl1z(I[k + 2]);

  // vector stores need D and the chunk start to be multiples of 4 (16-byte / 8-byte aligned)
  const bool vec = (D % 4 == 0) && (dper % 4 == 0);
  const InsideCorner corner =
      tile_inside_setup(a, (float)(strip * 62 - 1), (float)(strip * 62 + 62), (float)(y0 - 1),
                        (float)(y0 + TY), (float)y0, (float)(y0 + TY - 1));
  for (int dc0 = dbeg; dc0 < dend; dc0 += DC) {
  const int dcn = min(DC, dend - dc0);
  for (int d = dc0; d < dc0 + dcn; ++d) {
    const float idep = a.idep_min + (float)d * a.dep_sample;
    const uint32_t inside = tile_inside_mask(corner, idep, Wm1, Hm1);
    float acc[TY], cnt[TY];
#pragma unroll
    for (int t = 0; t < TY; ++t) { acc[t] = 0.0f; cnt[t] = 0.0f; }
    for (int s = 0; s < a.S; ++s) {
      const float* hm = a.hm[s];
      const float h11 = hm[0], h12 = hm[1], h13 = hm[2], h21 = hm[3], h22 = hm[4],
                  h23 = hm[5], h31 = hm[6], h32 = hm[7], h33 = hm[8];
      const float tx = hm[9] * idep, ty = hm[10] * idep, tz = hm[11] * idep;
      const QuadTex q(a.quad[s], (uint32_t)W, (uint32_t)H);
      // column parts of H (u, r, 1) + t idep
      const float Xu = h11 * fu + h13 + tx;
      const float Yu = h21 * fu + h23 + ty;
      const float Zu = h31 * fu + h33 + tz;
      // garbled warps, offsets from x = h11 u + h12 v + h13 (calc_cost.cu:49-51, 69-71)
      const float xu = h11 * fu + h13;
      const float g4x = xu + (h11 + tx), g4y = xu + (h21 + ty), g4z = xu + (h31 + tz);
      const float g8x = xu + (h11 - h12 + tx), g8y = xu + (h11 - h22 + ty),
                  g8z = xu + (h11 - h32 + tz);
      // one pass over the centre rows; INSIDE: every warp of the tile is valid (no tests)
      auto pass = [&](auto inside_tag) {
        constexpr bool INSIDE = decltype(inside_tag)::value;
        auto sample = [&](float X, float Y, float w) {
          return INSIDE ? warp_sample_inside(q, (uint32_t)W, X, Y, w)
                        : warp_sample(q, (uint32_t)W, Wm1, Hm1, X, Y, w);
        };
        // window over centre rows r-2, r-1: |I-S| (A, row r-1 only), its left-neighbour copy
        // (B), up terms (Up)
        float A1 = 0.0f, B0 = 0.0f, B1 = 0.0f, Up0 = 0.0f, Up1 = 0.0f;
#pragma unroll
        for (int k = 0; k < TY + 2; ++k) {   // centre row r = y0 - 1 + k
          const float fr = (float)(y0 - 1 + k);
          const float S = sample(fmaf(h12, fr, Xu), fmaf(h22, fr, Yu), fmaf(h32, fr, Zu));
          // signed differences; the |.| is taken where they are summed (a source modifier)
          const float A2 = I[k + 2] - S;            // I(u,r) - S(u,r)
          const float Up2 = I[k + 4] - S;           // I(u,r+2) - S(u,r)
          const float B2 = wave_shr1z(A2);          // the same at (u-1, r)
          if (k >= 2) {
            const int t = k - 2;                    // output row v = r - 1
            const float fv = (float)(y0 + t);
            const float Dn = I[k] - S;              // I(u,r-2) - S(u,r): term 2 of row v
            const float G4 = IR[t + 1] - sample(fmaf(h12, fv, g4x), fmaf(h12, fv, g4y),
                                                fmaf(h12, fv, g4z));
            const float G8 = IR[t] - sample(fmaf(h12, fv, g8x), fmaf(h12, fv, g8y),
                                            fmaf(h12, fv, g8z));
            float T = fabsf(A1) + fabsf(B0);        // terms 0, 5
            T += fabsf(B1);                         // term 3
            T += fabsf(wave_shl1z(A2));             // term 6
            T += fabsf(B2);                         // term 7
            T += fabsf(Up0) + fabsf(Dn);            // terms 1, 2
            T += fabsf(G4) + fabsf(G8);             // terms 4, 8
            if (INSIDE) {
              acc[t] += T;
              cnt[t] += 1.0f;
            } else {
              const bool ok = T < INF;
              acc[t] += ok ? T : 0.0f;
              cnt[t] += ok ? 1.0f : 0.0f;
            }
          }
          A1 = A2;
          B0 = B1; B1 = B2;
          Up0 = Up1; Up1 = Up2;
        }
      };
      if ((inside >> s) & 1u)
        pass(std::true_type{});
      else
        pass(std::false_type{});
    }
#pragma unroll
    for (int t = 0; t < TY; ++t) {   // mean over sources of (sum of 9 terms)/9, or -1
      tile[(t * DC + (d - dc0)) * 64 + lane] = cnt[t] > 0.0f ? acc[t] / (9.0f * cnt[t]) : -1.0f;
      if constexpr (CNT) ctile[(t * DC + (d - dc0)) * 64 + lane] = (uint8_t)cnt[t];
    }
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
  if constexpr (CNT) {
    const int items = TY * 62 * dcn;
    for (int it = lane; it < items; it += 64) {
      const int j = it % dcn, rest = it / dcn, px = rest % 62, t = rest / 62;
      const int v = y0 + t, x = x0 + px;
      if (v < H && x < W) cnt_out[((long)v * W + x) * D + dc0 + j] = ctile[(t * DC + j) * 64 + px + 1];
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

}  // namespace pm

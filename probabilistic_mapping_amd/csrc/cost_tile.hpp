// Plane-sweep cost tile body shared by the cost kernels (plain, batched, differentiable) and
// the cost adjoint.  Reference: ADCalcCostKernel, motion_stereo/calc_cost.cu:11-210.
#pragma once

#include <type_traits>

#include "common.hpp"

namespace pm {

// Structured buffer load of one 8-byte record (LLVM's intrinsic for the idxen form of
// buffer_load_dwordx2; clang has no builtin for it): address = base + vindex * stride + voffset.
typedef uint32_t u32x2_t __attribute__((ext_vector_type(2)));
__device__ u32x2_t struct_buffer_load_b64(__amdgpu_buffer_rsrc_t rsrc, int vindex, int voffset,
                                          int soffset, int aux)
    __asm("llvm.amdgcn.struct.ptr.buffer.load.v2i32");

// A measurement frame's quad texture ([H][W] texels of 8 bytes) read as a structured buffer of
// W H records with an 8-byte stride: wave-uniform base and size in scalar registers, a texel
// index per lane (one 24-bit multiply-add from the cell's row and column), and the hardware
// range check (an index past W H returns 0 instead of touching memory).  W H < 2^29 is checked
// on the host.
struct QuadTex {
  __amdgpu_buffer_rsrc_t r;
  __device__ __forceinline__ QuadTex(const uint64_t* q, uint32_t W, uint32_t H)
      : r(__builtin_amdgcn_make_buffer_rsrc(const_cast<uint64_t*>(q), 8, (int)(W * H),
                                            kBufferDword3)) {}
  __device__ __forceinline__ uint2 at(uint32_t texel) const {
    const u32x2_t v = struct_buffer_load_b64(r, (int)texel, 0, 0, 0);
    return make_uint2(v[0], v[1]);
  }
};

// Bilinear interpolation of texel t = (q00, q10 - q00, q01 - q00, q11 - q10 - q01 + q00) at
// weights (ax, ay).  The differences are integers in [-510, 510], exact in FP16.
__device__ __forceinline__ float bilinear(uint2 t, float ax, float ay) {
  const __half2 lo = *reinterpret_cast<const __half2*>(&t.x);
  const __half2 hi = *reinterpret_cast<const __half2*>(&t.y);
  const float top = fmaf(ax, __high2float(lo), __low2float(lo));
  const float dv = fmaf(ax, __high2float(hi), __low2float(hi));
  return fmaf(ay, dv, top);
}

// Bilinear sample of the quad texture at the warped point (U, V), in two halves: gather_uv
// issues the texel load, gather_value consumes it, so a caller can put independent work
// between the two.
// Checked form (INSIDE = false): the warp's validity (calc_cost.cu:85-93) is "the clamp to
// the image changes nothing" and w >= 0, so a NaN warp (0/0) is invalid and gather_value
// returns +inf for it; an invalid warp still reads an in-bounds texel (the clamped position),
// so every gather of a row can be in flight at once.
// Inside form (INSIDE = true): for a point known to be valid and inside the image (see
// tile_inside_mask), the same arithmetic without the clamp and the validity test, so it
// returns the same value.  Coordinates are non-negative there, so truncation is floor, and
// W, y <= 65535 fit the 24-bit multiply.
struct Gather {
  uint2 t;
  float ax, ay;
  bool ok;
};
template <bool INSIDE>
__device__ __forceinline__ Gather gather_uv(const QuadTex& quad, uint32_t W, float Wm1,
                                            float Hm1, float U, float V, float w) {
  Gather g;
  if (INSIDE) {
    g.t = quad.at(__umul24((uint32_t)V, W) + (uint32_t)U);
    g.ax = __builtin_amdgcn_fractf(U);
    g.ay = __builtin_amdgcn_fractf(V);
    g.ok = true;
  } else {
    const float Uc = __builtin_amdgcn_fmed3f(U, 0.0f, Wm1);
    const float Vc = __builtin_amdgcn_fmed3f(V, 0.0f, Hm1);
    g.ok = (w >= 0.0f) & (Uc == U) & (Vc == V);
    g.t = quad.at(__umul24((uint32_t)Vc, W) + (uint32_t)Uc);
    g.ax = __builtin_amdgcn_fractf(Uc);
    g.ay = __builtin_amdgcn_fractf(Vc);
  }
  return g;
}
template <bool INSIDE>
__device__ __forceinline__ float gather_value(const Gather& g) {
  const float s = bilinear(g.t, g.ax, g.ay);
  return (INSIDE || g.ok) ? s : __uint_as_float(0x7F800000u);
}

// Conservative whole-tile validity, per source.  Every warp a cost_tile wave evaluates for one
// (plane, source) is a linear-fractional map of the tile's pixel coordinates: the centre warps
// H (u, r, 1) + t idep over the rectangle u in [u_lo, u_hi], r in [r_lo, r_hi], and the two
// garbled warps (x + c_X, x + c_Y, x + c_Z) over x = h11 u + h12 v + h13 in [x_min, x_max].
// Where the denominator w is positive at the corners it is positive on the whole box, and U, V
// take their extremes at the corners; so if every corner has w >= kInsideW and U, V inside the
// image by kInsideMargin pixels, every warp of the tile is valid with room to spare for FP32
// rounding (< 4e-3 px at w >= kInsideW for |coordinates| < 2^16).  Then the validity tests can
// be skipped: gather_uv<true> returns what gather_uv<false> would.
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

// The garbled warps as a polynomial in rw = 1/w.  A garbled sample point is
// (U, V) = (1 + kx rw, 1 + ky rw) with per-(plane, source) constants kx, ky, and w =
// h11 u + h12 v + cw, so over a block of pixels it moves along one ray from (1, 1).  Where
// every sample of the block lies in one texel cell (i, j), the bilinear value there is, with
// ax = (1 - i) + kx rw and ay = (1 - j) + ky rw,
//   a + bx ax + ay (by + bxy ax) = c0 + c1 rw + c2 rw^2,
//   c0 = a + bx px + py B,  c1 = bx kx + py E + ky B,  c2 = ky E,
//   px = 1 - i, py = 1 - j, B = by + bxy px, E = bxy kx,
// which a pixel evaluates with two FMAs instead of the coordinates, cell index, fractions,
// gather and bilinear.  The bilinear surfaces of neighbouring cells agree on their common
// edge, so a sample that FP32 rounding puts just past the cell is still evaluated to within
// rounding.  garbled_poly tests the cell, and the warps' validity with the same margins as
// tile_inside_mask, at the extremes of w over the block (gmin + cw, gmax + cw, where g =
// h11 u + h12 v over the block's corners; U, V are monotone in w where w > 0), and returns
// the coefficients.
struct GarbledPoly {
  float c0, c1, c2;
  bool ok;
};
__device__ __forceinline__ GarbledPoly garbled_poly(const uint64_t* quad, uint32_t W, float Wm1,
                                                    float Hm1, float gmin, float gmax, float cw,
                                                    float kx, float ky) {
  const float wlo = gmin + cw, whi = gmax + cw;
  const float rlo = __builtin_amdgcn_rcpf(whi), rhi = __builtin_amdgcn_rcpf(wlo);
  const float ua = fmaf(kx, rlo, 1.0f), ub = fmaf(kx, rhi, 1.0f);
  const float va = fmaf(ky, rlo, 1.0f), vb = fmaf(ky, rhi, 1.0f);
  const float umin = fminf(ua, ub), umax = fmaxf(ua, ub);
  const float vmin = fminf(va, vb), vmax = fmaxf(va, vb);
  const float fi = floorf(umin), fj = floorf(vmin);
  GarbledPoly p;
  p.ok = wlo >= kInsideW && umin >= kInsideMargin && umax <= Wm1 - kInsideMargin &&
         vmin >= kInsideMargin && vmax <= Hm1 - kInsideMargin && floorf(umax) == fi &&
         floorf(vmax) == fj;
  const uint32_t texel = p.ok ? (uint32_t)fj * W + (uint32_t)fi : 0u;
  const uint2 t = reinterpret_cast<const uint2*>(quad)[texel];
  const __half2 lo = *reinterpret_cast<const __half2*>(&t.x);
  const __half2 hi = *reinterpret_cast<const __half2*>(&t.y);
  const float a = __low2float(lo), bx = __high2float(lo);
  const float by = __low2float(hi), bxy = __high2float(hi);
  const float px = 1.0f - fi, py = 1.0f - fj;
  const float B = fmaf(bxy, px, by), E = bxy * kx;
  p.c0 = fmaf(py, B, fmaf(bx, px, a));
  p.c1 = fmaf(ky, B, fmaf(py, E, bx * kx));
  p.c2 = ky * E;
  return p;
}

__device__ __forceinline__ uint32_t tile_inside_mask(const InsideCorner& c, float idep,
                                                     float Wm1, float Hm1) {
  const float X = fmaf(c.tx, idep, c.X0), Y = fmaf(c.ty, idep, c.Y0);
  const float Z = fmaf(c.tz, idep, c.Z0);
  // rcp is within 1 ulp: far inside the kInsideMargin / kInsideW slack
  const float rz = __builtin_amdgcn_rcpf(Z);
  const float U = X * rz, V = Y * rz;
  const bool ok = !c.live || (Z >= kInsideW && U >= kInsideMargin && U <= Wm1 - kInsideMargin &&
                              V >= kInsideMargin && V <= Hm1 - kInsideMargin);
  const uint64_t bad = __ballot(!ok);
  uint32_t m = 0;
#pragma unroll
  for (int s = 0; s < 8; ++s) m |= ((bad >> (8 * s)) & 0xFFull) == 0 ? (1u << s) : 0u;
  return m;
}

// Checked sample of the warp (X, Y, w), (U, V) = (X, Y) / w, with its derivatives with
// respect to the warped point: S = B(X/w, Y/w), dS/dX = a, dS/dY = b, dS/dw = c.  The value
// is computed as gather_uv<false> / gather_value compute it.  Inside a texel cell the bilinear slopes are
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
  const uint2 t = quad.at(__umul24((uint32_t)Vc, W) + (uint32_t)Uc);
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
  // The block's 4 dper planes (dper per wave) go in rounds of chunks: in round c the four waves
  // take four adjacent chunks, so their stores of a pixel's planes land side by side at about
  // the same time and leave the L2 as whole write granules (with each wave on a contiguous
  // range of its own, FP16 chunks of 4 planes went out as 8-byte pieces: 4x the volume's bytes).
  const int dper = (D + 4 * npg - 1) / (4 * npg);
  const int bbeg = pg * 4 * dper, bend = min(D, bbeg + 4 * dper);
  const int nfull = dper / DC, rem = dper - nfull * DC;
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
  // The garbled-warp polynomial (garbled_poly) serves the uint8 product forms; the
  // differentiable form keeps the gathers its adjoint differentiates.
  constexpr bool kPoly = std::is_same<RT, uint8_t>::value && !CNT;
  static_assert((DC & (DC - 1)) == 0, "staging chunk must be a power of two");
  for (int c = 0; c < nfull + (rem > 0 ? 1 : 0); ++c) {
  const int csz = c < nfull ? DC : rem;
  const int dc0 = bbeg + c * 4 * DC + wave * csz;
  const int dcn = min(csz, bend - dc0);
  if (dcn <= 0) break;
  // Per-(plane, source) constants of the warps.  At the start of each chunk of DC planes, the
  // lanes compute the records of every (source, plane) pair of the chunk and park each in its
  // plane's own output slot of the tile (16 floats per source from the slot's start, so
  // S <= TY * 64 / 16); a plane's results overwrite its slot once the plane is done.  With
  // x = h11 u + h12 v + h13 and (tx, ty, tz) = t idep:
  //   r0 = (h13 + tx, h23 + ty, h33 + tz, h13 + h31 + tz)   centre warps; w of term 4 at u = 0
  //   r1 = (h13 + h11 - h32 + tz, k4x, k4y, k8x)            w of term 8 at u = 0
  //   r2 = (k8y, c0, c1, c2 of term 4)                      garbled_poly coefficients
  //   r3 = (c0, c1, c2 of term 8, 1 if both polynomials hold on the whole band else 0)
  // Term 4 = (x + h11 + tx, x + h21 + ty, x + h31 + tz) has U = 1 + k4x / w, V = 1 + k4y / w with
  // k4x = h11 - h31 + tx - tz, k4y = h21 - h31 + ty - tz; term 8 = (x + h11 - h12 + tx,
  // x + h11 - h22 + ty, x + h11 - h32 + tz) has k8x = h32 - h12 + tx - tz, k8y = h32 - h22 + ty - tz.
  auto rec = [&](int j, int f) {   // float4 f of plane slot j
    return reinterpret_cast<float4*>(tile + ((f >> 4) * DC + j) * 64 + (f & 15) * 4);
  };
  for (int p = lane; p < a.S * DC; p += 64) {   // pair p: source p / DC, plane dc0 + p % DC
    const int s = p / DC, j = p % DC;
    if (j >= dcn) continue;
    // (h11, h12, h13, h21), (h22, h23, h31, h32), (h33, t1, t2, t3)
    const float4* hp = reinterpret_cast<const float4*>(a.hm[s]);
    const float4 ha = hp[0], hb = hp[1], hc = hp[2];
    const float idep = a.idep_min + (float)(dc0 + j) * a.dep_sample;
    const float tx = hc.y * idep, ty = hc.z * idep, tz = hc.w * idep;
    const float dxz = tx - tz, dyz = ty - tz;
    const float4 r0 = make_float4(ha.z + tx, hb.y + ty, hc.x + tz, (ha.z + hb.z) + tz);
    const float4 r1 = make_float4((ha.z + (ha.x - hb.w)) + tz, (ha.x - hb.z) + dxz,
                                  (ha.w - hb.z) + dyz, (hb.w - ha.y) + dxz);
    float4 r2 = make_float4((hb.w - hb.x) + dyz, 0.0f, 0.0f, 0.0f);
    float4 r3 = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    if constexpr (kPoly) {
      // h11 u + h12 v over the corners of the strip's 8-row band holding the tile (u over all
      // 64 lanes).  The band, not the tile, decides where the polynomial is used and fixes its
      // coefficients, so every tile height (2, 4 or 8 rows) gives the same bits.
      const float ul = (float)(strip * 62 - 1), uh = (float)(strip * 62 + 62);
      const float vl = (float)(y0 & ~7), vh = (float)((y0 & ~7) + 7);
      const float g00 = fmaf(ha.y, vl, ha.x * ul), g10 = fmaf(ha.y, vl, ha.x * uh);
      const float g01 = fmaf(ha.y, vh, ha.x * ul), g11 = fmaf(ha.y, vh, ha.x * uh);
      const float gmin = fminf(fminf(g00, g10), fminf(g01, g11));
      const float gmax = fmaxf(fmaxf(g00, g10), fmaxf(g01, g11));
      const GarbledPoly p4 = garbled_poly(a.quad[s], (uint32_t)W, Wm1, Hm1, gmin, gmax, r0.w,
                                          r1.y, r1.z);
      const GarbledPoly p8 = garbled_poly(a.quad[s], (uint32_t)W, Wm1, Hm1, gmin, gmax, r1.x,
                                          r1.w, r2.x);
      r2.y = p4.c0; r2.z = p4.c1; r2.w = p4.c2;
      r3 = make_float4(p8.c0, p8.c1, p8.c2, (p4.ok && p8.ok) ? 1.0f : 0.0f);
    }
    *rec(j, 4 * s) = r0;
    *rec(j, 4 * s + 1) = r1;
    *rec(j, 4 * s + 2) = r2;
    *rec(j, 4 * s + 3) = r3;
  }
  __builtin_amdgcn_wave_barrier();
  for (int d = dc0; d < dc0 + dcn; ++d) {
    const float idep = a.idep_min + (float)d * a.dep_sample;
    const int j = d - dc0;
    const uint32_t inside = tile_inside_mask(corner, idep, Wm1, Hm1);
    float acc[TY], cnt[TY];
    int n_inside = 0;   // sources that took the inside pass (each counts on every row)
#pragma unroll
    for (int t = 0; t < TY; ++t) { acc[t] = 0.0f; cnt[t] = 0.0f; }
    for (int s = 0; s < a.S; ++s) {
      const float* hm = a.hm[s];
      const float h11 = hm[0], h12 = hm[1], h21 = hm[3], h22 = hm[4], h31 = hm[6], h32 = hm[7];
      const float4 r0 = *rec(j, 4 * s), r1 = *rec(j, 4 * s + 1), r2 = *rec(j, 4 * s + 2);
      const QuadTex q(a.quad[s], (uint32_t)W, (uint32_t)H);
      // column parts of H (u, r, 1) + t idep, and of the garbled warps' w
      const float Xu = fmaf(fu, h11, r0.x);
      const float Yu = fmaf(fu, h21, r0.y);
      const float Zu = fmaf(fu, h31, r0.z);
      const float w4u = fmaf(fu, h11, r0.w);
      const float w8u = fmaf(fu, h11, r1.x);
      const float k4x = r1.y, k4y = r1.z, k8x = r1.w, k8y = r2.x;
      // one pass over the centre rows; INSIDE: every warp of the tile is valid (no tests);
      // POLY: the garbled terms are garbled_poly's polynomials (valid on the whole tile)
      auto pass = [&](auto inside_tag, auto poly_tag) {
        constexpr bool INSIDE = decltype(inside_tag)::value;
        constexpr bool POLY = decltype(poly_tag)::value;
        float4 r3 = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        if constexpr (POLY) r3 = *rec(j, 4 * s + 3);
        auto centre = [&](int k) {   // the centre warp of row y0 - 1 + k
          const float fr = (float)(y0 - 1 + k);
          const float w = fmaf(h32, fr, Zu);
          const float rw = __builtin_amdgcn_rcpf(w);
          return gather_uv<INSIDE>(q, (uint32_t)W, Wm1, Hm1, fmaf(h12, fr, Xu) * rw,
                                   fmaf(h22, fr, Yu) * rw, w);
        };
        auto garbled = [&](int t, float wu, float kx, float ky) {   // output row y0 + t
          const float w = fmaf(h12, (float)(y0 + t), wu);
          const float rw = __builtin_amdgcn_rcpf(w);
          return gather_uv<INSIDE>(q, (uint32_t)W, Wm1, Hm1, fmaf(kx, rw, 1.0f),
                                   fmaf(ky, rw, 1.0f), w);
        };
        auto garbled_value = [&](int t, float wu, float c0, float c1, float c2) {
          const float rw = __builtin_amdgcn_rcpf(fmaf(h12, (float)(y0 + t), wu));
          return fmaf(fmaf(c2, rw, c1), rw, c0);
        };
        // Row k's gathers are issued one row ahead (before row k-1's arithmetic), and the
        // scheduling barrier keeps the compiler from sinking them back to their uses: each
        // row's loads are in flight while the previous row computes.
        Gather gc[TY + 2], g4[TY], g8[TY];
        gc[0] = centre(0);
        // window over centre rows r-2, r-1: I-S (A0, A1) and the up terms (Up0, Up1)
        float A0 = 0.0f, A1 = 0.0f, Up0 = 0.0f, Up1 = 0.0f;
#pragma unroll
        for (int k = 0; k < TY + 2; ++k) {   // centre row r = y0 - 1 + k
          if (k + 1 < TY + 2) {
            gc[k + 1] = centre(k + 1);
            if (!POLY && k >= 1) {
              g4[k - 1] = garbled(k - 1, w4u, k4x, k4y);
              g8[k - 1] = garbled(k - 1, w8u, k8x, k8y);
            }
          }
          __builtin_amdgcn_sched_barrier(0);
          const float S = gather_value<INSIDE>(gc[k]);
          // signed differences; the |.| is taken where they are summed (a source modifier)
          const float A2 = I[k + 2] - S;            // I(u,r) - S(u,r)
          const float Up2 = I[k + 4] - S;           // I(u,r+2) - S(u,r)
          if (k >= 2) {
            const int t = k - 2;                    // output row v = r - 1
            const float Dn = I[k] - S;              // I(u,r-2) - S(u,r): term 2 of row v
            const float G4 = IR[t + 1] - (POLY ? garbled_value(t, w4u, r2.y, r2.z, r2.w)
                                               : gather_value<INSIDE>(g4[t]));
            const float G8 = IR[t] - (POLY ? garbled_value(t, w8u, r3.x, r3.y, r3.z)
                                           : gather_value<INSIDE>(g8[t]));
            // terms 5, 3, 7 are the left neighbour's |A| over rows v-1, v, v+1: one column sum
            // moved one lane right
            const float col = fabsf(A0) + fabsf(A1) + fabsf(A2);
            float T = wave_shr1z(col) + fabsf(A1);  // terms 3, 5, 7, 0
            T += fabsf(wave_shl1z(A2));             // term 6
            T += fabsf(Up0) + fabsf(Dn);            // terms 1, 2
            T += fabsf(G4) + fabsf(G8);             // terms 4, 8
            if (INSIDE) {
              acc[t] += T;
            } else {
              const bool ok = T < INF;
              acc[t] += ok ? T : 0.0f;
              cnt[t] += ok ? 1.0f : 0.0f;
            }
          }
          A0 = A1; A1 = A2;
          Up0 = Up1; Up1 = Up2;
        }
      };
      const bool ins = (inside >> s) & 1u;
      // the record's polynomial flag, read by every lane from one LDS word: wave-uniform
      const bool gp =
          kPoly && __builtin_amdgcn_readfirstlane(__float_as_uint(rec(j, 4 * s + 3)->w)) != 0u;
      if (gp) {
        if constexpr (kPoly) {
          if (ins) pass(std::true_type{}, std::true_type{});
          else pass(std::false_type{}, std::true_type{});
        }
      } else {
        if (ins) pass(std::true_type{}, std::false_type{});
        else pass(std::false_type{}, std::false_type{});
      }
      n_inside += ins ? 1 : 0;
    }
#pragma unroll
    for (int t = 0; t < TY; ++t) cnt[t] += (float)n_inside;
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

// One-wave-per-scanline SGM body shared by the plain, batched and differentiable kernels.
// Reference: sgm2<idx,start,dx,dy,n>, motion_stereo/sgm.cu:5-113.
#pragma once

#include <type_traits>

#include "common.hpp"

namespace pm {

__device__ __forceinline__ uint32_t ref_bits(uint8_t v) { return v; }
__device__ __forceinline__ uint32_t ref_bits(float v) { return __float_as_uint(v); }

// A cost element as the scan holds it between its load and its use.  The recurrence adds it
// as fma(c, w, x) with a wave-uniform w: 1 at a valid pixel (c * 1 + x, one rounding, the
// value of c + x) and 0 at an invalid one (x, the reference's COST_INF = 0 in place of c;
// costs are finite).  w comes from a scalar register the compiler cannot see into, so the
// FMA stays an FMA: for FP16 costs, held as raw 16-bit values, it becomes v_fma_mix and the
// conversion disappears; for every type the per-lane select of the cost disappears.
template <typename CT>
struct CostSlot {
  using T = float;
  static __device__ __forceinline__ T load(const Buf& b, uint32_t voff, uint32_t soff) {
    return buf_cost<CT>(b, voff, soff);
  }
  static __device__ __forceinline__ bool negative(T c) { return c < 0.0f; }
  static __device__ __forceinline__ float value(T c) { return c; }
  static __device__ __forceinline__ float add(T c, float x, float w) { return fmaf(c, w, x); }
};
template <>
struct CostSlot<__half> {
  using T = _Float16;
  static __device__ __forceinline__ T load(const Buf& b, uint32_t voff, uint32_t soff) {
    const unsigned short u = (unsigned short)b.u16(voff, soff);
    T h;
    __builtin_memcpy(&h, &u, 2);
    return h;
  }
  static __device__ __forceinline__ float value(T c) { return (float)c; }
  static __device__ __forceinline__ bool negative(T c) { return c < (T)0; }
  static __device__ __forceinline__ float add(T c, float x, float w) {
    return fmaf((float)c, w, x);
  }
};

__device__ __forceinline__ float opaque_one() {
  float one;
  asm("s_mov_b32 %0, 1.0" : "=s"(one));
  return one;
}

// Wave-uniform selects on the scalar unit.  Written as C++ ternaries these became per-lane
// v_cndmask instructions on the (issue-bound) vector unit in every scan step; the scalar unit
// issues beside it.  All operands are wave-uniform and live in scalar registers.
//   penalties: (P1s, P2s) where bit `bit` of `smooth` is set (|dI| < tau_so), else
//   (P1n, P2n), sgm.cu:78-82.
__device__ __forceinline__ void scalar_penalties(uint64_t smooth, uint32_t bit, float P1s,
                                                 float P1n, float P2s, float P2n, float& P1,
                                                 float& P2) {
  asm("s_bitcmp1_b64 %2, %3\n\t"
      "s_cselect_b32 %0, %4, %5\n\t"
      "s_cselect_b32 %1, %6, %7"
      : "=&s"(P1), "=&s"(P2)
      : "s"(smooth), "s"(bit), "s"(P1s), "s"(P1n), "s"(P2s), "s"(P2n)
      : "scc");
}
//   cost weight: 0 where the ballot of negative costs is non-zero (an invalid pixel adds
//   COST_INF = 0 in place of its cost), else `one`.
__device__ __forceinline__ float scalar_cost_weight(uint64_t neg_ballot, float one) {
  float w;
  asm("s_cmp_lg_u64 %1, 0\n\t"
      "s_cselect_b32 %0, 0, %2"
      : "=&s"(w)
      : "s"(neg_ballot), "s"(one)
      : "scc");
  return w;
}

// a + b on the scalar unit, opaque to the optimiser.  The scan advances its pixel offset by
// one step's stride per step; written as plain adds, the compiler folds them into
// base + i * stride for each of the unrolled block's slots, keeps those constants in scalar
// registers and, out of registers, spills them to vector lanes that every step must read
// back with a vector instruction.  Both operands are wave-uniform; readfirstlane states it
// where the compiler cannot prove it (it is free on a value already in a scalar register).
__device__ __forceinline__ uint32_t scalar_add(uint32_t a, uint32_t b) {
  uint32_t r;
  asm("s_add_u32 %0, %1, %2"
      : "=s"(r)
      : "s"(__builtin_amdgcn_readfirstlane(a)), "s"(__builtin_amdgcn_readfirstlane(b))
      : "scc");
  return r;
}

// Stores of one step's NV planes per lane (plane off[j] of the pixel at element offset soff /
// element size), with cache policy AUX (2 = non-temporal).  PACK: the planes are contiguous
// and the pixel's first element is aligned to NV elements, so a lane writes them with one
// 4/8/16-byte store and the wave's stores cover one contiguous span; written one element per
// instruction, each store instruction covers every NV-th element and the memory system moves
// whole granules for it (measured: 1.45-1.9x the bytes of the volume with two planes per
// lane).  Lanes past D (act false) store nothing; with PACK all NV planes of a lane are
// active or none is.
template <int NV, int AUX>
__device__ __forceinline__ void store_planes_f32(const Buf& b, const float (&v)[NV],
                                                 const bool (&act)[NV], bool pack,
                                                 const int (&off)[NV], uint32_t soff) {
  if (NV == 2 && pack) {
    if (act[0]) {
      const u32x2 x = {__float_as_uint(v[0]), __float_as_uint(v[NV > 1 ? 1 : 0])};
      b.put_b64<AUX>(x, (uint32_t)off[0] * 4u, soff);
    }
  } else if (NV == 4 && pack) {
    if (act[0]) {
      const u32x4 x = {__float_as_uint(v[0]), __float_as_uint(v[NV > 1 ? 1 : 0]),
                       __float_as_uint(v[NV > 2 ? 2 : 0]), __float_as_uint(v[NV > 3 ? 3 : 0])};
      b.put_b128<AUX>(x, (uint32_t)off[0] * 4u, soff);
    }
  } else {
#pragma unroll
    for (int j = 0; j < NV; ++j)
      if (act[j]) b.put_b32<AUX>(__float_as_uint(v[j]), (uint32_t)off[j] * 4u, soff);
  }
}
template <int NV, int AUX>
__device__ __forceinline__ void store_planes_u16(const Buf& b, const uint32_t (&v)[NV],
                                                 const bool (&act)[NV], bool pack,
                                                 const int (&off)[NV], uint32_t soff) {
  if (NV == 2 && pack) {
    if (act[0]) b.put_b32<AUX>(v[0] | (v[NV > 1 ? 1 : 0] << 16), (uint32_t)off[0] * 2u, soff);
  } else if (NV == 4 && pack) {
    if (act[0]) {
      const u32x2 x = {v[0] | (v[NV > 1 ? 1 : 0] << 16),
                       v[NV > 2 ? 2 : 0] | (v[NV > 3 ? 3 : 0] << 16)};
      b.put_b64<AUX>(x, (uint32_t)off[0] * 2u, soff);
    }
  } else {
#pragma unroll
    for (int j = 0; j < NV; ++j)
      if (act[j]) {
        if (AUX == 2) b.put_u16_nt(v[j], (uint32_t)off[j] * 2u, soff);
        else b.put_u16(v[j], (uint32_t)off[j] * 2u, soff);
      }
  }
}

__device__ __forceinline__ void scanline_start(const SgmArgs& a, int i, int& x, int& y) {
  if (a.dx != 0) {
    const int xs = a.dx > 0 ? 0 : a.W - 1;
    if (i < a.H) { x = xs; y = i; return; }
    const int j = i - a.H;
    x = a.dx > 0 ? 1 + j : j;
    y = a.dy > 0 ? 0 : a.H - 1;
    return;
  }
  x = i;
  y = a.dy > 0 ? 0 : a.H - 1;
}

// One wave per scanline.  All scanline geometry is wave-uniform (scalar registers); every
// load address is clamped in bounds (past-the-end steps re-read the last pixel, lanes past D
// re-read plane D-1 of the same pixel), so the prefetch has no branches and the B loads of
// the next block stay in flight while the current block computes.
// RT: reference element type (uint8_t, or float for the differentiable form).  AGG: accumulate
// into agg (else agg is untouched).  LOUT: also write this path's L to a.lout; with LMASK the
// stored value is 0 at invalid pixels (the path's contribution to the aggregate).  EQ: write
// the path's excess E = L - C' to a.eout as 16-bit fixed point; L itself is computed exactly as
// in the other forms.  The per-path volumes (LOUT, EQ) are written once and read once by the
// combine pass, so they go out with the non-temporal cache policy; the aggregate does not.
template <int NV, typename CT, bool FIRST, bool FULL, int B, typename RT = uint8_t,
          bool AGG = true, bool LOUT = false, bool LMASK = false, bool EQ = false>
__device__ __forceinline__ void sgm_scanline(const SgmArgs& a, int line,
                                             const CT* __restrict__ cost,
                                             float* __restrict__ agg) {
  const gptr<const RT> ref = to_global(static_cast<const RT*>(a.ref));
  // volumes through buffer resources: per step one uniform byte offset (the pixel's planes),
  // per lane a constant one (its planes); the host keeps W H D 4 bytes below 4 GiB
  const uint32_t V = (uint32_t)a.W * (uint32_t)a.H * (uint32_t)a.D;
  const Buf bcost(cost, V * (uint32_t)sizeof(CT));
  const Buf bagg(agg, AGG ? V * 4u : 0u);
  const Buf blout(a.lout, LOUT ? V * 4u : 0u);
  const Buf beout(a.eout, EQ ? V * 2u : 0u);
  const int lane = threadIdx.x & 63;
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
  // byte offsets of the lane's planes in a volume of elt-byte elements
  auto lb = [&](int j, uint32_t elt) { return (uint32_t)off[j] * elt; };
  // a lane's NV planes are contiguous; with D a multiple of NV each pixel's planes start on a
  // multiple of NV elements, so they go out as one aligned wide store
  const bool pack = NV > 1 && D % NV == 0;
  const float INF = __uint_as_float(0x7F800000u);

  // two register buffers of B steps each, used ping-pong: while one block computes, the
  // loads of the next are in flight (addresses past the end re-read the last pixel)
  using CS = CostSlot<CT>;
  using CV = typename CS::T;
  const float one = opaque_one();
  CV c0[B][NV], c1[B][NV];
  float g0[B][NV], g1[B][NV];
  uint32_t r0, r1;
  auto load_block = [&](int base, CV (&c)[B][NV], float (&g)[B][NV], uint32_t& r) {
    // lane i fetches the reference intensity of step base+i; issued first so that reading it
    // does not wait for the block's cost and aggregate loads
    r = ref_bits(ref[p0 + min(base + lane, n - 1) * pstep]);
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int i = 0; i < B; ++i) {
      const uint32_t p = (uint32_t)(p0 + min(base + i, n - 1) * pstep) * (uint32_t)D;
#pragma unroll
      for (int j = 0; j < NV; ++j) {
        c[i][j] = CS::load(bcost, lb(j, sizeof(CT)), p * (uint32_t)sizeof(CT));
        if (!FIRST && AGG) g[i][j] = bagg.f32(lb(j, 4), p * 4u);
      }
    }
  };

  float L[NV];
  uint32_t iq;        // intensity of the last pixel of the previous block
  uint64_t smooth;    // bit i: |I(step i) - I(step i-1)| < tau_so for the current block
  // gradient test of a whole block at once (sgm.cu:75-78): lane i compares step i with step
  // i-1 (lane 0 with the carried intensity) and a ballot packs the answers
  auto smooth_mask = [&](uint32_t r) {
    const uint32_t prev = dpp_u32<0x138>(iq, r);   // wave_shr:1, lane 0 takes iq
    iq = __builtin_amdgcn_readlane(r, B - 1);
    if constexpr (std::is_same<RT, float>::value)
      return (uint64_t)__ballot(fabsf(__uint_as_float(r) - __uint_as_float(prev)) < a.tau_f);
    else
      return (uint64_t)__ballot(abs((int)r - (int)prev) < a.tau_i);
  };
  // one step k >= 1 from slot i of a block buffer
  // L(q, d0-1) of the step, lane 0 = 0xFFFFFFFF throughout (see shr1_max_carry)
  uint32_t lo_carry = 0xFFFFFFFFu;
  // element offset of the current step's pixel, advanced by one stride per step
  uint32_t pk = (uint32_t)p0 * (uint32_t)D;
  const uint32_t pkstep = (uint32_t)pstep * (uint32_t)D;
  auto step = [&](int, int i, const CV (&c)[B][NV], const float (&g)[B][NV]) {
    pk = scalar_add(pk, pkstep);
    const uint32_t p = pk;
    bool neg = false;
#pragma unroll
    for (int j = 0; j < NV; ++j) neg = neg || CS::negative(c[i][j]);
    const uint64_t negb = __ballot(neg);
    const bool invalid = negb != 0;
    static_assert(kCostInf == 0.0f, "an invalid pixel adds 0 in place of its cost");
    const float w = scalar_cost_weight(negb, one);
    // m = min_d L(q, d) (sgm.cu:55-66)
    float lm = L[0];
#pragma unroll
    for (int j = 1; j < NV; ++j) lm = umin(lm, L[j]);
    const float m = wave_min_nonneg(lm);
    // gradient-adaptive penalties (sgm.cu:78-82), scalar
    float P1, P2;
    scalar_penalties(smooth, (uint32_t)i, a.P1s, a.P1n, a.P2s, a.P2n, P1, P2);
    // L(q, d0-1) and L(q, d0+NV) from the neighbouring lanes (0xFFFFFFFF where none)
    lo_carry = shr1_max_carry(lo_carry, L[NV - 1]);
    const uint32_t lo = lo_carry;
    const uint32_t hi = shl1_max(L[0]);
    const float mP2 = m + P2;
    float Ln[NV], out[NV], lo_[NV];
    uint32_t q[NV];
#pragma unroll
    for (int j = 0; j < NV; ++j) {
      const uint32_t left = j > 0 ? __float_as_uint(L[j - 1]) : lo;
      const uint32_t right = j + 1 < NV ? __float_as_uint(L[j + 1]) : hi;
      // min(L(q,d), L(q,d-1)+P1, L(q,d+1)+P1, m+P2) (sgm.cu:85-93)
      const float val = umin(umin(L[j], mP2), __uint_as_float(min(left, right)) + P1);
      Ln[j] = act[j] ? CS::add(c[i][j], val, w) - m : INF;   // sgm.cu:98
      out[j] = invalid ? kCostInf : (FIRST ? Ln[j] : g[i][j] + Ln[j]);   // sgm.cu:101-108
      lo_[j] = (LMASK && invalid) ? kCostInf : Ln[j];
      // val >= m exactly (every candidate is >= m), so E >= 0; rounding to nearest.  At an
      // invalid pixel the level is ignored (the combine finds the pixel from its cost).
      if (EQ) q[j] = excess_level_bits(val - m, a.eq_inv) & 0xFFFFu;
    }
    if (AGG) store_planes_f32<NV, 0>(bagg, out, act, pack, off, p * 4u);
    if (LOUT) store_planes_f32<NV, 2>(blout, lo_, act, pack, off, p * 4u);
    if (EQ) store_planes_u16<NV, 2>(beout, q, act, pack, off, p * 2u);
#pragma unroll
    for (int j = 0; j < NV; ++j) L[j] = Ln[j];
  };
  // a block whose B steps all exist: no exits, so the loads of the other buffer stay in flight
  auto run_full = [&](int base, const CV (&c)[B][NV], const float (&g)[B][NV], uint32_t r) {
    smooth = smooth_mask(r);
#pragma unroll
    for (int i = 0; i < B; ++i) step(base + i, i, c, g);
  };
  auto run_checked = [&](int base, const CV (&c)[B][NV], const float (&g)[B][NV],
                         uint32_t r) {
    smooth = smooth_mask(r);
#pragma unroll
    for (int i = 0; i < B; ++i) {
      if (base + i >= n) break;
      step(base + i, i, c, g);
    }
  };

  // step 0: L = C' (sgm.cu:21-43)
  {
    const uint32_t p = (uint32_t)p0 * (uint32_t)D;
    float c[NV], g[NV];
#pragma unroll
    for (int j = 0; j < NV; ++j) {
      c[j] = buf_cost<CT>(bcost, lb(j, sizeof(CT)), p * (uint32_t)sizeof(CT));
      g[j] = (FIRST || !AGG) ? 0.0f : bagg.f32(lb(j, 4), p * 4u);
    }
    iq = ref_bits(ref[p0]);   // intensity of step 0
    bool neg = false;
#pragma unroll
    for (int j = 0; j < NV; ++j) neg = neg || c[j] < 0.0f;
    const bool invalid = __ballot(neg) != 0;
#pragma unroll
    for (int j = 0; j < NV; ++j) {
      L[j] = act[j] ? (invalid ? kCostInf : c[j]) : INF;
      const float out = invalid ? kCostInf : (FIRST ? L[j] : g[j] + L[j]);
      if (AGG && act[j]) bagg.put_f32(out, lb(j, 4), p * 4u);
      if (LOUT && act[j])
        blout.put_f32_nt((LMASK && invalid) ? kCostInf : L[j], lb(j, 4), p * 4u);
      if (EQ && act[j]) beout.put_u16_nt(0u, lb(j, 2), p * 2u);
    }
  }
  int base = 1;
  load_block(base, c0, g0, r0);
  while (base + 2 * B <= n) {
    load_block(base + B, c1, g1, r1);
    run_full(base, c0, g0, r0);
    load_block(base + 2 * B, c0, g0, r0);
    run_full(base + B, c1, g1, r1);
    base += 2 * B;
  }
  run_checked(base, c0, g0, r0);
  if (base + B < n) {
    load_block(base + B, c1, g1, r1);
    run_checked(base + B, c1, g1, r1);
  }
}

// ---- Four scanlines per wave (compact form, 16-bit costs) -------------------------------
// A wave runs four scanlines of one path side by side: lanes 16r .. 16r+15 (row r of the
// wave) hold scanline line0 + r, lane k of a row holding planes NV*k .. NV*k+NV-1 (D = 16 NV).
// Against one scanline per wave (lane = plane) a step's minimum over planes is NV-1 in-lane
// minima plus four DPP stages that leave the row's minimum in every lane of the row (no
// cross-row broadcast, no readlane), and the per-step selects and shifts are shared by NV
// planes.  The recurrence and the stored levels are those of sgm_scanline<..., EQ = true>:
// the same operations in the same order on every element.

// min over the 16 lanes of each row, in every lane of the row
__device__ __forceinline__ uint32_t row_min16(uint32_t u) {
  u = umin_dpp<0xB1>(u);    // quad_perm [1,0,3,2]
  u = umin_dpp<0x4E>(u);    // quad_perm [2,3,0,1]
  u = umin_dpp<0x124>(u);   // row_ror:4
  u = umin_dpp<0x128>(u);   // row_ror:8
  return u;
}
// lane I of each row, in every lane of the row (row_newbcast:I)
template <int I>
__device__ __forceinline__ uint32_t row_bcast(uint32_t v) {
  return (uint32_t)__builtin_amdgcn_mov_dpp((int)v, 0x150 + I, 0xF, 0xF, true);
}
// row_bcast<i> for a slot index of a fully unrolled loop (the switch folds away)
template <int B>
__device__ __forceinline__ uint32_t row_bcast_at(uint32_t v, int i) {
  switch (i) {
    case 0: return row_bcast<0>(v);
    case 1: return row_bcast<1 % B>(v);
    case 2: return row_bcast<2 % B>(v);
    case 3: return row_bcast<3 % B>(v);
    case 4: return row_bcast<4 % B>(v);
    case 5: return row_bcast<5 % B>(v);
    case 6: return row_bcast<6 % B>(v);
    case 7: return row_bcast<7 % B>(v);
    case 8: return row_bcast<8 % B>(v);
    case 9: return row_bcast<9 % B>(v);
    case 10: return row_bcast<10 % B>(v);
    case 11: return row_bcast<11 % B>(v);
    case 12: return row_bcast<12 % B>(v);
    case 13: return row_bcast<13 % B>(v);
    case 14: return row_bcast<14 % B>(v);
    default: return row_bcast<15 % B>(v);
  }
}
// Bits 16r .. 16r+15 set where any bit of row r is set in b (wave-uniform, scalar unit).
__device__ __forceinline__ uint64_t row_any_mask(uint64_t b) {
  b |= b >> 1;
  b |= b >> 2;
  b |= b >> 4;
  b |= b >> 8;
  b &= 0x0001000100010001ull;
  return (b << 16) - b;
}
// 0 in the lanes whose bit of `mask` is set, else 1.0; opaque to the optimiser, so the cost's
// fma(c, w, x) stays an fma (v_fma_mix for FP16 costs).
__device__ __forceinline__ float lane_weight(uint64_t mask) {
  float w;
  asm("v_cndmask_b32_e64 %0, 1.0, 0, %1" : "=v"(w) : "s"(mask));
  return w;
}

typedef _Float16 h16x2 __attribute__((ext_vector_type(2)));

// One lane's NV 16-bit costs of one step, as NV/2 dwords (plane 2i in the low half of word i).
template <int NW>
struct CostWords {
  uint32_t w[NW];
};
template <int NW>
__device__ __forceinline__ CostWords<NW> load_words(const Buf& b, uint32_t voff, uint32_t soff) {
  CostWords<NW> c;
  if constexpr (NW == 1) {
    c.w[0] = __builtin_amdgcn_raw_buffer_load_b32(b.r, (int)voff, (int)soff, 0);
  } else if constexpr (NW == 2) {
    const u32x2 v = __builtin_amdgcn_raw_buffer_load_b64(b.r, (int)voff, (int)soff, 0);
    c.w[0] = v.x; c.w[1] = v.y;
  } else if constexpr (NW == 3) {
    const auto v = __builtin_amdgcn_raw_buffer_load_b96(b.r, (int)voff, (int)soff, 0);
    c.w[0] = v[0]; c.w[1] = v[1]; c.w[2] = v[2];
  } else {
    static_assert(NW == 4, "2 to 8 planes per lane");
    const u32x4 v = __builtin_amdgcn_raw_buffer_load_b128(b.r, (int)voff, (int)soff, 0);
    c.w[0] = v[0]; c.w[1] = v[1]; c.w[2] = v[2]; c.w[3] = v.w;
  }
  return c;
}
template <int NW, int AUX>
__device__ __forceinline__ void store_words(const Buf& b, const uint32_t (&w)[NW], uint32_t voff,
                                            uint32_t soff) {
  if constexpr (NW == 1) {
    __builtin_amdgcn_raw_buffer_store_b32(w[0], b.r, (int)voff, (int)soff, AUX);
  } else if constexpr (NW == 2) {
    const u32x2 v = {w[0], w[1]};
    __builtin_amdgcn_raw_buffer_store_b64(v, b.r, (int)voff, (int)soff, AUX);
  } else if constexpr (NW == 3) {
    const u32x3 v = {w[0], w[1], w[2]};
    b.put_b96<AUX>(v, voff, soff);
  } else {
    const u32x4 v = {w[0], w[1], w[2], w[3]};
    b.put_b128<AUX>(v, voff, soff);
  }
}

// Whether any 16-bit half of the lane's cost words, plus 0x7FFF modulo 2^16, is at most limit.
template <int NW>
__device__ __forceinline__ bool any_below(const CostWords<NW>& c, uint32_t limit);

// Plane j of a lane's cost words, as float, and whether any of the lane's planes is negative.
template <typename CT>
struct Cost16;
template <>
struct Cost16<__half> {
  template <int NW>
  static __device__ __forceinline__ float plane(const CostWords<NW>& c, int j) {
    h16x2 h;
    __builtin_memcpy(&h, &c.w[j >> 1], 4);
    return (float)((j & 1) ? h.y : h.x);
  }
  // Negative FP16 values (-0 and NaNs excluded, as for c < 0) are the bit patterns 0x8001 ..
  // 0xFC00 (-inf); adding 0x7FFF modulo 2^16 maps exactly those to 0 .. 0x7BFF, so one packed
  // add per word, a packed minimum and one compare test all of a lane's planes.
  template <int NW>
  static __device__ __forceinline__ bool any_negative(const CostWords<NW>& c) {
    return any_below<NW>(c, 0x7BFFu);
  }
};
template <int NW>
__device__ __forceinline__ bool any_below(const CostWords<NW>& c, uint32_t limit) {
  typedef unsigned short u16x2 __attribute__((ext_vector_type(2)));
  const u16x2 bias = {0x7FFF, 0x7FFF};
  u16x2 m;
  __builtin_memcpy(&m, &c.w[0], 4);
  m += bias;   // v_pk_add_u16, modulo 2^16
#pragma unroll
  for (int i = 1; i < NW; ++i) {
    u16x2 h;
    __builtin_memcpy(&h, &c.w[i], 4);
    h += bias;
    m = __builtin_elementwise_min(m, h);   // v_pk_min_u16
  }
  return (uint32_t)min(m.x, m.y) <= limit;
}
template <>
struct Cost16<__hip_bfloat16> {
  template <int NW>
  static __device__ __forceinline__ float plane(const CostWords<NW>& c, int j) {
    return __uint_as_float((j & 1) ? (c.w[j >> 1] & 0xFFFF0000u) : (c.w[j >> 1] << 16));
  }
  // negative BF16 values: 0x8001 .. 0xFF80 (-inf), mapped to 0 .. 0x7F7F as for FP16
  template <int NW>
  static __device__ __forceinline__ bool any_negative(const CostWords<NW>& c) {
    return any_below<NW>(c, 0x7F7Fu);
  }
};

// Scanlines line0 .. line0+3 of path (a.dx, a.dy), compact form: each writes its excess
// levels to a.eout.  Requires D == 16 NV with NV even (2..8) and a 16-bit cost type.
// Scanlines past a.nlines (the last wave of a path) compute a copy of the last scanline and
// store nothing.  Blocks of B steps in which every row's scanline still runs use one
// wave-uniform pixel offset per step (the row with the lowest start pixel's) plus a constant
// per-lane offset; the other blocks address every lane on its own, clamped to its scanline.
template <int NV, typename CT, int B>
__device__ __forceinline__ void sgm_rows_excess(const SgmArgs& a, int line0,
                                                const CT* __restrict__ cost) {
  static_assert(sizeof(CT) == 2 && NV % 2 == 0 && NV >= 2 && NV <= 8, "16-bit costs, NV even");
  static_assert(B >= 2 && B <= 16, "one block step per row lane");
  constexpr int NW = NV / 2;
  using CC = Cost16<CT>;
  const gptr<const uint8_t> ref = to_global(static_cast<const uint8_t*>(a.ref));
  const uint32_t D = (uint32_t)a.D;
  const uint32_t V = (uint32_t)a.W * (uint32_t)a.H * D;
  const Buf bcost(cost, V * 2u);
  const Buf beout(a.eout, V * 2u);
  const int lane = threadIdx.x & 63;
  const int k = lane & 15;
  const int row_line = line0 + (lane >> 4);
  const bool live = row_line < a.nlines;
  int x, y;
  scanline_start(a, live ? row_line : a.nlines - 1, x, y);
  int n = 1 << 30;
  if (a.dx > 0) n = min(n, a.W - x);
  if (a.dx < 0) n = min(n, x + 1);
  if (a.dy > 0) n = min(n, a.H - y);
  if (a.dy < 0) n = min(n, y + 1);
  const int nst = live ? n : 0;   // steps this row stores
  const int pstep = a.dy * a.W + a.dx;
  const int p0 = y * a.W + x;
  // wave-uniform extents: every row runs its steps < nmin; the wave runs to nmax
  int nmin = 1 << 30, nmax = 0, pbase = 1 << 30;
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    const int nr = __builtin_amdgcn_readlane(nst, 16 * r);
    nmin = min(nmin, nr);
    nmax = max(nmax, nr);
    pbase = min(pbase, __builtin_amdgcn_readlane(p0, 16 * r));
  }
  const uint32_t lofs = (uint32_t)k * (uint32_t)NV * 2u;             // the lane's planes
  const uint32_t voff = (uint32_t)(p0 - pbase) * D * 2u + lofs;      // ... in its row's pixel
  // byte offset of step s of this row, clamped to its scanline (own addressing)
  auto own = [&](int s) { return (uint32_t)(p0 + min(s, n - 1) * pstep) * D * 2u + lofs; };

  float L[NV], Lq[NV];
  uint32_t lo = 0xFFFFFFFFu, hi = 0xFFFFFFFFu;   // row-edge lanes keep these (see below)
  uint32_t Iprev;                                // intensity of the row's previous step
  // step 0: L = C' (sgm.cu:21-43); the stored excess is 0
  {
    const CostWords<NW> c = load_words<NW>(bcost, own(0), 0u);
    Iprev = (uint32_t)ref[p0];
    const uint64_t inv = row_any_mask(__ballot(CC::any_negative(c)));
    const bool invalid = (inv >> lane) & 1;
#pragma unroll
    for (int j = 0; j < NV; ++j) L[j] = invalid ? kCostInf : CC::plane(c, j);
    const uint32_t z[NW] = {};
    if (nst > 0) store_words<NW, 2>(beout, z, own(0), 0u);
  }

  CostWords<NW> c0[B], c1[B];
  uint32_t P1v, P2v;   // lane k: the penalties of step base + k of the row's current block
  // penalties of block [base, base + B): lane k compares the intensity of step base + k with
  // step base + k - 1 (lane 0 with the carried intensity), sgm.cu:75-82
  auto penalties = [&](int base) {
    const uint32_t I = (uint32_t)ref[p0 + min(base + min(k, B - 1), n - 1) * pstep];
    const uint32_t prev = (uint32_t)__builtin_amdgcn_update_dpp((int)Iprev, (int)I, 0x111, 0xF,
                                                                0xF, false);   // row_shr:1
    Iprev = row_bcast<B - 1>(I);
    const bool smooth = abs((int)I - (int)prev) < a.tau_i;
    P1v = __float_as_uint(smooth ? a.P1s : a.P1n);
    P2v = __float_as_uint(smooth ? a.P2s : a.P2n);
  };
  // uniform addressing: steps past nmin - 1 re-read step nmin - 1 (in bounds for every row)
  const uint32_t pkstep = (uint32_t)pstep * D * 2u;
  auto load_uniform = [&](int base, CostWords<NW> (&c)[B]) {
    uint32_t pk = (uint32_t)(pbase + min(base, nmin - 1) * pstep) * D * 2u;
#pragma unroll
    for (int i = 0; i < B; ++i) {
      c[i] = load_words<NW>(bcost, voff, pk);
      pk = scalar_add(pk, base + i + 1 < nmin ? pkstep : 0u);
    }
  };
  auto load_own = [&](int base, CostWords<NW> (&c)[B]) {
#pragma unroll
    for (int i = 0; i < B; ++i) c[i] = load_words<NW>(bcost, own(base + i), 0u);
  };
  // one step: the recurrence on the lane's planes and their packed levels
  auto step = [&](const CostWords<NW>& c, uint32_t P1u, uint32_t P2u, uint32_t (&qw)[NW]) {
    const uint64_t inv = row_any_mask(__ballot(CC::any_negative(c)));
    static_assert(kCostInf == 0.0f, "an invalid pixel adds 0 in place of its cost");
    const float w = lane_weight(inv);
    uint32_t lm = __float_as_uint(L[0]);
#pragma unroll
    for (int j = 1; j < NV; ++j) lm = min(lm, __float_as_uint(L[j]));
    const float m = __uint_as_float(row_min16(lm));   // sgm.cu:55-66
    const float P1 = __uint_as_float(P1u), mP2 = m + __uint_as_float(P2u);
    // L(q, d0-1) and L(q, d0+NV) from the neighbouring lanes of the row; the row's edge lanes
    // have no source lane and keep 0xFFFFFFFF, above every non-negative float's bit pattern
    lo = (uint32_t)__builtin_amdgcn_update_dpp((int)lo, (int)__float_as_uint(L[NV - 1]), 0x111,
                                               0xF, 0xF, false);   // row_shr:1
    hi = (uint32_t)__builtin_amdgcn_update_dpp((int)hi, (int)__float_as_uint(L[0]), 0x101, 0xF,
                                               0xF, false);        // row_shl:1
    uint32_t q[NV];
#pragma unroll
    for (int j = 0; j < NV; ++j) {
      const uint32_t left = j > 0 ? __float_as_uint(L[j - 1]) : lo;
      const uint32_t right = j + 1 < NV ? __float_as_uint(L[j + 1]) : hi;
      // min(L(q,d), L(q,d-1)+P1, L(q,d+1)+P1, m+P2) (sgm.cu:85-93)
      const float val = umin(umin(L[j], mP2), __uint_as_float(min(left, right)) + P1);
      q[j] = excess_level_bits(val - m, a.eq_inv);
      Lq[j] = fmaf(CC::plane(c, j), w, val) - m;   // sgm.cu:98
    }
#pragma unroll
    for (int j = 0; j < NV; ++j) L[j] = Lq[j];
    // the low halves of two level words side by side (v_perm_b32: bytes 1:0 of each)
#pragma unroll
    for (int i = 0; i < NW; ++i) qw[i] = __builtin_amdgcn_perm(q[2 * i + 1], q[2 * i], 0x05040100u);
  };
  auto run_uniform = [&](int base, const CostWords<NW> (&c)[B]) {
    penalties(base);
    uint32_t pk = (uint32_t)(pbase + base * pstep) * D * 2u;
#pragma unroll
    for (int i = 0; i < B; ++i) {
      uint32_t qw[NW];
      step(c[i], row_bcast_at<B>(P1v, i), row_bcast_at<B>(P2v, i), qw);
      store_words<NW, 2>(beout, qw, voff, pk);
      pk = scalar_add(pk, pkstep);
    }
  };
  auto run_own = [&](int base, const CostWords<NW> (&c)[B]) {
    penalties(base);
#pragma unroll
    for (int i = 0; i < B; ++i) {
      if (base + i >= nmax) break;
      uint32_t qw[NW];
      step(c[i], row_bcast_at<B>(P1v, i), row_bcast_at<B>(P2v, i), qw);
      if (base + i < nst) store_words<NW, 2>(beout, qw, own(base + i), 0u);
    }
  };

  int base = 1;
  if (base + B <= nmin) {
    // blocks every row runs in full, two register buffers used ping-pong: the loads of the
    // next block are in flight while the current one computes
    load_uniform(base, c0);
    while (base + 2 * B <= nmin) {
      load_uniform(base + B, c1);
      run_uniform(base, c0);
      load_uniform(base + 2 * B, c0);
      run_uniform(base + B, c1);
      base += 2 * B;
    }
    if (base + B <= nmin) {
      run_uniform(base, c0);
      base += B;
    }
  }
  // the rest (rows of unequal length, the last scanlines of a path), lane by lane
  while (base < nmax) {
    load_own(base, c0);
    run_own(base, c0);
    base += B;
  }
}

}  // namespace pm

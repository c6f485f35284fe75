// One-wave-per-scanline SGM body shared by the plain, batched and differentiable kernels.
// Reference: sgm2<idx,start,dx,dy,n>, motion_stereo/sgm.cu:5-113.
#pragma once

#include <type_traits>

#include "common.hpp"

namespace pm {

__device__ __forceinline__ uint32_t ref_bits(uint8_t v) { return v; }
__device__ __forceinline__ uint32_t ref_bits(float v) { return __float_as_uint(v); }

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
// the path's excess E = L - C' to a.eout as 16-bit fixed point (kExcessInvalid at invalid
// pixels); L itself is computed exactly as in the other forms.
template <int NV, typename CT, bool FIRST, bool FULL, int B, typename RT = uint8_t,
          bool AGG = true, bool LOUT = false, bool LMASK = false, bool EQ = false>
__device__ __forceinline__ void sgm_scanline(const SgmArgs& a, int line,
                                             const CT* __restrict__ cost,
                                             float* __restrict__ agg) {
  const RT* __restrict__ ref = static_cast<const RT*>(a.ref);
  float* __restrict__ lout = a.lout;
  uint16_t* __restrict__ eout = a.eout;
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
  const float INF = __uint_as_float(0x7F800000u);

  // two register buffers of B steps each, used ping-pong: while one block computes, the
  // loads of the next are in flight (addresses past the end re-read the last pixel)
  float c0[B][NV], g0[B][NV], c1[B][NV], g1[B][NV];
  uint32_t r0, r1;
  auto load_block = [&](int base, float (&c)[B][NV], float (&g)[B][NV], uint32_t& r) {
    // lane i fetches the reference intensity of step base+i; issued first so that reading it
    // does not wait for the block's cost and aggregate loads
    r = ref_bits(ref[p0 + min(base + lane, n - 1) * pstep]);
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int i = 0; i < B; ++i) {
      const long p = (long)(p0 + min(base + i, n - 1) * pstep) * D;
#pragma unroll
      for (int j = 0; j < NV; ++j) {
        c[i][j] = load_cost<CT>(cost + p + off[j]);
        if (!FIRST && AGG) g[i][j] = agg[p + off[j]];
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
  auto step = [&](int k, int i, const float (&c)[B][NV], const float (&g)[B][NV]) {
    const long p = (long)(p0 + k * pstep) * D;
    bool neg = false;
#pragma unroll
    for (int j = 0; j < NV; ++j) neg = neg || c[i][j] < 0.0f;
    const bool invalid = __ballot(neg) != 0;
    // m = min_d L(q, d) (sgm.cu:55-66)
    float lm = L[0];
#pragma unroll
    for (int j = 1; j < NV; ++j) lm = umin(lm, L[j]);
    const float m = wave_min_nonneg(lm);
    // gradient-adaptive penalties (sgm.cu:78-82), scalar
    const bool sm = (smooth >> i) & 1ull;
    const float P1 = sm ? a.P1s : a.P1n;
    const float P2 = sm ? a.P2s : a.P2n;
    const float lo = wave_shr1(L[NV - 1], INF);   // L(q, d0-1)
    const float hi = wave_shl1(L[0], INF);        // L(q, d0+NV)
    const float mP2 = m + P2;
    float Ln[NV];
#pragma unroll
    for (int j = 0; j < NV; ++j) {
      const float left = j > 0 ? L[j - 1] : lo;
      const float right = j + 1 < NV ? L[j + 1] : hi;
      float val = umin(L[j], mP2);                 // sgm.cu:85
      val = umin(val, umin(left, right) + P1);     // sgm.cu:86-93
      const float cc = invalid ? kCostInf : c[i][j];
      Ln[j] = act[j] ? (cc + val) - m : INF;       // sgm.cu:98
      const float out = invalid ? kCostInf : (FIRST ? Ln[j] : g[i][j] + Ln[j]);
      if (AGG && act[j]) agg[p + off[j]] = out;    // sgm.cu:101-108
      if (LOUT && act[j]) lout[p + off[j]] = (LMASK && invalid) ? kCostInf : Ln[j];
      if (EQ && act[j]) {
        // val >= m exactly (every candidate is >= m), so E >= 0; rounding to nearest
        const uint32_t q = min(kExcessMaxQ, (uint32_t)fmaf(val - m, a.eq_inv, 0.5f));
        eout[p + off[j]] = (uint16_t)(invalid ? kExcessInvalid : q);
      }
    }
#pragma unroll
    for (int j = 0; j < NV; ++j) L[j] = Ln[j];
  };
  // a block whose B steps all exist: no exits, so the loads of the other buffer stay in flight
  auto run_full = [&](int base, const float (&c)[B][NV], const float (&g)[B][NV], uint32_t r) {
    smooth = smooth_mask(r);
#pragma unroll
    for (int i = 0; i < B; ++i) step(base + i, i, c, g);
  };
  auto run_checked = [&](int base, const float (&c)[B][NV], const float (&g)[B][NV],
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
    const long p = (long)p0 * D;
    float c[NV], g[NV];
#pragma unroll
    for (int j = 0; j < NV; ++j) {
      c[j] = load_cost<CT>(cost + p + off[j]);
      g[j] = (FIRST || !AGG) ? 0.0f : agg[p + off[j]];
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
      if (AGG && act[j]) agg[p + off[j]] = out;
      if (LOUT && act[j]) lout[p + off[j]] = (LMASK && invalid) ? kCostInf : L[j];
      if (EQ && act[j]) eout[p + off[j]] = (uint16_t)(invalid ? kExcessInvalid : 0u);
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

}  // namespace pm

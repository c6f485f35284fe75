// Sweep form of the batched SGM pass: the paths that move down the image (T->B and, with 8
// paths, the two downward diagonals) advance together row by row, and so do the upward ones;
// only the two horizontal paths stay scanlines (sgm_excess_batch_kernel, excess slots 0 and 1).
//
// Reference: the recurrence of sgm2<idx,start,dx,dy,n>, motion_stereo/sgm.cu:5-113 (per path,
// per pixel, exactly as sgm_scanline computes it), and filterCostKernel, filter_cost.cu:3-55.
//
// Why: the compact form writes one 16-bit excess volume per path and reads all of them back
// (8 x 2 x 2 bytes per element, plus eight reads of the cost).  Here the downward sweep stores
// one volume (the sum of its paths' excess, quantised once) and the upward sweep rebuilds the
// aggregate in registers from the cost, the two horizontal volumes, that sum and its own
// paths' exact excess, then runs the WTA / subpixel filter on the spot.  Per element that is
// 3 volumes written and read instead of 8, and 4 cost reads instead of 8.
//
// What it costs: the recurrence and its per-step minimum over the planes are per path, so
// nothing of the VALU work is shared, and the halo adds a quarter more of it.  The scanline
// form is VALU-bound as well, so on the 752x480x64 workload this form is slower (about 0.4 ms
// per keyframe against 0.25 ms); its use is the scratch it needs, 3 volumes instead of 8.
//
// Layout of the work: one launch per block of kSweepRows rows; a workgroup of kSweepWaves
// waves owns kSweepBand columns of one keyframe.  A pixel's predecessor on a diagonal path is
// one column over on the previous row, so at a block's first row the workgroup starts
// kSweepRows - 1 columns beyond its band on each side (from the previous block's row-boundary
// states) and drops one column per row on each side: its own columns stay exact, and no
// workgroup ever waits for another.  Each wave holds kSweepCols columns (lanes = planes, as
// in the scanline kernels); the first and last columns of a wave exchange their diagonal
// states with the neighbouring waves through LDS once per row.
#include "common.hpp"

namespace pm {

namespace {

// Path slots of a sweep: 0 = vertical, 1 = from the column on the left (x-1), 2 = from the
// column on the right (x+1); the predecessor row is y-1 (downward) or y+1 (upward).
constexpr int kSwPaths = 3;

template <int NV, typename CT, bool FULL, bool UP>
__global__ __launch_bounds__(kSweepWaves * 64) void sgm_sweep_kernel(
    const KfDesc* __restrict__ kfs, const int* __restrict__ kfl, const int* __restrict__ first,
    int n, int rb) {
  // diagonal states of each wave's last column (from-left path) and first column (from-right
  // path) at the previous row, double-buffered by row parity; element 64 NV holds the minimum
  __shared__ float xl[2][kSweepWaves][64 * NV + 1];
  __shared__ float xr[2][kSweepWaves][64 * NV + 1];
  const int item = blockIdx.x;
  const int jk = owner_of(first, n, item);
  const KfDesc& k = kfs[kfl[jk]];
  const int W = k.c.W, H = k.c.H, D = k.c.D;
  const int nrb = sweep_blocks(H);
  if (rb >= nrb) return;   // the whole workgroup: this keyframe has fewer row blocks
  const int band = item - first[jk];
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  const int y0 = UP ? H - 1 - rb * kSweepRows : rb * kSweepRows;
  const int nrows = min(kSweepRows, UP ? y0 + 1 : H - y0);
  const int x0 = band * kSweepBand;
  const int xw = x0 - (kSweepRows - 1) + wave * kSweepCols;   // this wave's first column
  const bool diag = k.paths == 8;
  const int np = diag ? 3 : 1;
  const float P1n = uniform(k.P1n), P2n = uniform(k.P2n), P1s = uniform(k.P1s);
  const float P2s = uniform(k.P2s);
  const int tau = __builtin_amdgcn_readfirstlane(k.tau_i);
  const float INF = __uint_as_float(0x7F800000u);

  const uint32_t V = (uint32_t)W * (uint32_t)H * (uint32_t)D;
  const uint32_t WD = (uint32_t)W * (uint32_t)D;
  const Buf bcost(k.cost, V * (uint32_t)sizeof(CT));
  const Buf bE(k.ebuf, 3u * V * 2u);
  const Buf bS(k.sweep, 6u * WD * 4u);
  const uint8_t* ref = static_cast<const uint8_t*>(k.c.ref);
  int off[NV];
  bool act[NV];
#pragma unroll
  for (int j = 0; j < NV; ++j) {
    const int d = lane * NV + j;
    act[j] = FULL || d < D;
    off[j] = FULL ? d : min(d, D - 1);
  }
  auto lb = [&](int j, uint32_t elt) { return (uint32_t)off[j] * elt; };
  auto min_of = [&](const float (&v)[NV]) {
    float m = v[0];
#pragma unroll
    for (int j = 1; j < NV; ++j) m = umin(m, v[j]);
    return uniform(wave_min_nonneg(m));
  };
  // row-boundary states: half h, path p, column x
  auto state_off = [&](int h, int p, int x) {
    return ((uint32_t)(h * kSwPaths + p) * (uint32_t)W + (uint32_t)x) * (uint32_t)D * 4u;
  };

  // path states of the wave's columns at the previous row, and their minima over the planes
  float L[kSweepCols][kSwPaths][NV], M[kSweepCols][kSwPaths];
#pragma unroll
  for (int c = 0; c < kSweepCols; ++c)
#pragma unroll
    for (int p = 0; p < kSwPaths; ++p) {
      M[c][p] = 0.0f;
#pragma unroll
      for (int j = 0; j < NV; ++j) L[c][p][j] = INF;
    }
  float inl[NV], inr[NV], inlm = 0.0f, inrm = 0.0f;   // neighbours' diagonal states
#pragma unroll
  for (int j = 0; j < NV; ++j) { inl[j] = INF; inr[j] = INF; }
  const int rh = rb & 1;                               // half holding the previous block's row
  auto load_state = [&](int p, int x, float (&v)[NV], float& m) {
    if (rb > 0 && x >= 0 && x < W) {
#pragma unroll
      for (int j = 0; j < NV; ++j) v[j] = bS.f32(lb(j, 4), state_off(rh, p, x));
      m = min_of(v);
    } else {
#pragma unroll
      for (int j = 0; j < NV; ++j) v[j] = INF;
      m = 0.0f;
    }
  };
#pragma unroll
  for (int c = 0; c < kSweepCols; ++c)
#pragma unroll
    for (int p = 0; p < kSwPaths; ++p)
      if (p < np) load_state(p, xw + c, L[c][p], M[c][p]);
  if (diag) {   // the band edges' outer neighbours (read again below for inner waves)
    load_state(1, xw - 1, inl, inlm);
    load_state(2, xw + kSweepCols, inr, inrm);
  }
  int par = 0;
  // publish this wave's edge states and take the neighbours' (the waves at the workgroup's
  // ends keep what they read from the boundary states)
  auto exchange = [&]() {
    if (!diag) return;
#pragma unroll
    for (int j = 0; j < NV; ++j) {
      xl[par][wave][j * 64 + lane] = L[kSweepCols - 1][1][j];
      xr[par][wave][j * 64 + lane] = L[0][2][j];
    }
    if (lane == 0) {
      xl[par][wave][64 * NV] = M[kSweepCols - 1][1];
      xr[par][wave][64 * NV] = M[0][2];
    }
    __syncthreads();
    if (wave > 0) {
#pragma unroll
      for (int j = 0; j < NV; ++j) inl[j] = xl[par][wave - 1][j * 64 + lane];
      inlm = uniform(xl[par][wave - 1][64 * NV]);
    }
    if (wave + 1 < kSweepWaves) {
#pragma unroll
      for (int j = 0; j < NV; ++j) inr[j] = xr[par][wave + 1][j * 64 + lane];
      inrm = uniform(xr[par][wave + 1][64 * NV]);
    }
    par ^= 1;
  };
  exchange();

  const float step = k.estep;
  const float inv_down = k.einv / (float)np;    // the downward volume holds np paths' sum
  const float P = (float)k.paths;
  // One row's inputs: reference intensities at columns xw-1 .. xw+kSweepCols of the row and
  // of its predecessor row (lane l: column xw-1+l), and every column's cost (upward: also the
  // stored excess levels).  Row i+1's are issued before row i computes.
  // With two planes per lane the prefetched levels would not fit in the registers a
  // 1024-lane workgroup leaves each lane; they are then read where they are used.
  constexpr bool kPrefetchLevels = UP && NV == 1;
  struct Row {
    int Icur, Ipre;
    float cst[kSweepCols][NV];
    uint32_t qh[kSweepCols][NV], qd[kSweepCols][NV];
  };
  auto levels = [&](uint32_t e, int j, uint32_t& qh, uint32_t& qd) {
    qh = bE.u16(lb(j, 2), e * 2u) + bE.u16(lb(j, 2), (V + e) * 2u);
    qd = bE.u16(lb(j, 2), (2u * V + e) * 2u);
  };
  auto fetch = [&](int i, Row& r) {
    const int y = UP ? y0 - i : y0 + i;
    const int yp = UP ? y + 1 : y - 1;
    const bool pred_row = UP ? (y + 1 < H) : (y > 0);
    const int xi = min(max(xw - 1 + lane, 0), W - 1);
    r.Icur = ref[(long)y * W + xi];
    r.Ipre = ref[(long)(pred_row ? yp : y) * W + xi];
#pragma unroll
    for (int c = 0; c < kSweepCols; ++c) {
      const int x = min(max(xw + c, 0), W - 1);
      const uint32_t e = ((uint32_t)y * (uint32_t)W + (uint32_t)x) * (uint32_t)D;
#pragma unroll
      for (int j = 0; j < NV; ++j) {
        r.cst[c][j] = buf_cost<CT>(bcost, lb(j, sizeof(CT)), e * (uint32_t)sizeof(CT));
        if (kPrefetchLevels) levels(e, j, r.qh[c][j], r.qd[c][j]);
      }
    }
  };
  // row i from its inputs r
  auto process = [&](int i, const Row& r) {
    const int y = UP ? y0 - i : y0 + i;
    const bool pred_row = UP ? (y + 1 < H) : (y > 0);
    // columns computed at this row (the band plus what is left of the halo)
    const int lo = max(0, x0 - (kSweepRows - 1 - i));
    const int hi = min(W, x0 + kSweepBand + (kSweepRows - 1 - i));
    float es[kSweepCols][NV];
    bool invalid[kSweepCols];
    float sl[NV], slm = inlm;   // from-left state of the column on the left, before its update
#pragma unroll
    for (int j = 0; j < NV; ++j) sl[j] = inl[j];
#pragma unroll
    for (int c = 0; c < kSweepCols; ++c) {
      const int x = xw + c;
      float pl[NV], plm = slm;
#pragma unroll
      for (int j = 0; j < NV; ++j) { pl[j] = sl[j]; sl[j] = L[c][1][j]; }
      slm = M[c][1];
      es[c][0] = 0.0f;
      invalid[c] = false;
      if (x < lo || x >= hi) continue;   // wave-uniform; stays outside for the rest of the block
      bool neg = false;
#pragma unroll
      for (int j = 0; j < NV; ++j) neg = neg || (act[j] && r.cst[c][j] < 0.0f);
      invalid[c] = __ballot(neg) != 0;   // sgm.cu:33-38, 68-72
      float cc[NV];
#pragma unroll
      for (int j = 0; j < NV; ++j) {
        cc[j] = invalid[c] ? kCostInf : r.cst[c][j];
        es[c][j] = 0.0f;
      }
      const int Ip = __builtin_amdgcn_readlane(r.Icur, c + 1);
      float Ln[kSwPaths][NV], Mn[kSwPaths];
      // one path's step at (x, y) from predecessor state pv (minimum pm) whose reference
      // intensity is Iq; has = false starts the path here (L = C', sgm.cu:21-43)
      auto path = [&](int p, bool has, const float (&pv)[NV], float pm, int Iq) {
        const bool sm = abs(Ip - Iq) < tau;   // sgm.cu:75-82
        const float P1 = sm ? P1s : P1n, P2 = sm ? P2s : P2n;
        const uint32_t lo_n = shr1_max(pv[NV - 1]);
        const uint32_t hi_n = shl1_max(pv[0]);
        const float mP2 = pm + P2;
#pragma unroll
        for (int j = 0; j < NV; ++j) {
          const uint32_t left = j > 0 ? __float_as_uint(pv[j - 1]) : lo_n;
          const uint32_t right = j + 1 < NV ? __float_as_uint(pv[j + 1]) : hi_n;
          // min(L(q,d), L(q,d-1)+P1, L(q,d+1)+P1, m+P2) (sgm.cu:85-93)
          const float val = umin(umin(pv[j], mP2), __uint_as_float(min(left, right)) + P1);
          const float l = has ? (cc[j] + val) - pm : cc[j];   // sgm.cu:98
          Ln[p][j] = act[j] ? l : INF;
          es[c][j] += has ? val - pm : 0.0f;
        }
        Mn[p] = min_of(Ln[p]);
      };
      path(0, pred_row, L[c][0], M[c][0], __builtin_amdgcn_readlane(r.Ipre, c + 1));
      if (diag) {
        path(1, pred_row && x > 0, pl, plm, __builtin_amdgcn_readlane(r.Ipre, c));
        if (c + 1 < kSweepCols)
          path(2, pred_row && x + 1 < W, L[c + 1][2], M[c + 1][2],
               __builtin_amdgcn_readlane(r.Ipre, c + 2));
        else
          path(2, pred_row && x + 1 < W, inr, inrm, __builtin_amdgcn_readlane(r.Ipre, c + 2));
      }
#pragma unroll
      for (int p = 0; p < kSwPaths; ++p) {
        if (p >= np) continue;
        M[c][p] = Mn[p];
#pragma unroll
        for (int j = 0; j < NV; ++j) L[c][p][j] = Ln[p][j];
      }
    }
    // the band's own columns: store (downward) or combine and filter (upward)
#pragma unroll
    for (int c = 0; c < kSweepCols; ++c) {
      const int x = xw + c;
      const bool own = x >= x0 && x < x0 + kSweepBand && x >= lo && x < hi;   // wave-uniform
      if (!own) continue;
      const uint32_t e = ((uint32_t)y * (uint32_t)W + (uint32_t)min(max(x, 0), W - 1)) *
                         (uint32_t)D;
      if (!UP) {
        // the downward paths' excess sum, one quantisation (level step np * estep)
#pragma unroll
        for (int j = 0; j < NV; ++j) {
          const uint32_t q = min(kExcessMaxQ, (uint32_t)fmaf(es[c][j], inv_down, 0.5f));
          if (act[j]) bE.put_u16(q, lb(j, 2), (2u * V + e) * 2u);
        }
      } else {
        // aggregate = P C' + step (q_LR + q_RL) + np step q_down + sum of this sweep's excess
        float v[NV];
#pragma unroll
        for (int j = 0; j < NV; ++j) {
          uint32_t qh = 0, qd = 0;
          if (kPrefetchLevels) {
            qh = r.qh[c][j];
            qd = r.qd[c][j];
          } else {
            levels(e, j, qh, qd);
          }
          const float a = fmaf((float)qh, step, fmaf((float)qd, step * (float)np, P * r.cst[c][j])) +
                          es[c][j];
          v[j] = act[j] ? (invalid[c] ? kCostInf : a) : INF;
        }
        if (k.agg) {
#pragma unroll
          for (int j = 0; j < NV; ++j)
            if (act[j]) k.agg[(size_t)e + (size_t)(lane * NV + j)] = v[j];
        }
        // winner-take-all, lowest plane on ties (filter_cost.cu:22-30)
        float lm = v[0];
#pragma unroll
        for (int j = 1; j < NV; ++j) lm = umin(lm, v[j]);
        const float best = uniform(wave_min_nonneg(lm));
        uint32_t cand = 0xFFFFFFFFu;
#pragma unroll
        for (int j = NV - 1; j >= 0; --j)
          if (v[j] == best) cand = (uint32_t)(lane * NV + j);
        const int mi = (int)__builtin_amdgcn_readfirstlane(
            __float_as_uint(wave_min_nonneg(__uint_as_float(cand))));
        const int dm = mi > 0 ? mi - 1 : 0, dp = mi + 1 < D ? mi + 1 : D - 1;
        float sm_ = v[0], sp_ = v[0];
#pragma unroll
        for (int j = 1; j < NV; ++j) {
          if (dm % NV == j) sm_ = v[j];
          if (dp % NV == j) sp_ = v[j];
        }
        const float cm = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(sm_), dm / NV));
        const float cp = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(sp_), dp / NV));
        // rules 1-3 and the parabola (filter_cost.cu:37-51), without branches
        const float A = cm - 2.0f * best + cp;
        const float Bq = -cm + cp;
        const float s = (float)mi - Bq / (2.0f * A);
        const float parab = 1.0f / (k.c.idep_min + s * k.c.dep_sample);
        const bool r1 = (double)best < 1e-5 && (double)best > -1e-5;
        const bool r23 = mi == 0 || mi == D - 1 || cm + cp < 2.0f * best * k.var_scale;
        const float out = r1 ? kDepInf : (r23 ? 0.0f : parab);
        if (lane == 0) k.depth[(size_t)y * W + x] = out;
      }
    }
    exchange();
  };
  // two row buffers used alternately (named, not indexed, so they stay in registers); the
  // next row's loads are issued before this row's arithmetic
  Row ra, rb2;
  fetch(0, ra);
  for (int i = 0; i < nrows; i += 2) {
    if (i + 1 < nrows) fetch(i + 1, rb2);
    __builtin_amdgcn_sched_barrier(0);
    process(i, ra);
    if (i + 1 < nrows) {
      if (i + 2 < nrows) fetch(i + 2, ra);
      __builtin_amdgcn_sched_barrier(0);
      process(i + 1, rb2);
    }
  }
  // the block's last row, for the next block (own columns only)
  if (rb + 1 < nrb) {
#pragma unroll
    for (int c = 0; c < kSweepCols; ++c) {
      const int x = xw + c;
      if (x < x0 || x >= x0 + kSweepBand || x >= W) continue;
#pragma unroll
      for (int p = 0; p < kSwPaths; ++p) {
        if (p >= np) continue;
#pragma unroll
        for (int j = 0; j < NV; ++j)
          if (act[j]) bS.put_f32(L[c][p][j], lb(j, 4), state_off(rh ^ 1, p, x));
      }
    }
  }
}

template <int NV, typename CT, bool FULL>
void sweep_launch_v(bool up, const KfDesc* kfs, const int* kf, const int* first, int n,
                    int total, int rb, hipStream_t st) {
  const dim3 blk(kSweepWaves * 64);
  if (up)
    sgm_sweep_kernel<NV, CT, FULL, true><<<total, blk, 0, st>>>(kfs, kf, first, n, rb);
  else
    sgm_sweep_kernel<NV, CT, FULL, false><<<total, blk, 0, st>>>(kfs, kf, first, n, rb);
}

template <typename CT>
bool sweep_launch_t(int nv, bool full, bool up, const KfDesc* kfs, const int* kf,
                    const int* first, int n, int total, int rb, hipStream_t st) {
  if (nv == 1 && full) sweep_launch_v<1, CT, true>(up, kfs, kf, first, n, total, rb, st);
  else if (nv == 1) sweep_launch_v<1, CT, false>(up, kfs, kf, first, n, total, rb, st);
  else if (nv == 2 && full) sweep_launch_v<2, CT, true>(up, kfs, kf, first, n, total, rb, st);
  else if (nv == 2) sweep_launch_v<2, CT, false>(up, kfs, kf, first, n, total, rb, st);
  else return false;
  return true;
}

}  // namespace

bool launch_sgm_sweep_batch(pm_cost_type ct, int nv, bool full, bool up, const KfDesc* kfs,
                            const int* kf, const int* first, int n, int total, int rb,
                            hipStream_t st) {
  if (total <= 0) return true;
  switch (ct) {
    case PM_COST_F32: return sweep_launch_t<float>(nv, full, up, kfs, kf, first, n, total, rb, st);
    case PM_COST_F16: return sweep_launch_t<__half>(nv, full, up, kfs, kf, first, n, total, rb, st);
    case PM_COST_BF16: return sweep_launch_t<__hip_bfloat16>(nv, full, up, kfs, kf, first, n, total, rb, st);
    default: return false;
  }
}

}  // namespace pm

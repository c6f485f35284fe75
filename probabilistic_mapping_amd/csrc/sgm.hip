// Semi-global aggregation over 4 or 8 scanline directions.
//
// Reference: sgm2<idx,start,dx,dy,n> and its host driver, motion_stereo/sgm.cu:5-135.  The
// reference runs one block per scanline with one thread per plane, and finds min_d L with a
// shared-memory tree and log2(D) __syncthreads per step (sgm.cu:55-66).  Here a scanline is
// one wave64: lane l holds planes d = NV*l .. NV*l+NV-1, the minimum over planes is a DPP
// row reduction plus four readlanes, and the d-1 / d+1 neighbours come from DPP wave shifts,
// so a step has no LDS traffic and no barrier.  Cost and aggregate rows of the next B steps
// are loaded while the current B steps run.
//
// Recurrence (sgm.cu:75-98):
//   L(p,d) = C'(p,d) + min(L(q,d), L(q,d-1)+P1, L(q,d+1)+P1, m+P2) - m,  m = min_d L(q,d)
//   with P1, P2 scaled by q1, q2 where |I(p)-I(q)| < tau_so.
// A pixel is invalid if any plane's cost is < 0: there C' = 0 and the aggregate is set to 0
// (sgm.cu:33-38, 67-71, 101-104).  Every L is >= 0, so mins run on the unsigned bit patterns.
#include "common.hpp"

namespace pm {

struct SgmArgs {
  int W, H, D, dx, dy;
  long nlines;
  const uint8_t* ref;
  float pi1, pi2, tau_so, q1, q2;
};

__device__ __forceinline__ void scanline_start(const SgmArgs& a, long i, int& x, int& y) {
  if (a.dx != 0) {
    const int xs = a.dx > 0 ? 0 : a.W - 1;
    if (i < a.H) { x = xs; y = (int)i; return; }
    const long j = i - a.H;
    x = a.dx > 0 ? (int)(1 + j) : (int)j;
    y = a.dy > 0 ? 0 : a.H - 1;
    return;
  }
  x = (int)i;
  y = a.dy > 0 ? 0 : a.H - 1;
}

template <int NV, typename CT, bool FIRST, int B>
__global__ __launch_bounds__(256) void sgm_path_kernel(const SgmArgs a,
                                                       const CT* __restrict__ cost,
                                                       float* __restrict__ agg) {
  const int lane = threadIdx.x & 63;
  const long line = (long)blockIdx.x * 4 + (threadIdx.x >> 6);
  if (line >= a.nlines) return;
  int x, y;
  scanline_start(a, line, x, y);
  int n = 1 << 30;
  if (a.dx > 0) n = min(n, a.W - x);
  if (a.dx < 0) n = min(n, x + 1);
  if (a.dy > 0) n = min(n, a.H - y);
  if (a.dy < 0) n = min(n, y + 1);
  const int D = a.D;
  const int d0 = lane * NV;
  bool act[NV];
#pragma unroll
  for (int j = 0; j < NV; ++j) act[j] = d0 + j < D;
  const long pstep = (long)a.dy * a.W + a.dx;
  const long p0 = (long)y * a.W + x;
  const float INF = __uint_as_float(0x7F800000u);

  // register double buffer of B steps of cost and aggregate
  float cbuf[B][NV], abuf[B][NV];
  uint32_t ibuf[B];
  auto load_block = [&](int base, float (&cb)[B][NV], float (&ab)[B][NV], uint32_t (&ib)[B]) {
#pragma unroll
    for (int i = 0; i < B; ++i) {
      const int k = base + i;
      const long p = p0 + (long)k * pstep;
      const bool in = k < n;
      ib[i] = in ? (uint32_t)a.ref[p] : 0u;
#pragma unroll
      for (int j = 0; j < NV; ++j) {
        const bool ld = in && act[j];
        cb[i][j] = ld ? load_cost<CT>(cost + p * D + d0 + j) : 0.0f;
        if (!FIRST) ab[i][j] = ld ? agg[p * D + d0 + j] : 0.0f;
        else ab[i][j] = 0.0f;
      }
    }
  };

  load_block(0, cbuf, abuf, ibuf);
  float L[NV];
  uint32_t iprev = 0;
  // step 0: L = C' (sgm.cu:21-43)
  {
    bool neg = false;
#pragma unroll
    for (int j = 0; j < NV; ++j) neg = neg || (act[j] && cbuf[0][j] < 0.0f);
    const bool invalid = __ballot(neg) != 0;
#pragma unroll
    for (int j = 0; j < NV; ++j) {
      L[j] = act[j] ? (invalid ? kCostInf : cbuf[0][j]) : INF;
      if (act[j]) {
        const float out = invalid ? kCostInf : (FIRST ? L[j] : abuf[0][j] + L[j]);
        agg[p0 * D + d0 + j] = out;
      }
    }
    iprev = ibuf[0];
  }

  for (int base = 0; base < n; base += B) {
    float cn[B][NV], an[B][NV];
    uint32_t in_[B];
    load_block(base + B, cn, an, in_);
#pragma unroll
    for (int i = 0; i < B; ++i) {
      const int k = base + i;
      if (k == 0) continue;
      if (k >= n) break;
      const long p = p0 + (long)k * pstep;
      // m = min_d L(q, d) (sgm.cu:55-66)
      float lm = L[0];
#pragma unroll
      for (int j = 1; j < NV; ++j) lm = umin(lm, L[j]);
      const float m = wave_min_nonneg(lm);
      // gradient-adaptive penalties (sgm.cu:75-82)
      const float D1 = fabsf((float)ibuf[i] - (float)iprev);
      iprev = ibuf[i];
      float P1 = a.pi1, P2 = a.pi2;
      if (D1 < a.tau_so) {
        P1 *= a.q1;
        P2 *= a.q2;
      }
      const float lo = wave_shr1(L[NV - 1], INF);   // L(q, d0-1)
      const float hi = wave_shl1(L[0], INF);        // L(q, d0+NV)
      bool neg = false;
#pragma unroll
      for (int j = 0; j < NV; ++j) neg = neg || (act[j] && cbuf[i][j] < 0.0f);
      const bool invalid = __ballot(neg) != 0;
      const float mP2 = m + P2;
      float Ln[NV];
#pragma unroll
      for (int j = 0; j < NV; ++j) {
        const float left = j > 0 ? L[j - 1] : lo;
        const float right = j + 1 < NV ? L[j + 1] : hi;
        float val = umin(L[j], mP2);                 // sgm.cu:85
        val = umin(val, umin(left, right) + P1);     // sgm.cu:86-93
        const float c = invalid ? kCostInf : cbuf[i][j];
        Ln[j] = act[j] ? (c + val) - m : INF;        // sgm.cu:98
        if (act[j]) {
          const float out = invalid ? kCostInf : (FIRST ? Ln[j] : abuf[i][j] + Ln[j]);
          agg[p * D + d0 + j] = out;                 // sgm.cu:101-108
        }
      }
#pragma unroll
      for (int j = 0; j < NV; ++j) L[j] = Ln[j];
    }
#pragma unroll
    for (int i = 0; i < B; ++i) {
      ibuf[i] = in_[i];
#pragma unroll
      for (int j = 0; j < NV; ++j) {
        cbuf[i][j] = cn[i][j];
        abuf[i][j] = an[i][j];
      }
    }
  }
}

template <int NV, typename CT>
pm_status launch_paths(const SgmArgs& base, int paths, const CT* cost, float* agg,
                       hipStream_t st) {
  constexpr int B = 8;
  for (int path = 0; path < paths; ++path) {
    SgmArgs a = base;
    a.dx = kPathDx[path];
    a.dy = kPathDy[path];
    a.nlines = (a.dx != 0 && a.dy != 0) ? (long)a.W + a.H - 1 : (a.dx != 0 ? a.H : a.W);
    const unsigned blocks = (unsigned)((a.nlines + 3) / 4);
    if (path == 0)
      sgm_path_kernel<NV, CT, true, B><<<blocks, 256, 0, st>>>(a, cost, agg);
    else
      sgm_path_kernel<NV, CT, false, B><<<blocks, 256, 0, st>>>(a, cost, agg);
    pm_status s = check_launch("sgm_path_kernel");
    if (s != PM_OK) return s;
  }
  return PM_OK;
}

template <typename CT>
pm_status launch_sgm(const SgmArgs& a, int paths, const CT* cost, float* agg, hipStream_t st) {
  if (a.D <= 64) return launch_paths<1, CT>(a, paths, cost, agg, st);
  if (a.D <= 128) return launch_paths<2, CT>(a, paths, cost, agg, st);
  return launch_paths<4, CT>(a, paths, cost, agg, st);
}

}  // namespace pm

using namespace pm;

extern "C" pm_status pm_sgm(pm_shape shape, const uint8_t* ref, const void* cost,
                            pm_cost_type cost_type, pm_sgm_params params, float* agg,
                            pm_stream stream) {
  clear_error();
  if (!ref || !cost || !agg) return fail(PM_ERR_INVALID_ARGUMENT, "pm_sgm: null pointer");
  if (shape.width <= 0 || shape.height <= 0 || shape.width > 65535 || shape.height > 65535)
    return fail(PM_ERR_INVALID_ARGUMENT, "pm_sgm: bad image size");
  if (shape.planes <= 0 || shape.planes > PM_MAX_PLANES)
    return fail(PM_ERR_INVALID_ARGUMENT, "pm_sgm: planes must be in 1..256");
  if (params.paths != 4 && params.paths != 8)
    return fail(PM_ERR_INVALID_ARGUMENT, "pm_sgm: paths must be 4 or 8");
  const float kn[5] = {params.pi1, params.pi2, params.tau_so, params.sgm_q1, params.sgm_q2};
  for (float v : kn)
    if (!(std::isfinite(v) && v >= 0.0f))
      return fail(PM_ERR_INVALID_ARGUMENT,
                  "pm_sgm: pi1, pi2, tau_so, sgm_q1, sgm_q2 must be finite and >= 0");
  SgmArgs a{};
  a.W = shape.width; a.H = shape.height; a.D = shape.planes; a.ref = ref;
  a.pi1 = params.pi1; a.pi2 = params.pi2; a.tau_so = params.tau_so;
  a.q1 = params.sgm_q1; a.q2 = params.sgm_q2;
  hipStream_t st = (hipStream_t)stream;
  switch (cost_type) {
    case PM_COST_F32: return launch_sgm<float>(a, params.paths, (const float*)cost, agg, st);
    case PM_COST_F16: return launch_sgm<__half>(a, params.paths, (const __half*)cost, agg, st);
    case PM_COST_BF16:
      return launch_sgm<__hip_bfloat16>(a, params.paths, (const __hip_bfloat16*)cost, agg, st);
    default: return fail(PM_ERR_INVALID_ARGUMENT, "pm_sgm: unknown cost type");
  }
}

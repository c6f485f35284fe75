// Batched form: many keyframes of uneven size, source count and plane count per launch.
//
// A pm_batch owns a device table of keyframe descriptors and the work lists of each stage
// (which keyframe owns which work item).  Running the batch enqueues one cost launch for all
// keyframes, one launch per SGM direction and plane class (instead of 8 per keyframe), and one
// filter launch per lane class; the launches use persistent blocks that stride over the work
// list.  The enqueue sequence can be captured once into a hipGraph and replayed.  An
// autotuner picks the cost tile shape, the SGM prefetch depth and the persistent grid sizes by
// timing the stages on the batch itself.
//
// The stages are the reference's calc_cost / sgm2 / filter_cost (motion_stereo/calc_cost.cu,
// sgm.cu, filter_cost.cu) as restated by the single-keyframe kernels; fusion stays per stream
// (it is sequential in time) and is enqueued by the caller after the batch.
#include <algorithm>
#include <cmath>
#include <string>
#include <vector>

#include "common.hpp"

namespace pm {

namespace {

int round_up(int a, int b) { return (a + b - 1) / b * b; }

struct SgmClass {
  int nv = 1;
  bool full = false;
  std::vector<int> kf;       // keyframe indices of the class
  int kf_off = 0;            // offset of kf[] in the device int pool
  int first_off[8] = {};     // offsets of first[] per path in the pool
  int total[8] = {};
  // compact form: one work list over (keyframe, path) segments (sweep form: the horizontal
  // paths only); rest_*: the same over the keyframes no rows class takes
  int seg_off = 0, seg_first_off = 0, nseg = 0, seg_total = 0;
  int rest_seg_off = 0, rest_seg_first_off = 0, rest_nseg = 0, rest_seg_total = 0;
  // sweep form: workgroups (keyframe, band) and the most row blocks of any keyframe
  int sw_first_off = 0, sw_total = 0, sw_blocks = 0;
};

// Compact form, four scanlines per wave (16-bit costs, D = 16 nv): segments (keyframe, path)
// of 16-scanline items.
struct RowsClass {
  int nv = 4;
  std::vector<int> kf;
  int seg_off = 0, seg_first_off = 0, nseg = 0, seg_total = 0;
};

// The cost work list for one (tile rows, plane split) choice: item offsets and plane groups
// per keyframe, and the 8 XCD ranges [xr[x], xr[8 + x]) of about equal work.
struct CostList {
  int first_off = 0, npg_off = 0, xr_off = 0;
  int total = 0, max_range = 0;
};

struct FilterClass {
  int lp = 16;
  std::vector<int> kf;
  int kf_off = 0, first_off = 0, total = 0;
};

}  // namespace

}  // namespace pm

struct pm_batch {
  int nkf = 0;
  pm_cost_type ct = PM_COST_F32;
  pm_sgm_storage storage = PM_SGM_EXACT;
  uint16_t* d_ebuf = nullptr;   // compact / sweep form: excess volumes of every keyframe
  float* d_sweep = nullptr;     // sweep form: row-boundary path states of every keyframe
  int paths = 4;
  int cus = 256;
  int max_sources = 0;          // over the keyframes (bounds the cost tile rows)
  std::vector<pm::KfDesc> host;
  pm::KfDesc* d_kfs = nullptr;
  int* d_pool = nullptr;
  pm::CostList cost[3][2];              // [tile rows 2, 4, 8][plane split off, on]
  std::vector<pm::SgmClass> sgm;
  std::vector<pm::RowsClass> rows;   // compact form with 16-bit costs only
  std::vector<pm::FilterClass> filt;
  pm_batch_tuning tune{4, 8, 16, 0, 0, 0, 1};
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
};

using namespace pm;

namespace {

int grid_for(int total, int blocks_per_cu, int cus) {
  if (blocks_per_cu <= 0) return total;
  return std::max(8, std::min(round_up(total, 8), round_up(blocks_per_cu * cus, 8)));
}

// (tile rows, LDS staging planes) pairs with a compiled cost kernel; LDS per block is
// 4 waves x rows x planes x 64 lanes x 4 bytes <= 64 KB.
bool cost_shape_built(int rows, int planes) {
  return ((rows == 2 || rows == 4) && (planes == 8 || planes == 16)) ||
         (rows == 8 && (planes == 4 || planes == 8));
}

// A tile of r rows holds the warp records of cost_tile_max_sources(r) sources.
bool cost_rows_fit(int rows, int max_sources) {
  return max_sources <= cost_tile_max_sources(rows);
}

bool valid_tuning(const pm_batch_tuning& t, int max_sources) {
  return cost_shape_built(t.cost_tile_rows, t.cost_stage_planes) &&
         cost_rows_fit(t.cost_tile_rows, max_sources) &&
         (t.sgm_prefetch == 8 || t.sgm_prefetch == 16) && t.cost_blocks_per_cu >= 0 &&
         t.sgm_blocks_per_cu >= 0 && t.cost_blocks_per_cu <= 64 && t.sgm_blocks_per_cu <= 64 &&
         (t.cost_plane_split == 0 || t.cost_plane_split == 1) &&
         (t.sgm_rows == 0 || t.sgm_rows == 1);
}

// Plane groups of a keyframe's cost tiles: groups of 32 planes (16 when D is a multiple of 16
// but not of 32), so that each wave's share stays a multiple of 4 planes; else one group.
int cost_plane_groups(int D, bool split) {
  if (!split) return 1;
  if (D % 32 == 0) return D / 32;
  if (D % 16 == 0) return D / 16;
  return 1;
}

pm_status run_cost(pm_batch* b, hipStream_t st) {
  const int ti = b->tune.cost_tile_rows == 2 ? 0 : (b->tune.cost_tile_rows == 4 ? 1 : 2);
  const CostList& cl = b->cost[ti][b->tune.cost_plane_split ? 1 : 0];
  if (cl.total == 0) return PM_OK;
  // 0 blocks per CU: one block per item of the longest XCD range; else a persistent grid
  const int grid = b->tune.cost_blocks_per_cu <= 0
                       ? 8 * cl.max_range
                       : std::min(8 * cl.max_range,
                                  round_up(std::max(8, b->tune.cost_blocks_per_cu * b->cus), 8));
  if (!launch_cost_batch(b->ct, b->tune.cost_tile_rows, b->tune.cost_stage_planes, b->d_kfs,
                         b->d_pool + cl.first_off, b->d_pool + cl.npg_off, b->nkf,
                         b->d_pool + cl.xr_off, grid, st))
    return fail(PM_ERR_INVALID_ARGUMENT, "pm_batch: unsupported cost tuning");
  return check_launch("cost_batch_kernel");
}

// Sweep form: every row block of the downward (up = false) or upward sweep, in order.
pm_status run_sweeps(pm_batch* b, bool up, hipStream_t st) {
  for (const SgmClass& c : b->sgm)
    for (int rb = 0; rb < c.sw_blocks; ++rb) {
      if (!launch_sgm_sweep_batch(b->ct, c.nv, c.full, up, b->d_kfs, b->d_pool + c.kf_off,
                                  b->d_pool + c.sw_first_off, (int)c.kf.size(), c.sw_total, rb,
                                  st))
        return fail(PM_ERR_INVALID_ARGUMENT, "pm_batch: unsupported sweep class");
      const pm_status s = check_launch("sgm_sweep_kernel");
      if (s != PM_OK) return s;
    }
  return PM_OK;
}

pm_status run_sgm(pm_batch* b, hipStream_t st) {
  if (b->storage == PM_SGM_COMPACT || b->storage == PM_SGM_SWEEP) {
    const bool rows = b->tune.sgm_rows != 0 && !b->rows.empty();
    if (rows)
      for (const RowsClass& c : b->rows) {
        if (!launch_sgm_excess_rows_batch(b->ct, b->tune.sgm_prefetch, c.nv, b->d_kfs,
                                          b->d_pool + c.seg_off, b->d_pool + c.seg_first_off,
                                          c.nseg, c.seg_total,
                                          grid_for(c.seg_total, b->tune.sgm_blocks_per_cu,
                                                   b->cus), st))
          return fail(PM_ERR_INVALID_ARGUMENT, "pm_batch: unsupported sgm tuning");
        const pm_status s = check_launch("sgm_excess_rows_batch_kernel");
        if (s != PM_OK) return s;
      }
    for (const SgmClass& c : b->sgm) {
      const int total = rows ? c.rest_seg_total : c.seg_total;
      if (total == 0) continue;
      const int nseg = rows ? c.rest_nseg : c.nseg;
      const int* seg = b->d_pool + (rows ? c.rest_seg_off : c.seg_off);
      const int* first = b->d_pool + (rows ? c.rest_seg_first_off : c.seg_first_off);
      if (!launch_sgm_excess_batch(b->ct, b->tune.sgm_prefetch, c.nv, c.full, b->d_kfs, seg,
                                   first, nseg, total,
                                   grid_for(total, b->tune.sgm_blocks_per_cu, b->cus), st))
        return fail(PM_ERR_INVALID_ARGUMENT, "pm_batch: unsupported sgm tuning");
      const pm_status s = check_launch("sgm_excess_batch_kernel");
      if (s != PM_OK) return s;
    }
    return b->storage == PM_SGM_SWEEP ? run_sweeps(b, false, st) : PM_OK;
  }
  for (int p = 0; p < b->paths; ++p) {
    for (const SgmClass& c : b->sgm) {
      const int total = c.total[p];
      if (total == 0) continue;
      if (!launch_sgm_batch_path(b->ct, b->tune.sgm_prefetch, c.nv, c.full, b->d_kfs,
                                 b->d_pool + c.kf_off, b->d_pool + c.first_off[p],
                                 (int)c.kf.size(), total,
                                 grid_for(total, b->tune.sgm_blocks_per_cu, b->cus), p == 0,
                                 kPathDx[p], kPathDy[p], st))
        return fail(PM_ERR_INVALID_ARGUMENT, "pm_batch: unsupported sgm tuning");
      const pm_status s = check_launch("sgm_batch_kernel");
      if (s != PM_OK) return s;
    }
  }
  return PM_OK;
}

pm_status run_filter(pm_batch* b, hipStream_t st) {
  if (b->storage == PM_SGM_SWEEP) return run_sweeps(b, true, st);   // combines and filters
  for (const FilterClass& c : b->filt) {
    if (c.total == 0) continue;
    if (b->storage == PM_SGM_COMPACT)
      launch_combine_excess_batch(b->ct, c.lp, b->d_kfs, b->d_pool + c.kf_off,
                                  b->d_pool + c.first_off, (int)c.kf.size(), c.total, st);
    else
      launch_filter_batch(c.lp, b->d_kfs, b->d_pool + c.kf_off, b->d_pool + c.first_off,
                          (int)c.kf.size(), c.total, c.total, st);
    const pm_status s = check_launch("filter_batch_kernel");
    if (s != PM_OK) return s;
  }
  return PM_OK;
}

void drop_graph(pm_batch* b) {
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  b->exec = nullptr;
  b->graph = nullptr;
}

}  // namespace

extern "C" pm_status pm_batch_create(const pm_keyframe* kfs, int32_t count,
                                     pm_cost_type cost_type, pm_sgm_params params,
                                     float var_scale, pm_sgm_storage storage,
                                     pm_batch** out) {
  clear_error();
  if (!out) return fail(PM_ERR_INVALID_ARGUMENT, "pm_batch_create: null out");
  *out = nullptr;
  if (!kfs || count <= 0 || count > PM_MAX_BATCH)
    return fail(PM_ERR_INVALID_ARGUMENT, "pm_batch_create: count must be in 1..64");
  if (cost_type != PM_COST_F32 && cost_type != PM_COST_F16 && cost_type != PM_COST_BF16)
    return fail(PM_ERR_INVALID_ARGUMENT, "pm_batch_create: unknown cost type");
  if (params.paths != 4 && params.paths != 8)
    return fail(PM_ERR_INVALID_ARGUMENT, "pm_batch_create: paths must be 4 or 8");
  if (storage != PM_SGM_EXACT && storage != PM_SGM_COMPACT && storage != PM_SGM_SWEEP)
    return fail(PM_ERR_INVALID_ARGUMENT, "pm_batch_create: unknown SGM storage");
  const float kn[6] = {params.pi1, params.pi2, params.tau_so, params.sgm_q1, params.sgm_q2,
                       var_scale};
  for (float v : kn)
    if (!(std::isfinite(v) && v >= 0.0f))
      return fail(PM_ERR_INVALID_ARGUMENT, "pm_batch_create: knobs must be finite and >= 0");
  auto* b = new pm_batch();
  b->nkf = count;
  b->ct = cost_type;
  b->paths = params.paths;
  b->storage = storage;
  int dev = 0, cus = 0;
  if (hipGetDevice(&dev) == hipSuccess &&
      hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess &&
      cus > 0)
    b->cus = cus;
  // Keyframes are held largest first (cost volume W H D, then source count; equal ones in the
  // caller's order), so the persistent launches start their longest items first and end with
  // the short ones.  Each keyframe carries its own output buffers, so the order is internal.
  std::vector<int> order(count);
  for (int k = 0; k < count; ++k) order[k] = k;
  std::stable_sort(order.begin(), order.end(), [&](int i, int j) {
    const uint64_t vi = (uint64_t)std::max(kfs[i].width, 0) * (uint64_t)std::max(kfs[i].height, 0) *
                        (uint64_t)std::max(kfs[i].planes, 0);
    const uint64_t vj = (uint64_t)std::max(kfs[j].width, 0) * (uint64_t)std::max(kfs[j].height, 0) *
                        (uint64_t)std::max(kfs[j].planes, 0);
    return vi != vj ? vi > vj : kfs[i].num_sources > kfs[j].num_sources;
  });
  b->host.resize(count);
  for (int k = 0; k < count; ++k) {
    const pm_keyframe& in = kfs[order[k]];
    const std::string who = "pm_batch_create: keyframe " + std::to_string(order[k]) + ": ";
    if (in.width <= 0 || in.height <= 0 || in.width > 65535 || in.height > 65535 ||
        (long)in.width * in.height >= (1L << 29)) {
      delete b;
      return fail(PM_ERR_INVALID_ARGUMENT, who + "bad image size");
    }
    if (in.planes < 1 || in.planes > PM_MAX_PLANES || in.num_sources < 0 ||
        in.num_sources > PM_MAX_SOURCES ||
        (uint64_t)in.width * in.height * in.planes * 4 >= (1ull << 32)) {
      delete b;
      return fail(PM_ERR_INVALID_ARGUMENT, who + "planes or num_sources out of range");
    }
    if (storage == PM_SGM_SWEEP &&
        (in.planes > 128 || (uint64_t)in.width * in.height * in.planes * 6 >= (1ull << 32))) {
      delete b;
      return fail(PM_ERR_INVALID_ARGUMENT, who + "the sweep storage takes at most 128 planes "
                                                 "and W H D 6 bytes below 4 GiB");
    }
    if (!in.ref || !in.cost || (!in.agg && storage == PM_SGM_EXACT) || !in.depth ||
        !std::isfinite(in.idep_min) || !std::isfinite(in.dep_sample)) {
      delete b;
      return fail(PM_ERR_INVALID_ARGUMENT, who + "null buffer or non-finite planes");
    }
    KfDesc& d = b->host[k];
    d = KfDesc{};
    d.c.W = in.width; d.c.H = in.height; d.c.D = in.planes; d.c.S = in.num_sources;
    b->max_sources = std::max(b->max_sources, in.num_sources);
    d.c.idep_min = in.idep_min; d.c.dep_sample = in.dep_sample; d.c.ref = in.ref;
    for (int s = 0; s < in.num_sources; ++s) {
      if (!in.src_quads[s]) {
        delete b;
        return fail(PM_ERR_INVALID_ARGUMENT, who + "null source quad");
      }
      d.c.quad[s] = in.src_quads[s];
      for (int j = 0; j < 9; ++j) d.c.hm[s][j] = in.homs[s].h[j];
      for (int j = 0; j < 3; ++j) d.c.hm[s][9 + j] = in.homs[s].t[j];
      for (int j = 0; j < 12; ++j)
        if (!std::isfinite(d.c.hm[s][j])) {
          delete b;
          return fail(PM_ERR_INVALID_ARGUMENT, who + "non-finite homography");
        }
    }
    d.cost = in.cost; d.agg = in.agg; d.depth = in.depth;
    d.P1n = params.pi1; d.P2n = params.pi2;
    d.P1s = params.pi1 * params.sgm_q1;   // the same float products as sgm.cu:80-81
    d.P2s = params.pi2 * params.sgm_q2;
    d.tau_i = (int)std::min(256.0f, std::ceil(params.tau_so));
    d.paths = params.paths;
    d.var_scale = var_scale;
    d.estep = excess_step(d.P2n, d.P2s);
    d.einv = d.estep > 0.0f ? 1.0f / d.estep : 0.0f;
  }
  // work lists, packed in one int pool
  std::vector<int> pool;
  auto put = [&](const std::vector<int>& v) {
    const int off = (int)pool.size();
    pool.insert(pool.end(), v.begin(), v.end());
    return off;
  };
  for (int ti = 0; ti < 3; ++ti)
    for (int sp = 0; sp < 2; ++sp) {
      std::vector<int> first(count), npg(count);
      std::vector<double> kw(count);   // work of one item of keyframe k
      int acc = 0;
      double work = 0.0;
      for (int k = 0; k < count; ++k) {
        const CostArgs& c = b->host[k].c;
        first[k] = acc;
        npg[k] = cost_plane_groups(c.D, sp == 1);
        const int items = cost_items(c.W, c.H, 2 << ti) * npg[k];
        // per (plane, source) passes plus the per-plane record and staging work
        kw[k] = (double)((c.D + npg[k] - 1) / npg[k]) * (c.S + 1);
        acc += items;
        work += kw[k] * items;
      }
      // XCD x takes the items whose work midpoint falls in [x, x + 1) / 8 of the total
      std::vector<int> xr(16, acc);
      int x = 0;
      double cum = 0.0;
      xr[0] = 0;
      for (int k = 0; k < count && x < 7; ++k) {
        const int items = (k + 1 < count ? first[k + 1] : acc) - first[k];
        for (int j = 0; j < items && x < 7; ++j) {
          const double mid = cum + 0.5 * kw[k];
          while (x < 7 && mid >= work * (x + 1) / 8.0) xr[++x] = first[k] + j;
          cum += kw[k];
        }
      }
      while (x < 7) xr[++x] = acc;
      for (int i = 0; i < 8; ++i) xr[8 + i] = i < 7 ? xr[i + 1] : acc;
      CostList& cl = b->cost[ti][sp];
      cl.first_off = put(first);
      cl.npg_off = put(npg);
      cl.xr_off = put(xr);
      cl.total = acc;
      cl.max_range = 0;
      for (int i = 0; i < 8; ++i) cl.max_range = std::max(cl.max_range, xr[8 + i] - xr[i]);
    }
  for (int k = 0; k < count; ++k) {
    const int D = b->host[k].c.D;
    const int nv = D <= 64 ? 1 : (D <= 128 ? 2 : 4);
    const bool full = nv < 4 && D == 64 * nv;
    auto it = std::find_if(b->sgm.begin(), b->sgm.end(),
                           [&](const SgmClass& c) { return c.nv == nv && c.full == full; });
    if (it == b->sgm.end()) {
      b->sgm.push_back(SgmClass{});
      it = b->sgm.end() - 1;
      it->nv = nv;
      it->full = full;
    }
    it->kf.push_back(k);
    const int lp = filter_lanes(D);
    auto ft = std::find_if(b->filt.begin(), b->filt.end(),
                           [&](const FilterClass& c) { return c.lp == lp; });
    if (ft == b->filt.end()) {
      b->filt.push_back(FilterClass{});
      ft = b->filt.end() - 1;
      ft->lp = lp;
    }
    ft->kf.push_back(k);
  }
  for (SgmClass& c : b->sgm) {
    c.kf_off = put(c.kf);
    for (int p = 0; p < b->paths; ++p) {
      std::vector<int> first(c.kf.size());
      int acc = 0;
      for (size_t j = 0; j < c.kf.size(); ++j) {
        first[j] = acc;
        const CostArgs& a = b->host[c.kf[j]].c;
        acc += (sgm_lines(a.W, a.H, kPathDx[p], kPathDy[p]) + 3) / 4;
      }
      c.first_off[p] = put(first);
      c.total[p] = acc;
    }
    // compact form: segments (keyframe, path) in keyframe order, so the waves in flight at
    // any time cover all paths of a few keyframes (whose cost volumes stay cache-resident)
    std::vector<int> seg, sfirst;
    int acc = 0;
    const int seg_paths = storage == PM_SGM_SWEEP ? 2 : b->paths;   // sweep: L->R, R->L
    for (int kf : c.kf)
      for (int p = 0; p < seg_paths; ++p) {
        const CostArgs& a = b->host[kf].c;
        seg.push_back(kf);
        seg.push_back(p);
        sfirst.push_back(acc);
        acc += (sgm_lines(a.W, a.H, kPathDx[p], kPathDy[p]) + 3) / 4;
      }
    c.seg_off = put(seg);
    c.seg_first_off = put(sfirst);
    c.nseg = (int)sfirst.size();
    c.seg_total = acc;
    // the same over the keyframes that the rows classes do not take
    std::vector<int> rseg, rfirst;
    int racc = 0;
    if (storage == PM_SGM_COMPACT)
      for (int kf : c.kf) {
        const CostArgs& a = b->host[kf].c;
        if (rows_supported(cost_type, a.D)) continue;
        for (int p = 0; p < b->paths; ++p) {
          rseg.push_back(kf);
          rseg.push_back(p);
          rfirst.push_back(racc);
          racc += (sgm_lines(a.W, a.H, kPathDx[p], kPathDy[p]) + 3) / 4;
        }
      }
    c.rest_seg_off = put(rseg);
    c.rest_seg_first_off = put(rfirst);
    c.rest_nseg = (int)rfirst.size();
    c.rest_seg_total = racc;
    std::vector<int> wfirst(c.kf.size());
    int wacc = 0;
    for (size_t j = 0; j < c.kf.size(); ++j) {
      const CostArgs& a = b->host[c.kf[j]].c;
      wfirst[j] = wacc;
      wacc += sweep_bands(a.W);
      c.sw_blocks = std::max(c.sw_blocks, sweep_blocks(a.H));
    }
    c.sw_first_off = put(wfirst);
    c.sw_total = wacc;
  }
  // rows classes (compact storage, 16-bit costs, D = 32, 64, 96, 128)
  if (storage == PM_SGM_COMPACT)
    for (int k = 0; k < count; ++k) {
      const CostArgs& a = b->host[k].c;
      if (!rows_supported(cost_type, a.D)) continue;
      const int nv = rows_nv(a.D);
      auto it = std::find_if(b->rows.begin(), b->rows.end(),
                             [&](const RowsClass& c) { return c.nv == nv; });
      if (it == b->rows.end()) {
        b->rows.push_back(RowsClass{});
        it = b->rows.end() - 1;
        it->nv = nv;
      }
      it->kf.push_back(k);
    }
  for (RowsClass& c : b->rows) {
    std::vector<int> seg, sfirst;
    int acc = 0;
    for (int kf : c.kf)
      for (int p = 0; p < b->paths; ++p) {
        const CostArgs& a = b->host[kf].c;
        seg.push_back(kf);
        seg.push_back(p);
        sfirst.push_back(acc);
        acc += (sgm_lines(a.W, a.H, kPathDx[p], kPathDy[p]) + 15) / 16;
      }
    c.seg_off = put(seg);
    c.seg_first_off = put(sfirst);
    c.nseg = (int)sfirst.size();
    c.seg_total = acc;
  }
  for (FilterClass& c : b->filt) {
    c.kf_off = put(c.kf);
    std::vector<int> first(c.kf.size());
    int acc = 0;
    for (size_t j = 0; j < c.kf.size(); ++j) {
      first[j] = acc;
      const CostArgs& a = b->host[c.kf[j]].c;
      acc += filter_items(a.W, a.H, a.D);
    }
    c.first_off = put(first);
    c.total = acc;
  }
  hipError_t e = hipSuccess;
  if (storage == PM_SGM_COMPACT || storage == PM_SGM_SWEEP) {
    // every keyframe's volumes start 16-byte aligned (8 uint16); the sweep form keeps 3
    const size_t slots = storage == PM_SGM_SWEEP ? 3 : (size_t)b->paths;
    auto span = [&](const KfDesc& d) {
      return (slots * d.c.W * d.c.H * d.c.D + 7) / 8 * 8;
    };
    size_t total = 0;
    for (const KfDesc& d : b->host) total += span(d);
    e = hipMalloc(&b->d_ebuf, total * sizeof(uint16_t));
    size_t off = 0;
    for (KfDesc& d : b->host) {
      d.ebuf = e == hipSuccess ? b->d_ebuf + off : nullptr;
      off += span(d);
    }
  }
  if (e == hipSuccess && storage == PM_SGM_SWEEP) {
    // two halves of 3 x [W][D] row-boundary states per keyframe (16-byte aligned)
    auto span = [&](const KfDesc& d) { return ((size_t)6 * d.c.W * d.c.D + 3) / 4 * 4; };
    size_t total = 0;
    for (const KfDesc& d : b->host) total += span(d);
    e = hipMalloc(&b->d_sweep, total * sizeof(float));
    size_t off = 0;
    for (KfDesc& d : b->host) {
      d.sweep = e == hipSuccess ? b->d_sweep + off : nullptr;
      off += span(d);
    }
  }
  if (e == hipSuccess) e = hipMalloc(&b->d_kfs, sizeof(KfDesc) * count);
  if (e == hipSuccess) e = hipMalloc(&b->d_pool, sizeof(int) * std::max<size_t>(pool.size(), 1));
  if (e == hipSuccess)
    e = hipMemcpy(b->d_kfs, b->host.data(), sizeof(KfDesc) * count, hipMemcpyHostToDevice);
  if (e == hipSuccess && !pool.empty())
    e = hipMemcpy(b->d_pool, pool.data(), sizeof(int) * pool.size(), hipMemcpyHostToDevice);
  if (e != hipSuccess) {
    pm_batch_destroy(b);
    return fail(e == hipErrorOutOfMemory ? PM_ERR_OUT_OF_MEMORY : PM_ERR_DEVICE,
                std::string("pm_batch_create: ") + hipGetErrorString(e));
  }
  *out = b;
  return PM_OK;
}

extern "C" void pm_batch_destroy(pm_batch* b) {
  if (!b) return;
  drop_graph(b);
  if (b->d_kfs) (void)hipFree(b->d_kfs);
  if (b->d_pool) (void)hipFree(b->d_pool);
  if (b->d_ebuf) (void)hipFree(b->d_ebuf);
  if (b->d_sweep) (void)hipFree(b->d_sweep);
  delete b;
}

extern "C" pm_status pm_batch_set_tuning(pm_batch* b, const pm_batch_tuning* t) {
  clear_error();
  if (!b || !t) return fail(PM_ERR_INVALID_ARGUMENT, "pm_batch_set_tuning: null argument");
  if (!valid_tuning(*t, b->max_sources))
    return fail(PM_ERR_INVALID_ARGUMENT, "pm_batch_set_tuning: bad values");
  b->tune = *t;
  drop_graph(b);   // a captured graph holds the old launch shapes
  return PM_OK;
}

extern "C" pm_status pm_batch_get_tuning(const pm_batch* b, pm_batch_tuning* t) {
  clear_error();
  if (!b || !t) return fail(PM_ERR_INVALID_ARGUMENT, "pm_batch_get_tuning: null argument");
  *t = b->tune;
  return PM_OK;
}

extern "C" pm_status pm_batch_run(pm_batch* b, pm_stream stream) {
  clear_error();
  if (!b) return fail(PM_ERR_INVALID_ARGUMENT, "pm_batch_run: null batch");
  hipStream_t st = (hipStream_t)stream;
  pm_status s = run_cost(b, st);
  if (s == PM_OK) s = run_sgm(b, st);
  if (s == PM_OK) s = run_filter(b, st);
  return s;
}

extern "C" pm_status pm_batch_run_stage(pm_batch* b, int32_t stage, pm_stream stream) {
  clear_error();
  if (!b) return fail(PM_ERR_INVALID_ARGUMENT, "pm_batch_run_stage: null batch");
  hipStream_t st = (hipStream_t)stream;
  switch (stage) {
    case 0: return run_cost(b, st);
    case 1: return run_sgm(b, st);
    case 2: return run_filter(b, st);
    default: return fail(PM_ERR_INVALID_ARGUMENT, "pm_batch_run_stage: stage must be 0, 1 or 2");
  }
}

extern "C" pm_status pm_batch_capture(pm_batch* b, pm_stream stream) {
  clear_error();
  if (!b) return fail(PM_ERR_INVALID_ARGUMENT, "pm_batch_capture: null batch");
  if (!stream)
    return fail(PM_ERR_INVALID_ARGUMENT, "pm_batch_capture: capture needs a non-default stream");
  drop_graph(b);
  hipStream_t st = (hipStream_t)stream;
  hipError_t e = hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal);
  if (e != hipSuccess)
    return fail(PM_ERR_DEVICE, std::string("pm_batch_capture: ") + hipGetErrorString(e));
  const pm_status s = pm_batch_run(b, stream);
  hipGraph_t g = nullptr;
  e = hipStreamEndCapture(st, &g);
  if (s != PM_OK) {
    if (g) (void)hipGraphDestroy(g);
    return s;
  }
  if (e != hipSuccess)
    return fail(PM_ERR_DEVICE, std::string("pm_batch_capture: ") + hipGetErrorString(e));
  b->graph = g;
  e = hipGraphInstantiate(&b->exec, g, nullptr, nullptr, 0);
  if (e != hipSuccess) {
    drop_graph(b);
    return fail(PM_ERR_DEVICE, std::string("pm_batch_capture: ") + hipGetErrorString(e));
  }
  return PM_OK;
}

extern "C" pm_status pm_batch_launch(pm_batch* b, pm_stream stream) {
  clear_error();
  if (!b || !b->exec)
    return fail(PM_ERR_INVALID_STATE, "pm_batch_launch: no captured graph");
  const hipError_t e = hipGraphLaunch(b->exec, (hipStream_t)stream);
  if (e != hipSuccess)
    return fail(PM_ERR_LAUNCH, std::string("pm_batch_launch: ") + hipGetErrorString(e));
  return PM_OK;
}

// Times each stage under every candidate setting (reps runs after one warm-up, mean per run)
// and keeps the fastest per stage.  The batch's output buffers are overwritten.
extern "C" pm_status pm_batch_autotune(pm_batch* b, pm_stream stream, int32_t reps,
                                       pm_batch_tuning* best) {
  clear_error();
  if (!b) return fail(PM_ERR_INVALID_ARGUMENT, "pm_batch_autotune: null batch");
  if (reps < 1 || reps > 100) return fail(PM_ERR_INVALID_ARGUMENT, "pm_batch_autotune: reps");
  hipStream_t st = (hipStream_t)stream;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess)
    return fail(PM_ERR_DEVICE, "pm_batch_autotune: event creation failed");
  pm_status status = PM_OK;
  auto time_stage = [&](pm_status (*fn)(pm_batch*, hipStream_t), float* ms) {
    pm_status s = fn(b, st);
    if (s != PM_OK) return s;
    (void)hipEventRecord(e0, st);
    for (int r = 0; r < reps && s == PM_OK; ++r) s = fn(b, st);
    (void)hipEventRecord(e1, st);
    if (hipEventSynchronize(e1) != hipSuccess) return fail(PM_ERR_DEVICE, "autotune sync");
    (void)hipEventElapsedTime(ms, e0, e1);
    *ms /= (float)reps;
    return s;
  };
  pm_batch_tuning cur = b->tune, win = b->tune;
  float best_ms = 1e30f;
  const int rows[3] = {2, 4, 8}, planes[3] = {4, 8, 16}, bpc[4] = {0, 4, 8, 16};
  for (int r : rows)
    for (int p : planes)
      for (int g : bpc)
        for (int sp = 0; sp < 2; ++sp) {
          if (status != PM_OK) break;
          if (!cost_shape_built(r, p) || !cost_rows_fit(r, b->max_sources)) continue;
          b->tune = cur;
          b->tune.cost_tile_rows = r;
          b->tune.cost_stage_planes = p;
          b->tune.cost_blocks_per_cu = g;
          b->tune.cost_plane_split = sp;
          float ms = 0;
          status = time_stage(run_cost, &ms);
          if (status == PM_OK && ms < best_ms) {
            best_ms = ms;
            win.cost_tile_rows = r;
            win.cost_stage_planes = p;
            win.cost_blocks_per_cu = g;
            win.cost_plane_split = sp;
          }
        }
  best_ms = 1e30f;
  const int pref[2] = {8, 16};
  const int nrows = b->rows.empty() ? 1 : 2;
  for (int rw = 0; rw < nrows; ++rw)
    for (int f : pref)
      for (int g : bpc) {
        if (status != PM_OK) break;
        b->tune = win;
        b->tune.sgm_prefetch = f;
        b->tune.sgm_blocks_per_cu = g;
        b->tune.sgm_rows = rw;
        float ms = 0;
        status = time_stage(run_sgm, &ms);
        if (status == PM_OK && ms < best_ms) {
          best_ms = ms;
          win.sgm_prefetch = f;
          win.sgm_blocks_per_cu = g;
          win.sgm_rows = rw;
        }
      }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  b->tune = status == PM_OK ? win : cur;
  drop_graph(b);
  if (status == PM_OK && best) *best = win;
  return status;
}

// Plane-sweep matching cost, fused over all measurement frames of a keyframe.
//
// Reference: ADCalcCostKernel + ad_calc_cost, motion_stereo/calc_cost.cu:11-253, called once
// per measurement frame by StereoMapper::update (stereo_mapper.cpp:73-91) on a volume reset
// to -1 (stereo_mapper.cpp:69-70).  The reference keeps a running mean in the cost volume and
// a second float volume of counts, re-reading and re-writing both per source (16 B per element
// per source).  Here one launch visits every source while the per-(pixel, plane) sum and count
// stay in registers, and the volume is written once.
//
// Mapping (cost_tile, cost_tile.hpp): lanes are image columns, each pixel's centre warp is
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
#include "cost_tile.hpp"

namespace pm {

// Tile count below which a single-keyframe launch splits each tile's planes over two blocks.
constexpr long kCostSplitTiles = 4096;

// Single-keyframe launch shape (tile rows, LDS staging planes) per cost storage type, measured
// on 752x480x64 keyframes with 5 sources: FP32 0.234 ms with 4 x 8, 0.223 ms with 8 x 4;
// FP16 0.219 ms with 4 x 8, 0.220 ms with 8 x 4.
template <typename CT> struct SingleShape { static constexpr int TY = 4, DC = 8; };
template <> struct SingleShape<float> { static constexpr int TY = 8, DC = 4; };

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

template <typename CT, int TY, int DC>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4))) void cost_tile_kernel(const CostArgs a, CT* __restrict__ cost,
                                                        int npg, int nblocks) {
  static_assert(cost_tile_max_sources(TY) >= PM_MAX_SOURCES, "tile too short for the sources");
  __shared__ float stage[4][TY * DC * 64];
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  // One item per block (nblocks = the grid).  Written as a loop and held to 4 waves per SIMD,
  // the 8 x 4 FP32 tile body takes 128 VGPRs with 9 spilled values, reloaded only in the
  // per-chunk record setup; straight-line it took 168 (3 waves per SIMD), and held to 4 waves
  // it spilled 32 into the inner loop.
  for (int b = blockIdx.x; b < nblocks; b += gridDim.x)
    cost_tile<CT, TY, DC>(a, xcd_block(b, nblocks), cost, stage[wave], nullptr, nullptr, npg);
}

// Batched form: the work items of every keyframe of a batch form one list (keyframe k owns
// items first[k] .. first[k+1]-1, each a tile or, with npg[k] > 1, one plane group of a tile).
// The list is cut into 8 contiguous ranges of about equal work, [xr[x], xr[8 + x]) for XCD x:
// workgroups are dispatched round-robin over the XCDs (block i runs on XCD i % 8, the grid is
// a multiple of 8), so XCD x's blocks stride through range x, and neighbouring tiles share
// that XCD's L2.
template <typename CT, int TY, int DC>
__global__ __launch_bounds__(256) void cost_batch_kernel(const KfDesc* __restrict__ kfs,
                                                         const int* __restrict__ first,
                                                         const int* __restrict__ npg, int nkf,
                                                         const int* __restrict__ xr) {
  __shared__ float stage[4][TY * DC * 64];
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int x = blockIdx.x & 7;
  const int end = xr[8 + x];
  for (int j = xr[x] + (int)(blockIdx.x >> 3); j < end; j += (int)(gridDim.x >> 3)) {
    const int k = owner_of(first, nkf, j);
    cost_tile<CT, TY, DC>(kfs[k].c, j - first[k], static_cast<CT*>(kfs[k].cost), stage[wave],
                          nullptr, nullptr, npg[k]);
  }
}

int cost_items(int W, int H, int tile_rows) {
  return tile_rows == 2 ? cost_tiles<2>(W, H)
                        : (tile_rows == 4 ? cost_tiles<4>(W, H) : cost_tiles<8>(W, H));
}

template <typename CT>
bool launch_cost_batch_t(int tile_rows, int stage_planes, const KfDesc* kfs, const int* first,
                         const int* npg, int nkf, const int* xr, int grid, hipStream_t st) {
  if (tile_rows == 2 && stage_planes == 8)
    cost_batch_kernel<CT, 2, 8><<<grid, 256, 0, st>>>(kfs, first, npg, nkf, xr);
  else if (tile_rows == 2 && stage_planes == 16)
    cost_batch_kernel<CT, 2, 16><<<grid, 256, 0, st>>>(kfs, first, npg, nkf, xr);
  else if (tile_rows == 4 && stage_planes == 8)
    cost_batch_kernel<CT, 4, 8><<<grid, 256, 0, st>>>(kfs, first, npg, nkf, xr);
  else if (tile_rows == 4 && stage_planes == 16)
    cost_batch_kernel<CT, 4, 16><<<grid, 256, 0, st>>>(kfs, first, npg, nkf, xr);
  else if (tile_rows == 8 && stage_planes == 8)   // 64 KB of LDS per block
    cost_batch_kernel<CT, 8, 8><<<grid, 256, 0, st>>>(kfs, first, npg, nkf, xr);
  else if (tile_rows == 8 && stage_planes == 4)
    cost_batch_kernel<CT, 8, 4><<<grid, 256, 0, st>>>(kfs, first, npg, nkf, xr);
  else
    return false;
  return true;
}

bool launch_cost_batch(pm_cost_type ct, int tile_rows, int stage_planes, const KfDesc* kfs,
                       const int* first, const int* npg, int nkf, const int* xr, int grid,
                       hipStream_t st) {
  if (grid <= 0 || grid % 8 != 0) return false;
  switch (ct) {
    case PM_COST_F32:
      return launch_cost_batch_t<float>(tile_rows, stage_planes, kfs, first, npg, nkf, xr, grid,
                                        st);
    case PM_COST_F16:
      return launch_cost_batch_t<__half>(tile_rows, stage_planes, kfs, first, npg, nkf, xr, grid,
                                         st);
    case PM_COST_BF16:
      return launch_cost_batch_t<__hip_bfloat16>(tile_rows, stage_planes, kfs, first, npg, nkf,
                                                 xr, grid, st);
    default:
      return false;
  }
}

}  // namespace pm

using namespace pm;

namespace pm {
// One keyframe's tiles are about 1.5 rounds of resident blocks at 752x480: splitting each
// tile's planes over several blocks (4 planes per wave at the least, 4-aligned) leaves a short
// tail.  Measured at 752x480x64 with 4-row tiles: 0.32 ms unsplit, 0.289 ms in two, 0.284 ms
// in four.
template <typename CT>
void launch_cost_single(const CostArgs& a, CT* cost, hipStream_t st) {
  constexpr int TY = SingleShape<CT>::TY, DC = SingleShape<CT>::DC;
  const long tiles = cost_tiles<TY>(a.W, a.H);
  const int npg = tiles >= kCostSplitTiles ? 1 : (a.D % 64 == 0 ? 4 : (a.D % 32 == 0 ? 2 : 1));
  const int nblocks = (int)(tiles * npg);
  cost_tile_kernel<CT, TY, DC><<<(unsigned)nblocks, 256, 0, st>>>(a, cost, npg, nblocks);
}
}  // namespace pm

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
      (long)shape.width * shape.height >= (1L << 29))
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
  hipStream_t st = (hipStream_t)stream;
  switch (cost_type) {
    case PM_COST_F32:
      launch_cost_single<float>(a, (float*)cost, st);
      break;
    case PM_COST_F16:
      launch_cost_single<__half>(a, (__half*)cost, st);
      break;
    case PM_COST_BF16:
      launch_cost_single<__hip_bfloat16>(a, (__hip_bfloat16*)cost, st);
      break;
    default:
      return fail(PM_ERR_INVALID_ARGUMENT, "pm_calc_cost: unknown cost type");
  }
  return check_launch("cost_tile_kernel");
}

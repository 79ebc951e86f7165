// Launchers of the kernels templated on the image storage type (reference
// patch, init scoring, plane sweep, prior re-init, candidate and refinement costs,
// fallback).
// Included once per storage type, by pm_cost_f32.hip and pm_cost_f16.hip,
// with PM_STORAGE_T (the texel type) and PM_STORAGE_FN (the launcher name),
// so the two sets of instantiations compile in parallel.
#pragma once
#include "pm_body.h"

namespace pmhip {
namespace {

template <int HP, typename T, int TX>
void launch_tiled(int kernel, const PmArgs* a, const PmBatch* bt, dim3 grid, size_t lds, hipStream_t s) {
  switch (kernel) {
    case PM_K_CANDIDATES:
      if (a) hipLaunchKernelGGL((candidates_kernel<HP, T, TX>), grid, dim3(kWave, 9), 0, s, *a);
      else hipLaunchKernelGGL((candidates_batch<HP, T, TX>), grid, dim3(kWave, 9), 0, s, *bt);
      break;
    case PM_K_REFINE_COST:
      if (a) hipLaunchKernelGGL((refine_cost_kernel<HP, T, TX>), grid, dim3(kWave, 6), 0, s, *a);
      else hipLaunchKernelGGL((refine_cost_batch<HP, T, TX>), grid, dim3(kWave, 6), 0, s, *bt);
      break;
    default:  // PM_K_FALLBACK; lds: dynamic LDS that only caps its occupancy
      if (a) hipLaunchKernelGGL((fallback_kernel<HP, T, TX>), grid, dim3(kWave, kFbWaves), lds, s, *a);
      else hipLaunchKernelGGL((fallback_batch<HP, T, TX>), grid, dim3(kWave, kFbWaves), lds, s, *bt);
      break;
  }
}

template <int HP, typename T>
hipError_t launch_hp(int kernel, const PmArgs* a, const PmBatch* bt, dim3 grid, size_t lds, int tx, hipStream_t s) {
  switch (kernel) {
    case PM_K_REFPATCH:
      if (a) hipLaunchKernelGGL((refpatch_kernel<HP, T>), grid, dim3(kWave), 0, s, *a);
      else hipLaunchKernelGGL((refpatch_batch<HP, T>), grid, dim3(kWave), 0, s, *bt);
      break;
    case PM_K_INIT_VIEWS:
      if (a) hipLaunchKernelGGL((init_views_kernel<HP, T>), grid, dim3(kWave), lds, s, *a);
      else hipLaunchKernelGGL((init_views_batch<HP, T>), grid, dim3(kWave), lds, s, *bt);
      break;
    case PM_K_SWEEP:  // single-reference launches only
      if (!a) return hipErrorInvalidValue;
      hipLaunchKernelGGL((plane_sweep_kernel<HP, T>), grid, dim3(kWave), lds, s, *a);
      break;
    case PM_K_PRIOR:
      if (a) hipLaunchKernelGGL((prior_kernel<HP, T>), grid, dim3(kWave), lds, s, *a);
      else hipLaunchKernelGGL((prior_batch<HP, T>), grid, dim3(kWave), lds, s, *bt);
      break;
    case PM_K_CANDIDATES:
    case PM_K_REFINE_COST:
    case PM_K_FALLBACK:
      switch (tx) {
        case 4: launch_tiled<HP, T, 4>(kernel, a, bt, grid, lds, s); break;
        case 8: launch_tiled<HP, T, 8>(kernel, a, bt, grid, lds, s); break;
        case 16: launch_tiled<HP, T, 16>(kernel, a, bt, grid, lds, s); break;
        default: return hipErrorInvalidValue;
      }
      break;
    default:
      return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

}  // namespace

hipError_t PM_STORAGE_FN(int kernel, const PmArgs* a, const PmBatch* bt, unsigned grid, size_t lds, int half_patch,
                         int tile_x, hipStream_t stream) {
  if ((a == nullptr) == (bt == nullptr)) return hipErrorInvalidValue;
  if (grid == 0) return hipSuccess;
  const dim3 g(grid);
  switch (half_patch) {
    case 1: return launch_hp<1, PM_STORAGE_T>(kernel, a, bt, g, lds, tile_x, stream);
    case 2: return launch_hp<2, PM_STORAGE_T>(kernel, a, bt, g, lds, tile_x, stream);
    case 3: return launch_hp<3, PM_STORAGE_T>(kernel, a, bt, g, lds, tile_x, stream);
    default: return hipErrorInvalidValue;
  }
}

}  // namespace pmhip

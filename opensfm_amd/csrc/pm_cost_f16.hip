// Kernels templated on the image storage type, float16 images.
#define PM_STORAGE_T _Float16
#define PM_STORAGE_FN launch_storage_f16
// On float16 images the cost kernels run fastest at 5 waves per SIMD
// (candidates 20.0 vs 21.1 ms, refine_cost 14.3 vs 14.9 ms at 1600x1200 with
// 8 sources); on float32 images the compiler's own choice is fastest.
#define PM_COST_WAVES_PER_EU 5
// The textureless fallback is fastest at up to 4 waves per SIMD on float16
// images (26.3 vs 28.1 ms per step at 1600x1200 with 8 sources) and at 3 on
// float32 images (32.3 vs 32.9 ms), the default.
#define PM_FALLBACK_WAVES_MAX 4
#include "pm_cost_impl.h"

// Kernels templated on the image storage type, float16 images.
#define PM_STORAGE_T _Float16
#define PM_STORAGE_FN launch_storage_f16
// On float16 images the candidates kernel runs fastest at 5 waves per SIMD
// (20.0 vs 21.1 ms at 1600x1200 with 8 sources); refine_cost runs fastest
// at the compiler's own choice (13.2 vs 13.9 ms there, 1.04 vs 1.15 s per
// step of ten 4000x3000 references with 10 sources), as both do on float32
// images.
#define PM_COST_WAVES_PER_EU 5
// The textureless fallback is fastest at up to 4 waves per SIMD on float16
// images (26.3 vs 28.1 ms per step at 1600x1200 with 8 sources) and at 3 on
// float32 images (32.3 vs 32.9 ms), the default.
#define PM_FALLBACK_WAVES_MAX 4
#include "pm_cost_impl.h"

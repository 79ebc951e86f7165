// Kernels templated on the image storage type, float32 images.
#define PM_STORAGE_T float
#define PM_STORAGE_FN launch_storage_f32
#include "pm_cost_impl.h"

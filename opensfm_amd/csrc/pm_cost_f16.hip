// Kernels templated on the image storage type, float16 images.
#define PM_STORAGE_T _Float16
#define PM_STORAGE_FN launch_storage_f16
#include "pm_cost_impl.h"

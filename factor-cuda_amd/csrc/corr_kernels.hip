// correlation kernels (placeholder translation unit)
#include "corr_internal.h"

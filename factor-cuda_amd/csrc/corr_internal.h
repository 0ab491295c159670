// Internal launch interfaces of the correlation kernels (filled in by corr_kernels.hip).
#pragma once
#include <hip/hip_runtime.h>

#include "fc_hip.h"
#include <hip/hip_runtime.h>
extern "C" int fc_stock_corr(const void*, int, const uint8_t*, int, int, int, double*, fc_stock_corr_stats*) { return (int)hipErrorNotSupported; }
extern "C" int fc_factor_corr(const void*, int, const uint8_t*, int, int, int, int, double*) { return (int)hipErrorNotSupported; }
extern "C" int fc_stock_corr_dev(const float*, const uint8_t*, int, int, int, double*, fc_workspace*, void*) { return (int)hipErrorNotSupported; }
extern "C" int fc_factor_corr_dev(const float*, const uint8_t*, int, int, int, int, double*, fc_workspace*, void*) { return (int)hipErrorNotSupported; }

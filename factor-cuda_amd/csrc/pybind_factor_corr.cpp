#include <pybind11/pybind11.h>
PYBIND11_MODULE(factor_corr_pybind, m) {}

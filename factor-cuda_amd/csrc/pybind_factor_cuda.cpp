#include <pybind11/pybind11.h>
PYBIND11_MODULE(factor_cuda_pybind, m) {}

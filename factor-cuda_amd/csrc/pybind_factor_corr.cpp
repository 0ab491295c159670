// factor_corr_pybind: binding of factor_corr_f64 with the reference's
// signature and error mapping (src/factor_corr_pybind.cpp:38-108): any
// non-zero status is a RuntimeError. Calls only the C ABI of libfc_hip.so.
#include <pybind11/numpy.h>
#include <pybind11/pybind11.h>

#include <cstdint>
#include <stdexcept>
#include <string>

#include "fc_hip.h"

namespace py = pybind11;

namespace {

using ArrF32 = py::array_t<float, py::array::c_style>;
using ArrD = py::array_t<double, py::array::c_style | py::array::forcecast>;
using ArrU8 = py::array_t<uint8_t, py::array::c_style | py::array::forcecast>;

int precision_code(py::object p) {
  if (p.is_none()) return FC_PREC_CONTRACT;
  const std::string s = py::cast<std::string>(p);
  if (s == "contract" || s == "fp64") return FC_PREC_CONTRACT;
  if (s == "fp32") return FC_PREC_FP32;
  if (s == "bf16x3") return FC_PREC_BF16X3;
  throw py::value_error("precision must be 'contract', 'fp32' or 'bf16x3'");
}

py::array_t<double> factor_corr_f64(py::object F3, py::object mask, py::object precision) {
  py::array in;
  int vb = 8;
  if (py::isinstance<py::array>(F3) &&
      py::reinterpret_borrow<py::array>(F3).dtype() == py::dtype::of<float>()) {
    in = py::cast<ArrF32>(F3);
    vb = 4;
  } else {
    in = py::cast<ArrD>(F3);  // float64 (or forcecast as in the reference binding)
  }
  if (in.ndim() != 3) throw std::invalid_argument("F3 shape must be (T,N,F) with T,N,F >= 1");
  const py::ssize_t T = in.shape(0), N = in.shape(1), F = in.shape(2);
  if (T < 1 || N < 1 || F < 1) throw std::invalid_argument("F3 shape must be (T,N,F) with T,N,F >= 1");
  if (static_cast<int64_t>(T) * N > INT32_MAX)
    throw std::invalid_argument("T*N exceeds INT32_MAX (implementation cap)");
  if (F > 128) throw std::invalid_argument("F > 128 (factor pair grid cap)");
  const int prec = precision_code(precision);
  ArrU8 mh;
  const uint8_t* mp = nullptr;
  if (!mask.is_none()) {
    mh = mask.cast<ArrU8>();
    if (mh.ndim() != 2 || mh.shape(0) != T || mh.shape(1) != N)
      throw std::invalid_argument("mask shape must be (T,N) matching F3");
    mp = mh.data();
  }
  py::array_t<double> result({F, F});
  double* out = result.mutable_data();
  const void* fp = in.data();
  int rc = 0;
  {
    py::gil_scoped_release release;
    rc = fc_factor_corr(fp, vb, mp, static_cast<int>(T), static_cast<int>(N), static_cast<int>(F), prec,
                        out);
  }
  if (rc != 0) throw std::runtime_error("factor_corr_gpu failed with error code " + std::to_string(rc));
  return result;
}

}  // namespace

PYBIND11_MODULE(factor_corr_pybind, m) {
  m.doc() =
      "MI355X factor_corr: (F,F) pooled pairwise-complete Pearson over (T,N) rows.\n"
      "  factor_corr_f64(F3, mask=None, precision='contract') -> (F,F) float64";
  m.def("factor_corr_f64", &factor_corr_f64, py::arg("F3"), py::arg("mask") = py::none(),
        py::arg("precision") = py::none(),
        "Compute the (F,F) factor correlation matrix from a (T,N,F) panel and an optional "
        "(T,N) bool/uint8 mask.");
}

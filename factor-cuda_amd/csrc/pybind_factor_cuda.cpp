// factor_cuda_pybind: the low-level binding of cs_rank / rolling_ic /
// stock_corr / parameter_scan, same module, function, keyword and default
// names, dtype gates, error mapping and return schemas as the reference
// binding (src/factor_cuda_pybind.cpp:134-536). It calls only the C ABI of
// libfc_hip.so (include/fc_hip.h); the GIL is released around each call and
// every host pointer handed across is owned by a named local that outlives
// the call.
#include <pybind11/numpy.h>
#include <pybind11/pybind11.h>

#include <chrono>
#include <cstdint>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

#include "fc_hip.h"

namespace py = pybind11;

namespace {

using ArrD = py::array_t<double, py::array::c_style | py::array::forcecast>;
using ArrF32 = py::array_t<float, py::array::c_style>;
using ArrF64 = py::array_t<double, py::array::c_style>;
using ArrU8 = py::array_t<uint8_t, py::array::c_style | py::array::forcecast>;

void require_ndarray(py::handle obj, const char* name) {
  if (!py::isinstance<py::array>(obj)) throw py::type_error(std::string(name) + " must be a numpy array");
}

py::dtype dtype_of(py::handle obj) { return py::reinterpret_borrow<py::array>(obj).dtype(); }

// Strict float32 (cs_rank / parameter_scan): a float64 panel is a TypeError
// (src/factor_cuda_pybind.cpp:66-74).
ArrF32 require_f32(py::handle obj, const char* name) {
  require_ndarray(obj, name);
  if (!(dtype_of(obj) == py::dtype::of<float>()))
    throw py::type_error(std::string(name) +
                         " must be float32 (float64 downcast is adapter responsibility)");
  return py::cast<ArrF32>(obj);
}

// float32 or float64 accepted (src/factor_cuda_pybind.cpp:76-84). The values
// stay in their own precision: float32 panels are ranked/summed from the
// float32 values, which is order- and value-equivalent to the exact upcast.
struct RealArray {
  py::array arr;
  int value_bytes = 0;
};
RealArray require_real(py::handle obj, const char* name) {
  require_ndarray(obj, name);
  py::dtype dt = dtype_of(obj);
  RealArray r;
  if (dt == py::dtype::of<float>()) {
    r.arr = py::cast<ArrF32>(obj);
    r.value_bytes = 4;
  } else if (dt == py::dtype::of<double>()) {
    r.arr = py::cast<ArrF64>(obj);
    r.value_bytes = 8;
  } else {
    throw py::type_error(std::string(name) + " must be float32 or float64");
  }
  return r;
}

// Optional mask (None -> nullptr); bool/uint8 only, anything else a
// ValueError (src/factor_cuda_pybind.cpp:86-104).
const uint8_t* mask_or_null(py::object obj, const char* name, py::ssize_t T, py::ssize_t N,
                            ArrU8& holder) {
  if (obj.is_none()) return nullptr;
  require_ndarray(obj, name);
  py::dtype dt = dtype_of(obj);
  if (!(dt == py::dtype::of<bool>() || dt == py::dtype::of<uint8_t>()))
    throw py::value_error(std::string(name) + " must be bool or uint8");
  holder = py::cast<ArrU8>(obj);
  if (holder.ndim() != 2 || holder.shape(0) != T || holder.shape(1) != N)
    throw std::invalid_argument(std::string(name) + " shape must be (T,N) matching input");
  return holder.data();
}

int require_int(py::handle obj, const char* name) {
  if (!py::isinstance<py::int_>(obj)) throw std::invalid_argument(std::string(name) + " must be an integer");
  return py::cast<int>(obj);
}

// negative status -> ValueError, positive -> RuntimeError
// (src/factor_cuda_pybind.cpp:117-123)
void check_rc(int rc, const char* fn) {
  if (rc == 0) return;
  if (rc < 0) throw std::invalid_argument(std::string(fn) + " contract error code " + std::to_string(rc));
  throw std::runtime_error(std::string(fn) + " failed with error code " + std::to_string(rc));
}

void check_2d(const py::array& a, const char* name, py::ssize_t& T, py::ssize_t& N) {
  if (a.ndim() != 2) throw std::invalid_argument(std::string(name) + " must be (T,N)");
  T = a.shape(0);
  N = a.shape(1);
  if (T < 1 || N < 1) throw std::invalid_argument(std::string(name) + " must be (T,N) with T,N >= 1");
}

// Workspace handle: owns one fc_workspace; clear() releases the device
// buffers (idempotent) and the destructor clears before delete, so a handle
// dropped without clear() frees its buffers.
struct Workspace {
  fc_workspace* ws = nullptr;
  Workspace() {
    if (fc_workspace_create(&ws) != 0) throw std::runtime_error("fc_workspace_create failed");
  }
  ~Workspace() { fc_workspace_destroy(ws); }
  Workspace(const Workspace&) = delete;
  Workspace& operator=(const Workspace&) = delete;
  void clear() { fc_workspace_clear(ws); }
};
struct CsRankWorkspace : Workspace {};
struct RollingIcWorkspace : Workspace {};

fc_workspace* ws_of(py::object w) {
  if (w.is_none()) return nullptr;
  if (py::isinstance<CsRankWorkspace>(w)) return w.cast<CsRankWorkspace&>().ws;
  if (py::isinstance<RollingIcWorkspace>(w)) return w.cast<RollingIcWorkspace&>().ws;
  throw py::type_error("workspace must be a CsRankWorkspace / RollingIcWorkspace or None");
}

py::array_t<float> cs_rank_f32(py::object X, py::object mask, bool descending, py::object workspace) {
  ArrF32 xa = require_f32(X, "X");
  py::ssize_t T, N;
  check_2d(xa, "X", T, N);
  if (N > (1 << 24)) throw std::invalid_argument("N > 2^24 (rank precision cap)");
  if (static_cast<int64_t>(T) * N > INT32_MAX)
    throw std::invalid_argument("T*N exceeds INT32_MAX (implementation cap)");
  ArrU8 m_holder;
  const uint8_t* mp = mask_or_null(mask, "mask", T, N, m_holder);
  fc_workspace* ws = ws_of(workspace);
  py::array_t<float> result({T, N});
  float* out = result.mutable_data();
  const float* xp = xa.data();
  int rc = 0;
  {
    py::gil_scoped_release release;
    rc = fc_cs_rank_f32(xp, mp, static_cast<int>(T), static_cast<int>(N), descending ? 1 : 0, out, ws);
  }
  check_rc(rc, "cs_rank_gpu");
  return result;
}

py::object rolling_ic_f64(py::object F, py::object R, py::object fmask, py::object rmask,
                          py::object min_valid_obj, bool return_ranks, py::object workspace) {
  RealArray fa = require_real(F, "F");
  RealArray ra = require_real(R, "R");
  py::ssize_t T, N, T2, N2;
  check_2d(fa.arr, "F", T, N);
  check_2d(ra.arr, "R", T2, N2);
  if (T2 != T || N2 != N) throw std::invalid_argument("R shape must match F (T,N)");
  if (static_cast<int64_t>(T) * N > INT32_MAX)
    throw std::invalid_argument("T*N exceeds INT32_MAX (implementation cap)");
  const int min_valid = require_int(min_valid_obj, "min_valid");
  if (min_valid < 2) throw std::invalid_argument("min_valid must be >= 2");
  // mixed precisions: both sides go to float64 (exact upcast)
  if (fa.value_bytes != ra.value_bytes) {
    fa.arr = py::cast<ArrF64>(py::module_::import("numpy").attr("asarray")(fa.arr, "float64"));
    ra.arr = py::cast<ArrF64>(py::module_::import("numpy").attr("asarray")(ra.arr, "float64"));
    fa.value_bytes = ra.value_bytes = 8;
  }
  ArrU8 fh, rh;
  const uint8_t* fmp = mask_or_null(fmask, "fmask", T, N, fh);
  const uint8_t* rmp = mask_or_null(rmask, "rmask", T, N, rh);
  fc_workspace* ws = ws_of(workspace);
  py::array_t<double> ic({T});
  py::array_t<double> rf, rr;
  double* rfp = nullptr;
  double* rrp = nullptr;
  if (return_ranks) {
    rf = py::array_t<double>({T, N});
    rr = py::array_t<double>({T, N});
    rfp = rf.mutable_data();
    rrp = rr.mutable_data();
  }
  double* icp = ic.mutable_data();
  const void* fp = fa.arr.data();
  const void* rp = ra.arr.data();
  const int vb = fa.value_bytes;
  int rc = 0;
  {
    py::gil_scoped_release release;
    rc = fc_rolling_ic(fp, rp, vb, fmp, rmp, static_cast<int>(T), static_cast<int>(N), min_valid, icp,
                       rfp, rrp, ws);
  }
  check_rc(rc, "rolling_ic_gpu");
  if (!return_ranks) return ic;
  return py::make_tuple(ic, rf, rr);
}

int precision_code(py::object p) {
  if (p.is_none()) return FC_PREC_CONTRACT;
  const std::string s = py::cast<std::string>(p);
  if (s == "contract" || s == "fp64") return FC_PREC_CONTRACT;
  if (s == "fp32") return FC_PREC_FP32;
  if (s == "bf16x3") return FC_PREC_BF16X3;
  throw py::value_error("precision must be 'contract', 'fp32' or 'bf16x3'");
}

py::object stock_corr_f64(py::object X, py::object mask, bool return_stats, py::object precision) {
  RealArray xa = require_real(X, "X");
  py::ssize_t T, N;
  check_2d(xa.arr, "X", T, N);
  if (static_cast<int64_t>(T) * N > INT32_MAX)
    throw std::invalid_argument("T*N exceeds INT32_MAX (implementation cap)");
  if (static_cast<int64_t>(N) * N > INT32_MAX)
    throw std::invalid_argument("N*N exceeds INT32_MAX (output grid cap, N <= 46340)");
  const int prec = precision_code(precision);
  ArrU8 mh;
  const uint8_t* mp = mask_or_null(mask, "mask", T, N, mh);
  py::array_t<double> result({N, N});
  double* out = result.mutable_data();
  const void* xp = xa.arr.data();
  const int vb = xa.value_bytes;
  fc_stock_corr_stats st{0, 0};
  int rc = 0;
  {
    py::gil_scoped_release release;
    rc = fc_stock_corr(xp, vb, mp, static_cast<int>(T), static_cast<int>(N), prec, out, &st);
  }
  check_rc(rc, "stock_corr_gpu");
  if (!return_stats) return result;
  py::dict sd;
  sd["selected_path"] = st.selected_path;
  sd["fallback_count"] = st.fallback_count;
  return py::make_tuple(result, sd);
}

py::dict parameter_scan_f32(py::object X, py::object mask, bool return_timing, py::object active_groups,
                            py::object workspace) {
  ArrF32 xa = require_f32(X, "X");
  py::ssize_t T, N;
  check_2d(xa, "X", T, N);
  if (N > (1 << 24)) throw std::invalid_argument("N > 2^24 (rank precision cap)");
  const int64_t total64 = static_cast<int64_t>(T) * N;
  if (total64 > INT32_MAX) throw std::invalid_argument("T*N exceeds INT32_MAX (implementation cap)");
  int act[4] = {1, 1, 1, 1};
  const int* actp = nullptr;
  int nact = 4;
  if (!active_groups.is_none()) {
    py::sequence seq = py::cast<py::sequence>(active_groups);
    if (py::len(seq) != 4) throw std::invalid_argument("active_groups must have length 4");
    nact = 0;
    for (int g = 0; g < 4; ++g) {
      const int v = py::cast<int>(seq[g]);
      if (v != 0 && v != 1) throw std::invalid_argument("active_groups values must be 0 or 1");
      act[g] = v;
      nact += v;
    }
    actp = act;
  }
  ArrU8 mh;
  const uint8_t* mp = nullptr;
  if (act[0] || act[2]) {
    if (mask.is_none())
      throw std::invalid_argument("parameter_scan requires a mask when any masked group is active");
    mp = mask_or_null(mask, "mask", T, N, mh);
  }
  const size_t total = static_cast<size_t>(total64);
  if (static_cast<size_t>(nact) * total * sizeof(float) > (4ull << 30))
    throw std::invalid_argument("active*T*N*4 bytes exceeds the host output budget");
  py::array_t<float> outs[4];
  float* hp[4] = {nullptr, nullptr, nullptr, nullptr};
  for (int g = 0; g < 4; ++g) {
    if (act[g]) {
      outs[g] = py::array_t<float>({T, N});
      hp[g] = outs[g].mutable_data();
    }
  }
  int status[4];
  double tms[4] = {0, 0, 0, 0};
  float tgms[4] = {0, 0, 0, 0};
  double diag = 0.0;
  const float* xp = xa.data();
  fc_workspace* ws = ws_of(workspace);
  int rc = 0;
  {
    py::gil_scoped_release release;
    const auto t0 = return_timing ? std::chrono::steady_clock::now() : std::chrono::steady_clock::time_point{};
    rc = fc_parameter_scan_f32(xp, mp, static_cast<int>(T), static_cast<int>(N), hp, status,
                               return_timing ? tms : nullptr, return_timing ? tgms : nullptr, actp, ws);
    if (return_timing)
      diag = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  }
  check_rc(rc, "parameter_scan_gpu");
  py::dict result;
  py::list groups, sl;
  for (int g = 0; g < 4; ++g) {
    sl.append(status[g]);
    if (status[g] == 0) {
      groups.append(outs[g]);
    } else if (status[g] == 9 || status[g] == 701 || status[g] == FC_GROUP_NOT_EXECUTED) {
      groups.append(py::none());
    } else {
      throw std::runtime_error("parameter_scan_gpu group-level error code " + std::to_string(status[g]));
    }
  }
  result["groups"] = groups;
  result["group_status"] = sl;
  if (return_timing) {
    py::list a, b;
    for (int g = 0; g < 4; ++g) {
      a.append(tms[g]);
      b.append(tgms[g]);
    }
    result["time_ms"] = a;
    result["time_gpu_ms"] = b;
    result["_elapsed_ms_diag"] = diag;
  }
  return result;
}

}  // namespace

PYBIND11_MODULE(factor_cuda_pybind, m) {
  m.doc() =
      "MI355X (gfx950) cross-sectional kernels, low-level binding over libfc_hip.so.\n"
      "  cs_rank_f32(X, mask=None, descending=False, workspace=None) -> (T,N) f32\n"
      "  rolling_ic_f64(F, R, fmask=None, rmask=None, min_valid=30,\n"
      "                 return_ranks=False, workspace=None) -> (T,) f64 | (ic,rf,rr)\n"
      "  stock_corr_f64(X, mask=None, return_stats=False, precision='contract') -> (N,N) f64\n"
      "  parameter_scan_f32(X, mask, return_timing=False, active_groups=None,\n"
      "                     workspace=None) -> {groups, group_status}\n"
      "  CsRankWorkspace() / RollingIcWorkspace(): cached device buffers; .clear().";
  py::class_<CsRankWorkspace>(m, "CsRankWorkspace")
      .def(py::init<>())
      .def("clear", &CsRankWorkspace::clear, "Release all cached device buffers (idempotent).");
  py::class_<RollingIcWorkspace>(m, "RollingIcWorkspace")
      .def(py::init<>())
      .def("clear", &RollingIcWorkspace::clear, "Release all cached device buffers (idempotent).");
  m.def("cs_rank_f32", &cs_rank_f32, py::arg("X"), py::arg("mask") = py::none(),
        py::arg("descending") = false, py::arg("workspace") = py::none(),
        "Stable ordinal cross-sectional rank of a (T,N) float32 panel; invalid -> quiet NaN.");
  m.def("rolling_ic_f64", &rolling_ic_f64, py::arg("F"), py::arg("R"), py::arg("fmask") = py::none(),
        py::arg("rmask") = py::none(), py::arg("min_valid") = 30, py::arg("return_ranks") = false,
        py::arg("workspace") = py::none(),
        "Daily cross-sectional Spearman IC (optional rank diagnostics, 0 = not ranked).");
  m.def("stock_corr_f64", &stock_corr_f64, py::arg("X"), py::arg("mask") = py::none(),
        py::arg("return_stats") = false, py::arg("precision") = py::none(),
        "(N,N) stock correlation matrix. precision: 'contract' (default, FP64 with "
        "cancellation fallback), 'fp32', or 'bf16x3'.");
  m.def("parameter_scan_f32", &parameter_scan_f32, py::arg("X"), py::arg("mask"),
        py::arg("return_timing") = false, py::arg("active_groups") = py::none(),
        py::arg("workspace") = py::none(),
        "G=4 rank parameter scan (direction x mask_mode) from one launch.");
  m.def("version", []() { return std::string(fc_version()); });
}

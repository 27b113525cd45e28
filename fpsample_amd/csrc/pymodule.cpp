// pymodule.cpp -- pybind11 module fpsample_amd._fpsample.
//
// Mirrors the reference's private module fpsample._fpsample
// (src/lib.cpp:581-660): same function names, argument meaning, validation
// order, exception types and messages (src/lib.cpp:52-109, 249-270,
// 522-579).  The work itself goes through the C ABI of libfpsx.so
// (include/fpsx.h), with the GIL released while the GPU runs.
//
// Differences from the reference, all deliberate:
//   * points are taken as Fortran-order float32 for the vanilla entry, which
//     is the SoA layout the GPU reads, so the array the Python wrapper builds
//     with np.asfortranarray (src/fpsample/__init__.py:62) is used without the
//     C-order copy the reference makes (src/lib.cpp:250);
//   * an empty start-index array raises ValueError instead of reading past
//     the end of the array (src/lib.cpp:172).
#include <pybind11/numpy.h>
#include <pybind11/pybind11.h>
#include <pybind11/stl.h>

#include <cstring>
#include <string>
#include <vector>

#include "fpsx.h"

namespace py = pybind11;

namespace {

struct StartIndex {
  enum Type { SINGLE, ARRAY } type;
  size_t single_idx = 0;
  py::array_t<size_t> array_idx;
};

StartIndex to_start_index(const py::object &obj) {
  StartIndex s;
  if (py::isinstance<py::int_>(obj)) {
    s.type = StartIndex::SINGLE;
    s.single_idx = obj.cast<size_t>();
  } else if (py::isinstance<py::array_t<size_t>>(obj)) {
    s.type = StartIndex::ARRAY;
    s.array_idx = obj.cast<py::array_t<size_t>>();
  } else {
    throw py::type_error("start_idx must be int or 1D numpy array of size_t");
  }
  return s;
}

// Validation of src/lib.cpp:52-109, same order and messages.
template <typename Arr>
void check_py_input(const Arr &points, size_t n_samples, const StartIndex &start_idx) {
  if (points.ndim() != 2)
    throw py::value_error("points must be a 2D array, but got shape " +
                          std::to_string(points.ndim()));
  const py::ssize_t P = points.shape(0);
  const py::ssize_t C = points.shape(1);
  if (C == 0) throw py::value_error("points must have at least one column");
  if (n_samples > static_cast<size_t>(P))
    throw py::value_error("n_samples must be less than the number of points: n_samples=" +
                          std::to_string(n_samples) + ", P=" + std::to_string(P));
  if (start_idx.type == StartIndex::SINGLE) {
    if (start_idx.single_idx >= static_cast<size_t>(P))
      throw py::value_error("start_idx must be less than the number of points: start_idx=" +
                            std::to_string(start_idx.single_idx) + ", P=" + std::to_string(P));
  } else {
    auto arr = start_idx.array_idx.template unchecked<1>();
    if (static_cast<size_t>(arr.shape(0)) > n_samples)
      throw py::value_error(
          "The number of start indices must be less than or equal to n_samples: " +
          std::to_string(arr.shape(0)) + ", n_samples=" + std::to_string(n_samples));
    for (py::ssize_t i = 0; i < arr.shape(0); ++i)
      if (arr(i) >= static_cast<size_t>(P))
        throw py::value_error(
            "All indices in start_idx must be less than the number of points: " +
            std::to_string(arr(i)) + ", P=" + std::to_string(P));
  }
}

[[noreturn]] void raise_rc(int rc, const char *what) {
  throw std::runtime_error(std::string(what) + " failed with error code " + std::to_string(rc) +
                           ": " + fpsx_last_error());
}

using FArray = py::array_t<float, py::array::f_style | py::array::forcecast>;
using CArray = py::array_t<float, py::array::c_style | py::array::forcecast>;

// src/lib.cpp:249-270 (dispatch), :188-246 and :111-186 (the loops, on the GPU).
py::array_t<size_t> fps_sampling_py(FArray points, size_t n_samples, py::object start_idx_obj) {
  StartIndex start_idx = to_start_index(start_idx_obj);
  check_py_input(points, n_samples, start_idx);
  const py::ssize_t P = points.shape(0);
  const py::ssize_t C = points.shape(1);
  if (start_idx.type == StartIndex::SINGLE) {
    if (P <= 0 || C <= 0)
      throw py::value_error("points must be a 2D array with at least one column");
  } else if (P <= 0 || C <= 0) {
    throw std::runtime_error("points must be a 2D array");
  }
  py::array_t<size_t> out(n_samples);
  if (n_samples == 0) return out;

  std::vector<uint64_t> starts;
  if (start_idx.type == StartIndex::SINGLE) {
    starts.push_back(start_idx.single_idx);
  } else {
    auto arr = start_idx.array_idx.unchecked<1>();
    if (arr.shape(0) == 0) throw py::value_error("start_idx must not be an empty array");
    for (py::ssize_t i = 0; i < arr.shape(0); ++i) starts.push_back(arr(i));
  }
  const float *data = points.data();
  uint64_t *optr = reinterpret_cast<uint64_t *>(out.mutable_data());
  int rc;
  {
    py::gil_scoped_release nogil;
    rc = fpsx_fps_sampling(data, (size_t)P, (size_t)C, FPSX_LAYOUT_COLS, n_samples,
                           starts.data(), starts.size(), optr);
  }
  if (rc != FPSX_OK) raise_rc(rc, "fps_sampling");
  return out;
}

// src/lib.cpp:522-579.
py::array_t<size_t> bucket_fps_kdline_sampling_py(CArray points, size_t n_samples, size_t height,
                                                  py::object start_idx_obj) {
  StartIndex start_idx = to_start_index(start_idx_obj);
  if (start_idx.type == StartIndex::ARRAY) {
    PyErr_SetString(PyExc_NotImplementedError, "Array of start indices not implemented yet");
    throw py::error_already_set();
  }
  if (points.ndim() != 2) throw py::value_error("points must be a 2D float32 array");
  const py::ssize_t P = points.shape(0);
  const py::ssize_t C = points.shape(1);
  if (start_idx.single_idx >= static_cast<size_t>(P))
    throw py::value_error("start_idx out of range");
  if (n_samples == 0 || n_samples > static_cast<size_t>(P))
    throw py::value_error("n_samples must be in [1, num_points]");
  if (height == 0) throw py::value_error("height must be >= 1");

  py::array_t<size_t> out(n_samples);
  const float *data = points.data();
  size_t *optr = out.mutable_data();
  int rc;
  {
    py::gil_scoped_release nogil;
    rc = bucket_fps_kdline(data, (size_t)P, (size_t)C, n_samples, start_idx.single_idx, height,
                           optr);
  }
  if (rc != 0) {
    if (rc == FPSX_EBADDIM || rc == FPSX_EBADSTART)
      throw std::runtime_error("bucket_fps_kdline failed with error code " + std::to_string(rc));
    raise_rc(rc, "bucket_fps_kdline");
  }
  return out;
}

// src/lib.cpp:342-366 (dispatch) and :272-340 (the loop, on the GPU).
py::array_t<size_t> fps_npdu_sampling_py(CArray points, size_t n_samples, size_t k,
                                         py::object start_idx_obj) {
  StartIndex start_idx = to_start_index(start_idx_obj);
  check_py_input(points, n_samples, start_idx);
  if (start_idx.type == StartIndex::ARRAY) {
    PyErr_SetString(PyExc_NotImplementedError, "Array of start indices not implemented yet");
    throw py::error_already_set();
  }
  const py::ssize_t P = points.shape(0);
  const py::ssize_t C = points.shape(1);
  if (P <= 0 || C <= 0)
    throw py::value_error("points must be a 2D array with at least one column");
  if (start_idx.single_idx >= static_cast<size_t>(P))
    throw py::value_error("start_idx out of range");
  py::array_t<size_t> out(n_samples);
  if (n_samples == 0) return out;
  const float *data = points.data();
  uint64_t *optr = reinterpret_cast<uint64_t *>(out.mutable_data());
  int rc;
  {
    py::gil_scoped_release nogil;
    rc = fpsx_fps_npdu_sampling(data, (size_t)P, (size_t)C, n_samples, k, start_idx.single_idx,
                                optr);
  }
  if (rc != FPSX_OK) raise_rc(rc, "fps_npdu_sampling");
  return out;
}

// Batched form (no counterpart in the reference): every cloud sampled as
// fps_sampling would sample it alone, all in one persistent GPU launch.
py::list fps_sampling_batch_py(py::list clouds, std::vector<size_t> n_samples,
                               std::vector<size_t> starts) {
  const size_t nc = clouds.size();
  if (nc == 0) throw py::value_error("clouds must not be empty");
  if (n_samples.size() != nc || starts.size() != nc)
    throw py::value_error("n_samples and start_idx must have one entry per cloud");
  std::vector<CArray> arrs;
  arrs.reserve(nc);
  std::vector<uint64_t> offsets(nc + 1, 0);
  py::ssize_t dim = -1;
  for (size_t c = 0; c < nc; ++c) {
    CArray a = CArray::ensure(clouds[c]);
    if (!a) throw py::value_error("cloud " + std::to_string(c) + " is not convertible to float32");
    if (a.ndim() != 2) throw py::value_error("every cloud must be a 2D array");
    if (dim < 0) dim = a.shape(1);
    if (a.shape(1) != dim || dim == 0)
      throw py::value_error("every cloud must have the same, non-zero number of columns");
    offsets[c + 1] = offsets[c] + (uint64_t)a.shape(0);
    arrs.push_back(a);
  }
  std::vector<float> rows(offsets[nc] * (size_t)dim);
  for (size_t c = 0; c < nc; ++c)
    std::memcpy(rows.data() + offsets[c] * dim, arrs[c].data(),
                (size_t)arrs[c].shape(0) * dim * sizeof(float));
  size_t total_m = 0;
  for (size_t v : n_samples) total_m += v;
  std::vector<uint64_t> ns(n_samples.begin(), n_samples.end());
  std::vector<uint64_t> st(starts.begin(), starts.end());
  std::vector<uint64_t> out(total_m);
  int rc;
  {
    py::gil_scoped_release nogil;
    rc = fpsx_fps_batch(rows.data(), (size_t)dim, offsets.data(), nc, ns.data(), st.data(),
                        out.data());
  }
  if (rc == FPSX_EINVAL || rc == FPSX_EBADSTART || rc == FPSX_EBADDIM)
    throw py::value_error(fpsx_last_error());
  if (rc != FPSX_OK) raise_rc(rc, "fps_sampling_batch");
  py::list res;
  size_t off = 0;
  for (size_t c = 0; c < nc; ++c) {
    py::array_t<size_t> a(n_samples[c]);
    std::memcpy(a.mutable_data(), out.data() + off, n_samples[c] * sizeof(uint64_t));
    off += n_samples[c];
    res.append(a);
  }
  return res;
}

}  // namespace

// Declared as the reference declares its module (src/lib.cpp:581): safe
// without the GIL (the device calls release it anyway and the C ABI is
// thread-safe) and under per-interpreter GILs (no static Python state).
PYBIND11_MODULE(_fpsample, m, py::mod_gil_not_used(),
                py::multiple_interpreters::per_interpreter_gil()) {
  m.doc() = R"pbdoc(
        MI355X farthest point sampling (FPS)
        -----------------------------------

        .. currentmodule:: fpsample_amd

        .. autosummary::
           :toctree: _generate

           _fps_sampling
           _fps_npdu_sampling
           _bucket_fps_kdline_sampling
    )pbdoc";

  m.def("_fps_sampling", &fps_sampling_py, py::arg("points"), py::arg("n_samples"),
        py::arg("start_idx"), R"pbdoc(
            Farthest Point Sampling (FPS) on the GPU.
            Args:
                points (np.ndarray[float32, 2D]): N x C point array.
                n_samples (int): number of samples to pick.
                start_idx (int or np.ndarray[uint64, 1D]): initial index or indices.
            Returns:
                np.ndarray[uint64]: sampled point indices.
    )pbdoc");

  m.def("_fps_npdu_sampling", &fps_npdu_sampling_py, py::arg("points"), py::arg("n_samples"),
        py::arg("k"), py::arg("start_idx"), R"pbdoc(
            FPS with Nearest Point Distance Update on the GPU.
            Args:
                points (np.ndarray[float32, 2D]): N x C point array.
                n_samples (int): number of samples to pick.
                k (int): window of indices updated around each pick.
                start_idx (int): initial index to start FPS.
            Returns:
                np.ndarray[uint64]: sampled point indices.
    )pbdoc");

  m.def("_bucket_fps_kdline_sampling", &bucket_fps_kdline_sampling_py, py::arg("points"),
        py::arg("n_samples"), py::arg("height"), py::arg("start_idx"), R"pbdoc(
            Bucket FPS (KD-line tree order) on the GPU, same picks as QuickFPS.
            Args:
                points (np.ndarray[float32, 2D]): N x C point array.
                n_samples (int): number of samples to pick.
                height (int): height of the KD-line tree.
                start_idx (int): position of the first pick in the tree order.
            Returns:
                np.ndarray[uint64]: sampled point indices.
    )pbdoc");

  m.def("_fps_sampling_batch", &fps_sampling_batch_py, py::arg("clouds"), py::arg("n_samples"),
        py::arg("start_idx"), R"pbdoc(
            Batched FPS: one persistent GPU launch over many clouds.
            Args:
                clouds (list[np.ndarray[float32, 2D]]): clouds with the same number of columns.
                n_samples (list[int]): picks per cloud.
                start_idx (list[int]): first pick per cloud.
            Returns:
                list[np.ndarray[uint64]]: per-cloud picks, each equal to _fps_sampling on that cloud.
    )pbdoc");

  m.attr("__version__") = fpsx_version();
}

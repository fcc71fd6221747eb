// CPython module `estimate_normals` with pptk's signature, validation,
// verbose output and return shapes
// (pptk/processing/estimate_normals/estimate_normals.cpp:135-375):
//   estimate_normals(points, k, r, subsample=None, output_eigenvalues=False,
//                    output_all_eigenvectors=False,
//                    output_neighborhood_sizes=False, verbose=True,
//                    num_procs=-1)
// plus one keyword of this library, precision=None ('mixed' | 'fp64').
// The tree is built on the GPU for each call, as the reference builds one
// per call (estimate_normals.cpp:31-35).
#include <chrono>
#include <cmath>
#include <limits>
#include <string>
#include <vector>

#include "pyutil.hpp"

using pyx::Ref;

namespace {

double now_s() {
  using namespace std::chrono;
  return duration<double>(steady_clock::now().time_since_epoch()).count();
}

PyObject* EstimateNormals(PyObject*, PyObject* args, PyObject* kwargs) {
  static const char* keywords[] = {"points", "k", "r", "subsample", "output_eigenvalues",
                                   "output_all_eigenvectors", "output_neighborhood_sizes",
                                   "verbose", "num_procs", "precision", nullptr};
  PyObject* p = nullptr;
  PyObject* subsample = Py_None;
  int k;
  float r;
  int output_eigenvalues = 0, output_all_eigenvectors = 0, output_neighborhood_sizes = 0;
  int verbose = 1, num_procs = -1;
  PyObject* obj_prec = Py_None;
  if (!PyArg_ParseTupleAndKeywords(args, kwargs, "Oif|OiiiiiO", (char**)keywords, &p, &k, &r,
                                   &subsample, &output_eigenvalues, &output_all_eigenvectors,
                                   &output_neighborhood_sizes, &verbose, &num_procs, &obj_prec)) {
    PyErr_SetString(PyExc_RuntimeError, "Failed to parse inputs");
    return nullptr;
  }
  if (r <= 0.0f) {
    PyErr_SetString(PyExc_ValueError, "r must be positive");
    return nullptr;
  } else if (k == 0) {
    PyErr_SetString(PyExc_ValueError, "k cannot be zero");
    return nullptr;
  } else if (k < 0 && r == std::numeric_limits<float>::infinity()) {
    PyErr_SetString(PyExc_ValueError, "invalid combo: r == inf and k < 0");
    return nullptr;
  }
  pptk_precision prec = PPTK_PREC_AUTO;
  if (obj_prec != Py_None) {
    const char* s = PyUnicode_Check(obj_prec) ? PyUnicode_AsUTF8(obj_prec) : nullptr;
    if (!s) {
      PyErr_SetString(PyExc_TypeError, "precision must be None, 'mixed' or 'fp64'");
      return nullptr;
    }
    const std::string ps(s);
    if (ps == "mixed") prec = PPTK_PREC_MIXED;
    else if (ps == "fp64") prec = PPTK_PREC_FP64;
    else if (ps == "auto") prec = PPTK_PREC_AUTO;
    else {
      PyErr_SetString(PyExc_ValueError, "precision must be None, 'auto', 'mixed' or 'fp64'");
      return nullptr;
    }
  }

  pyx::Array2D arr;
  if (!pyx::extract_array2d(p, arr)) {
    if (!PyErr_Occurred())
      PyErr_SetString(PyExc_TypeError, "points must be interpretable as 0, 1 or 2-d array");
    return nullptr;
  }
  std::vector<int32_t> sub;
  const bool has_sub = subsample != Py_None;
  if (has_sub) {
    if (!pyx::extract_indices(subsample, arr.m, sub)) {
      if (!PyErr_Occurred())
        PyErr_SetString(PyExc_TypeError,
                        "subsample must be interpretable as specifying a subset of points");
      return nullptr;
    }
  }
  int typenum;
  pptk_dtype dt;
  if (arr.type_num == NPY_FLOAT32) {
    typenum = NPY_FLOAT32;
    dt = PPTK_F32;
  } else if (arr.type_num == NPY_FLOAT64) {
    typenum = NPY_FLOAT64;
    dt = PPTK_F64;
  } else {
    PyErr_SetString(PyExc_TypeError, "points must be float32 or float64");
    return nullptr;
  }
  if (arr.n != 3) {
    PyErr_Format(PyExc_ValueError, "points must have 3 columns (got %d)", (int)arr.n);
    return nullptr;
  }

  const int64_t m = has_sub ? (int64_t)sub.size() : arr.m;
  const int num_eigen = output_all_eigenvectors ? 3 : 1;
  npy_intp d1[3] = {(npy_intp)m, 3, 3};
  Ref evecs(PyArray_SimpleNew(num_eigen == 3 ? 3 : 2, d1, typenum));
  if (!evecs.p) return nullptr;
  Ref evals, sizes;
  if (output_eigenvalues) {
    npy_intp d2[2] = {(npy_intp)m, 3};
    evals.reset(PyArray_SimpleNew(num_eigen == 3 ? 2 : 1, d2, typenum));
    if (!evals.p) return nullptr;
  }
  if (output_neighborhood_sizes) {
    npy_intp d3[1] = {(npy_intp)m};
    sizes.reset(PyArray_SimpleNew(1, d3, NPY_INT));
    if (!sizes.p) return nullptr;
  }

  if (verbose) {
    PySys_WriteStdout("(estimate_normals) using %d GPU\n", pptk_device_count() > 0 ? 1 : 0);
  }
  pptk_build_params params;
  params.max_leaf_size = 0;
  params.num_proc = num_procs;
  params.empty_split = 0.2;
  pptk_tree* tree = nullptr;
  pptk_status st;
  const void* data = arr.data();
  const int64_t n = arr.m;
  double t0 = now_s();
  Py_BEGIN_ALLOW_THREADS
  st = pptk_tree_build(data, n, 3, dt, &params, PPTK_MEM_HOST, nullptr, &tree);
  Py_END_ALLOW_THREADS
  if (st != PPTK_OK) return pyx::raise_status(st);
  const double build_time = now_s() - t0;
  if (verbose) {
    PySys_WriteStdout("(estimate_normals) k-d tree build time (s): %g\n", build_time);
    PySys_WriteStdout("(estimate_normals) neighborhood parameters: \n  k = %d\n  r = %g\n", k,
                      (double)r);
  }
  void* pe = PyArray_DATA((PyArrayObject*)evecs.p);
  void* pv = evals.p ? PyArray_DATA((PyArrayObject*)evals.p) : nullptr;
  int32_t* ps = sizes.p ? (int32_t*)PyArray_DATA((PyArrayObject*)sizes.p) : nullptr;
  const int32_t* rows = has_sub ? sub.data() : nullptr;
  const double rr = (double)r;
  t0 = now_s();
  Py_BEGIN_ALLOW_THREADS
  st = pptk_normals(tree, k, rr, rows, m, num_eigen, prec, pe, pv, ps, nullptr, PPTK_MEM_HOST,
                    nullptr);
  pptk_tree_free(tree);
  Py_END_ALLOW_THREADS
  if (st != PPTK_OK) return pyx::raise_status(st);
  if (verbose)
    PySys_WriteStdout("(estimate_normals) PCA time (s): %g\n", now_s() - t0);

  if (!evals.p && !sizes.p) return evecs.release();
  PyObject* out = PyTuple_New(3);
  if (!out) return nullptr;
  PyTuple_SET_ITEM(out, 0, evecs.release());
  if (evals.p) {
    PyTuple_SET_ITEM(out, 1, evals.release());
  } else {
    Py_INCREF(Py_None);
    PyTuple_SET_ITEM(out, 1, Py_None);
  }
  if (sizes.p) {
    PyTuple_SET_ITEM(out, 2, sizes.release());
  } else {
    Py_INCREF(Py_None);
    PyTuple_SET_ITEM(out, 2, Py_None);
  }
  return out;
}

const char kDoc[] =
    "Estimates normals at all points using principal component analysis\n"
    "(PCA). Specifically, computes eigenvectors of the covariance matrix C\n"
    "based on the local neighborhoods of each point.\n\n"
    "    C = 1/N sum_i (p_i - mu)(p_i - mu)^T\n\n"
    "Parameters\n----------\n"
    "points : 3-column numpy array of type float32 or float64\n"
    "k : int\n    Number of neighbors to use. For pure r-near neighborhoods, set this to -1.\n"
    "r : float\n    Use neighbors within r of query point. For pure k-nearest neighborhoods,\n"
    "    set this to np.inf.\n"
    "subsample : 1-d array of bool or int, optional\n"
    "output_eigenvalues : bool, optional\n"
    "output_all_eigenvectors : bool, optional\n"
    "output_neighborhood_sizes : bool, optional\n"
    "verbose : bool, optional\n"
    "num_procs : int, optional (accepted for compatibility)\n"
    "precision : None, 'mixed' or 'fp64', optional\n"
    "    None: FP64 for float64 points, 'mixed' for float32 points.\n\n"
    "Returns\n-------\n"
    "results : ndarray (m x 3, or m x 3 x 3 with all eigenvectors) or the tuple\n"
    "    (eigenvectors, eigenvalues or None, neighborhood sizes or None) when\n"
    "    eigenvalues or sizes are requested. The j-th eigenvector of the i-th\n"
    "    row is [i, j, :]; eigenvectors are sorted by increasing eigenvalue.\n";

PyMethodDef kMethods[] = {{"estimate_normals", (PyCFunction)(void (*)(void))EstimateNormals,
                           METH_VARARGS | METH_KEYWORDS, kDoc},
                          {nullptr, nullptr, 0, nullptr}};

PyModuleDef kModule = {PyModuleDef_HEAD_INIT, "estimate_normals", nullptr, -1, kMethods,
                       nullptr, nullptr, nullptr, nullptr};

}  // namespace

PyMODINIT_FUNC PyInit_estimate_normals(void) {
  import_array();
  return PyModule_Create(&kModule);
}

// CPython module `kdtree` with the same methods, argument parsing, errors and
// return types as pptk's (pptk/kdtree/kdtree_wrapper.cpp:281-419):
//   _build(data, numprocs=-1, maxleafsize=10, emptysplit=0.2) -> capsule
//   _query(kdtree, queries, k=1, dmax=inf, numprocs=-1) -> list of int32 arrays
// The body of each call is argument parsing, one C-ABI call into
// libpptk_amd.so with the GIL released, and result marshalling.
#include <cmath>
#include <limits>
#include <vector>

#include "pyutil.hpp"

using pyx::Ref;

namespace {

void delete_tree(PyObject* cap) {
  pptk_tree* t = (pptk_tree*)PyCapsule_GetPointer(cap, nullptr);
  if (t) pptk_tree_free(t);
}

PyObject* Build(PyObject*, PyObject* args, PyObject* kwargs) {
  static const char* keywords[] = {"data", "numprocs", "maxleafsize", "emptysplit", nullptr};
  PyObject* data = nullptr;
  int numprocs = -1;
  PyObject* obj_leaf = nullptr;
  double emptysplit = 0.2;
  if (!PyArg_ParseTupleAndKeywords(args, kwargs, "O|iOd", (char**)keywords, &data, &numprocs,
                                   &obj_leaf, &emptysplit)) {
    PyErr_SetString(PyExc_TypeError, "Build(): failed parsing arguments");
    return nullptr;
  }
  // maxleafsize must be an int like the reference's "i" format; when it is
  // omitted the library picks its tuned bucket size.
  int leaf = 0;
  if (obj_leaf) {
    Ref idx(PyNumber_Index(obj_leaf));
    if (!idx.p) {
      PyErr_Clear();
      PyErr_SetString(PyExc_TypeError, "Build(): failed parsing arguments");
      return nullptr;
    }
    const long v = PyLong_AsLong(idx.p);
    if (v == -1 && PyErr_Occurred()) return nullptr;
    leaf = v < 1 ? 1 : (v > 256 ? 256 : (int)v);
  }
  pyx::Array2D x;
  if (!pyx::extract_array2d(data, x)) {
    if (!PyErr_Occurred())
      PyErr_Format(PyExc_TypeError, "Build(): points array type %s unsupported",
                   Py_TYPE(data)->tp_name);
    return nullptr;
  }
  if (x.n < 2 || x.n > 4) {
    PyErr_Format(PyExc_RuntimeError, "PerformAction(): Invalid k-d tree dim = %d.", (int)x.n);
    return nullptr;
  }
  pptk_dtype dt;
  if (!pyx::map_dtype(x.type_num, &dt)) {
    PyErr_Format(PyExc_RuntimeError, "PerformAction_(): Invalid k-d tree type_num = %d.",
                 x.type_num);
    return nullptr;
  }
  pptk_build_params params;
  params.max_leaf_size = leaf;
  params.num_proc = numprocs;
  params.empty_split = emptysplit;
  pptk_tree* tree = nullptr;
  pptk_status st;
  const void* ptr = x.data();
  const int64_t m = x.m;
  const int dim = (int)x.n;
  Py_BEGIN_ALLOW_THREADS
  st = pptk_tree_build(ptr, m, dim, dt, &params, PPTK_MEM_HOST, nullptr, &tree);
  Py_END_ALLOW_THREADS
  if (st != PPTK_OK) return pyx::raise_status(st);
  PyObject* cap = PyCapsule_New((void*)tree, nullptr, delete_tree);
  if (!cap) pptk_tree_free(tree);
  return cap;
}

// kNN: list of m arrays of up to k neighbours, ascending (d, index).
PyObject* run_knn(pptk_tree* tree, const pptk_queries& q, int64_t m, long k, double dmax,
                  int64_t nvalid) {
  // No row can hold more than num_valid neighbours.
  long keff = k;
  if (keff > nvalid) keff = nvalid > 0 ? (long)nvalid : 1;
  npy_intp dims[2] = {(npy_intp)m, (npy_intp)keff};
  Ref idx(PyArray_SimpleNew(2, dims, NPY_INT32));
  if (!idx.p) return nullptr;
  std::vector<int32_t> counts((size_t)m);
  int32_t* pidx = (int32_t*)PyArray_DATA((PyArrayObject*)idx.p);
  pptk_status st;
  Py_BEGIN_ALLOW_THREADS
  st = pptk_knn(tree, &q, (int)keff, dmax, pidx, counts.data(), nullptr, PPTK_MEM_HOST, nullptr);
  Py_END_ALLOW_THREADS
  if (st != PPTK_OK) return pyx::raise_status(st);
  return pyx::list_from_rows(idx.p, pidx, m, keff, counts.data());
}

// r-near: list of m arrays (tree order within a list).
PyObject* run_rnear(pptk_tree* tree, const pptk_queries& q, int64_t m, double dmax) {
  npy_intp odims[1] = {(npy_intp)(m + 1)};
  Ref off(PyArray_SimpleNew(1, odims, NPY_INT64));
  if (!off.p) return nullptr;
  int64_t* poff = (int64_t*)PyArray_DATA((PyArrayObject*)off.p);
  int64_t total = 0;
  pptk_status st;
  Py_BEGIN_ALLOW_THREADS
  st = pptk_rnear_count(tree, &q, dmax, poff, &total, PPTK_MEM_HOST, nullptr);
  Py_END_ALLOW_THREADS
  if (st != PPTK_OK) return pyx::raise_status(st);
  npy_intp idims[1] = {(npy_intp)total};
  Ref idx(PyArray_SimpleNew(1, idims, NPY_INT32));
  if (!idx.p) return nullptr;
  int32_t* pidx = (int32_t*)PyArray_DATA((PyArrayObject*)idx.p);
  Py_BEGIN_ALLOW_THREADS
  st = pptk_rnear_fill(tree, &q, dmax, poff, pidx, PPTK_MEM_HOST, nullptr);
  Py_END_ALLOW_THREADS
  if (st != PPTK_OK) return pyx::raise_status(st);
  return pyx::list_from_csr(idx.p, pidx, m, poff);
}

PyObject* Query(PyObject*, PyObject* args, PyObject* kwargs) {
  static const char* keywords[] = {"kdtree", "queries", "k", "dmax", "numprocs", nullptr};
  PyObject* obj_tree = nullptr;
  PyObject* obj_q = nullptr;
  PyObject* obj_k = nullptr;
  PyObject* obj_dmax = nullptr;
  int numprocs = -1;
  if (!PyArg_ParseTupleAndKeywords(args, kwargs, "OO|OOi", (char**)keywords, &obj_tree, &obj_q,
                                   &obj_k, &obj_dmax, &numprocs)) {
    PyErr_SetString(PyExc_TypeError, "Query(): failed parsing arguments");
    return nullptr;
  }
  // k: omitted -> 1, None -> r-near (kdtree_wrapper.cpp:357-372)
  long k;
  if (!obj_k) {
    k = 1;
  } else if (obj_k == Py_None) {
    k = -1;
  } else {
    Ref num(PyNumber_Long(obj_k));
    if (!num.p) {
      if (!PyErr_Occurred())
        PyErr_Format(PyExc_RuntimeError, "Query(): could not interpret k of type %s as C long.",
                     Py_TYPE(obj_k)->tp_name);
      return nullptr;
    }
    k = PyLong_AsLong(num.p);
    if (k == -1 && PyErr_Occurred()) return nullptr;
    if (k <= 0) {
      PyErr_Format(PyExc_ValueError, "Query(): encountered negative k=%ld", k);
      return nullptr;
    }
  }
  // dmax: rounded through float32 like CastAsDouble (python_util.h:393-400)
  double dmax = std::numeric_limits<double>::infinity();
  if (obj_dmax && obj_dmax != Py_None) {
    Ref num(PyNumber_Float(obj_dmax));
    if (!num.p) {
      if (!PyErr_Occurred())
        PyErr_Format(PyExc_RuntimeError,
                     "Query(): could not interpret dmax of type %s as C double.",
                     Py_TYPE(obj_dmax)->tp_name);
      return nullptr;
    }
    dmax = (double)(float)PyFloat_AS_DOUBLE(num.p);
    if (dmax < 0.0) {
      PyErr_Format(PyExc_ValueError, "Query(): encountered negative dmax %lf", dmax);
      return nullptr;
    }
  }
  if (!PyCapsule_CheckExact(obj_tree)) {
    PyErr_SetString(PyExc_TypeError,
                    "Query(): first arg must be capsule of a KdTreeStruct pointer");
    return nullptr;
  }
  pptk_tree* tree = (pptk_tree*)PyCapsule_GetPointer(obj_tree, nullptr);
  if (!tree || !pptk_tree_check(tree)) {
    PyErr_Clear();
    PyErr_SetString(PyExc_TypeError,
                    "Query(): first arg must be capsule of a KdTreeStruct pointer");
    return nullptr;
  }
  pptk_tree_info info;
  pptk_status st = pptk_tree_get_info(tree, &info);
  if (st != PPTK_OK) return pyx::raise_status(st);

  pptk_queries q;
  q.kind = PPTK_Q_ALL;
  q.m = 0;
  q.indices = nullptr;
  q.points = nullptr;
  q.dtype = PPTK_F32;
  q.mem = PPTK_MEM_HOST;
  int64_t m = 0;
  pyx::Array2D qa;
  std::vector<int32_t> qidx;
  const char* who = "QueryWithIndices()";
  if (obj_q == Py_None) {
    q.kind = PPTK_Q_ALL;
    m = info.num_valid;
  } else if (PySlice_Check(obj_q)) {
    PyErr_SetString(PyExc_NotImplementedError,
                    "QueryTreeAction::Perform(): slice-based query not yet implemented");
    return nullptr;
  } else if (PyArray_Check(obj_q) && PyArray_NDIM((PyArrayObject*)obj_q) == 2) {
    if (!pyx::extract_array2d(obj_q, qa)) return nullptr;
    if (qa.n != info.dim) {
      PyErr_Format(PyExc_ValueError,
                   "QueryTreeAction::Perform(): query point dim = %d (expecting dim = %d).",
                   (int)qa.n, info.dim);
      return nullptr;
    }
    pptk_dtype qd;
    if (!pyx::map_dtype(qa.type_num, &qd)) {
      PyErr_Format(PyExc_TypeError,
                   "QueryTreeAction::Perform(): query points of type_num = %d are not supported.",
                   qa.type_num);
      return nullptr;
    }
    q.kind = PPTK_Q_POINTS;
    q.m = m = qa.m;
    q.points = qa.data();
    q.dtype = qd;
    who = "QueryWithPoints()";
  } else if (pyx::extract_indices(obj_q, info.num_valid, qidx)) {
    q.kind = PPTK_Q_INDICES;
    q.m = m = (int64_t)qidx.size();
    q.indices = qidx.data();
  } else {
    if (!PyErr_Occurred())
      PyErr_Format(PyExc_TypeError,
                   "QueryTreeAction::Perform(): could not use object of type %s as query input.",
                   Py_TYPE(obj_q)->tp_name);
    return nullptr;
  }
  if (k > 0) return run_knn(tree, q, m, k, dmax, info.num_valid);
  if (dmax != std::numeric_limits<double>::infinity()) return run_rnear(tree, q, m, dmax);
  PyErr_Format(PyExc_ValueError, "%s: k = %ld and dmax = %lf is an invalid combination.", who, k,
               dmax);
  return nullptr;
}

const char kBuildDoc[] =
    "Builds k-d tree.\n\n"
    "Parameters\n----------\n"
    "points : n x k numpy array\n"
    "    Currently supports any numeric type and any k in [2, 4].\n"
    "numprocs : int, optional\n"
    "    Accepted for compatibility; the tree is built on the GPU.\n"
    "maxleafsize : int, optional\n"
    "    Points per leaf bucket (1..256). Default: tuned for the GPU.\n"
    "emptysplit : float, optional\n"
    "    Accepted for compatibility; node boxes are always tight.\n\n"
    "Returns\n-------\n"
    "tree : capsule object\n"
    "    This object is meant to be passed into the kdtree._query method.\n";

const char kQueryDoc[] =
    "Query k-d tree for nearest neighbors.\n\n"
    "Query points can be specified by one of the following ways\n"
    "(n is the number of points in the k-d tree and m the number of queries):\n\n"
    "    1. k-d points\n"
    "       (m x k numpy array)\n"
    "    2. integer indices\n"
    "       (list or numpy array of m integers; negative integers supported)\n"
    "    3. boolean selection mask on the n points used to build the k-d tree\n"
    "       (numpy array of n bools)\n\n"
    "For each query point, this function returns up to k nearest neighbors\n"
    "that have distances strictly less than dmax from the query point.\n"
    "Neighbours are ordered by distance, ties by point index.\n\n"
    "Parameters\n----------\n"
    "tree : capsule object\n"
    "    A k-d tree as produced by kdtree._build.\n"
    "queries : numpy array\n"
    "    Specify query points by one of the ways above.\n"
    "k : int, optional\n"
    "    Default: 1\n"
    "dmax : float or double, optional\n"
    "    Default: inf\n"
    "numprocs : int, optional\n"
    "    Accepted for compatibility.\n\n"
    "Returns\n-------\n"
    "results: list of m numpy arrays of integer indices\n";

PyMethodDef kMethods[] = {
    {"_build", (PyCFunction)(void (*)(void))Build, METH_VARARGS | METH_KEYWORDS, kBuildDoc},
    {"_query", (PyCFunction)(void (*)(void))Query, METH_VARARGS | METH_KEYWORDS, kQueryDoc},
    {nullptr, nullptr, 0, nullptr}};

PyModuleDef kModule = {PyModuleDef_HEAD_INIT, "kdtree", nullptr, -1, kMethods,
                       nullptr, nullptr, nullptr, nullptr};

}  // namespace

PyMODINIT_FUNC PyInit_kdtree(void) {
  import_array();
  return PyModule_Create(&kModule);
}

// NumPy / Python argument marshalling shared by the kdtree and
// estimate_normals CPython modules. Behaviour follows pptk's helpers
// (pptk/include/python_util.h): arrays of ndim 0..2 or Python scalars are
// accepted as point sets (CheckAndExtractArray2D, python_util.h:334-352);
// query / subsample indices may be a list, an integer or bool ndarray, or an
// integer scalar, are range-checked against [-n, n-1] and made non-negative
// (CheckAndExtractIndices, python_util.h:576-610).
#pragma once

#define PY_SSIZE_T_CLEAN
#include <Python.h>
#ifndef NO_IMPORT_ARRAY
#define PY_ARRAY_UNIQUE_SYMBOL PPTK_AMD_ARRAY_API
#endif
#define NPY_NO_DEPRECATED_API NPY_1_7_API_VERSION
#include <numpy/arrayobject.h>

#include <cstdint>
#include <string>
#include <vector>

#include "pptk_amd.h"

namespace pyx {

// Owned reference.
struct Ref {
  PyObject* p = nullptr;
  Ref() = default;
  explicit Ref(PyObject* o) : p(o) {}
  ~Ref() { Py_XDECREF(p); }
  Ref(const Ref&) = delete;
  Ref& operator=(const Ref&) = delete;
  PyObject* release() { PyObject* r = p; p = nullptr; return r; }
  void reset(PyObject* o) { Py_XDECREF(p); p = o; }
};

// Raises the Python exception matching a library status.
inline PyObject* raise_status(pptk_status st) {
  const char* msg = pptk_last_error();
  PyObject* exc = PyExc_RuntimeError;
  switch (st) {
    case PPTK_E_VALUE: exc = PyExc_ValueError; break;
    case PPTK_E_TYPE: exc = PyExc_TypeError; break;
    case PPTK_E_MEMORY: exc = PyExc_MemoryError; break;
    case PPTK_E_NOTIMPL: exc = PyExc_NotImplementedError; break;
    default: exc = PyExc_RuntimeError; break;
  }
  PyErr_SetString(exc, (msg && *msg) ? msg : "pptk_amd: device call failed");
  return nullptr;
}

// Maps a NumPy type number onto the library's element types.
inline bool map_dtype(int type_num, pptk_dtype* out) {
  PyArray_Descr* d = PyArray_DescrFromType(type_num);
  if (!d) { PyErr_Clear(); return false; }
  const char kind = d->kind;
  const int size = (int)PyDataType_ELSIZE(d);
  Py_DECREF(d);
  if (kind == 'f') {
    if (size == 4) { *out = PPTK_F32; return true; }
    if (size == 8) { *out = PPTK_F64; return true; }
    return false;
  }
  if (kind == 'i') {
    switch (size) {
      case 1: *out = PPTK_I8; return true;
      case 2: *out = PPTK_I16; return true;
      case 4: *out = PPTK_I32; return true;
      case 8: *out = PPTK_I64; return true;
    }
    return false;
  }
  if (kind == 'u') {
    switch (size) {
      case 1: *out = PPTK_U8; return true;
      case 2: *out = PPTK_U16; return true;
      case 4: *out = PPTK_U32; return true;
      case 8: *out = PPTK_U64; return true;
    }
    return false;
  }
  return false;
}

// A C-contiguous 2-D view (m rows x n columns) of a point set.
struct Array2D {
  Ref arr;   // contiguous array holding the data
  int64_t m = 0, n = 0;
  int type_num = -1;
  const void* data() const { return PyArray_DATA((PyArrayObject*)arr.p); }
};

// Returns true and fills `x` for an ndarray of ndim 0..2 or a Python/NumPy
// scalar. Returns false with ValueError set for ndim > 2, or false with no
// error set for objects of any other kind.
inline bool extract_array2d(PyObject* obj, Array2D& x) {
  PyObject* src = nullptr;
  Ref tmp;
  if (PyArray_Check(obj)) {
    src = obj;
  } else if (PyArray_IsScalar(obj, Generic) || PyFloat_Check(obj) || PyLong_Check(obj) ||
             PyBool_Check(obj)) {
    tmp.reset(PyArray_FROM_O(obj));
    if (!tmp.p) return false;
    src = tmp.p;
  } else {
    return false;
  }
  PyArrayObject* a = (PyArrayObject*)src;
  const int ndim = PyArray_NDIM(a);
  if (ndim > 2) {
    PyErr_Format(PyExc_ValueError,
                 "ExtractArray2DFromPyArray(): only handles ndim of 0, 1 or 2; encountered ndim=%d",
                 ndim);
    return false;
  }
  if (ndim == 0) { x.m = 1; x.n = 1; }
  else if (ndim == 1) { x.m = 1; x.n = PyArray_DIM(a, 0); }
  else { x.m = PyArray_DIM(a, 0); x.n = PyArray_DIM(a, 1); }
  x.type_num = PyArray_TYPE(a);
  x.arr.reset((PyObject*)PyArray_GETCONTIGUOUS(a));
  return x.arr.p != nullptr;
}

inline bool is_integral_type(int type_num) {
  return PyTypeNum_ISINTEGER(type_num);
}

// Index extraction. Returns true on success; false with an exception set on
// a bad index object, or false without an exception for an unusable object.
inline bool extract_indices(PyObject* obj, int64_t n, std::vector<int32_t>& out) {
  std::vector<long long> raw;
  if (PyList_Check(obj)) {
    const Py_ssize_t len = PyList_Size(obj);
    raw.reserve((size_t)len);
    for (Py_ssize_t i = 0; i < len; ++i) {
      PyObject* item = PyList_GetItem(obj, i);
      Ref idx(PyNumber_Index(item));
      if (!idx.p) {
        if (PyErr_ExceptionMatches(PyExc_TypeError)) {
          PyErr_Clear();
          PyErr_Format(PyExc_TypeError,
                       "ExtractIndicesFromPyList(): Encountered non-index %s object at the %d-th item",
                       Py_TYPE(item)->tp_name, (int)i);
        }
        return false;
      }
      const long long v = PyLong_AsLongLong(idx.p);
      if (v == -1 && PyErr_Occurred()) return false;
      raw.push_back(v);
    }
  } else if (PyArray_Check(obj)) {
    PyArrayObject* a = (PyArrayObject*)obj;
    if (PyArray_NDIM(a) > 1) {
      PyErr_Format(PyExc_ValueError,
                   "ExtractIndicesFromPyArray(): PyArray has dim = %d (expected dim = 0 or 1).",
                   PyArray_NDIM(a));
      return false;
    }
    const int tn = PyArray_TYPE(a);
    if (tn == NPY_BOOL) {
      const int64_t len = PyArray_NDIM(a) == 0 ? 1 : PyArray_DIM(a, 0);
      if (len != n) {
        PyErr_Format(PyExc_ValueError,
                     "ExtractIndicesFromPyArray(): binary indexing mask has incorrect size %d "
                     "(expected %d).",
                     (int)len, (int)n);
        return false;
      }
      Ref c((PyObject*)PyArray_GETCONTIGUOUS(a));
      if (!c.p) return false;
      const npy_bool* m = (const npy_bool*)PyArray_DATA((PyArrayObject*)c.p);
      for (int64_t i = 0; i < len; ++i)
        if (m[i]) raw.push_back(i);
    } else if (is_integral_type(tn)) {
      Ref c(PyArray_FROM_OTF(obj, NPY_LONGLONG, NPY_ARRAY_C_CONTIGUOUS | NPY_ARRAY_FORCECAST));
      if (!c.p) return false;
      const int64_t len = PyArray_SIZE((PyArrayObject*)c.p);
      const long long* v = (const long long*)PyArray_DATA((PyArrayObject*)c.p);
      raw.assign(v, v + len);
    } else {
      PyErr_Format(PyExc_ValueError,
                   "ExtractIndicesFromPyArray(): encountered unsupported type_num = %d "
                   "(expected boolean or integral type).",
                   tn);
      return false;
    }
  } else if (PyLong_Check(obj) && !PyBool_Check(obj)) {
    const long long v = PyLong_AsLongLong(obj);
    if (v == -1 && PyErr_Occurred()) return false;
    raw.push_back(v);
  } else if (PyArray_IsScalar(obj, Integer)) {
    Ref idx(PyNumber_Index(obj));
    if (!idx.p) return false;
    const long long v = PyLong_AsLongLong(idx.p);
    if (v == -1 && PyErr_Occurred()) return false;
    raw.push_back(v);
  } else if (PyFloat_Check(obj) || PyBool_Check(obj) || PyArray_IsScalar(obj, Generic)) {
    PyErr_SetString(PyExc_ValueError,
                    "CheckAndExtractIndices(): encountered non-integral scalar");
    return false;
  } else {
    return false;
  }
  out.resize(raw.size());
  for (size_t i = 0; i < raw.size(); ++i) {
    long long v = raw[i];
    if (v < -n || v >= n) {
      PyErr_Format(PyExc_RuntimeError,
                   "CheckAndExtractIndices(): %lu-th query index is outside of [%d,%d]",
                   (unsigned long)i, (int)-n, (int)(n - 1));
      return false;
    }
    if (v < 0) v += n;
    out[i] = (int32_t)v;
  }
  return true;
}

// One 1-D int32 view of `base` (kept alive by the view).
inline PyObject* int32_view(PyObject* base, const int32_t* ptr, npy_intp len) {
  PyObject* v = PyArray_New(&PyArray_Type, 1, &len, NPY_INT32, nullptr, (void*)ptr, 0,
                            NPY_ARRAY_CARRAY, nullptr);
  if (!v) return nullptr;
  Py_INCREF(base);
  if (PyArray_SetBaseObject((PyArrayObject*)v, base) < 0) {
    Py_DECREF(v);
    return nullptr;
  }
  return v;
}

// kNN rows: list of m arrays, row i = idx[i, :counts[i]] of an (m, k) array.
inline PyObject* list_from_rows(PyObject* base, const int32_t* idx, int64_t m, int64_t k,
                                const int32_t* counts) {
  Ref list(PyList_New((Py_ssize_t)m));
  if (!list.p) return nullptr;
  for (int64_t i = 0; i < m; ++i) {
    PyObject* v = int32_view(base, idx + i * k, (npy_intp)counts[i]);
    if (!v) return nullptr;
    PyList_SET_ITEM(list.p, (Py_ssize_t)i, v);
  }
  return list.release();
}

// CSR lists: row i = idx[offsets[i] : offsets[i+1]].
inline PyObject* list_from_csr(PyObject* base, const int32_t* idx, int64_t m,
                               const int64_t* offsets) {
  Ref list(PyList_New((Py_ssize_t)m));
  if (!list.p) return nullptr;
  for (int64_t i = 0; i < m; ++i) {
    PyObject* v = int32_view(base, idx + offsets[i], (npy_intp)(offsets[i + 1] - offsets[i]));
    if (!v) return nullptr;
    PyList_SET_ITEM(list.p, (Py_ssize_t)i, v);
  }
  return list.release();
}

}  // namespace pyx

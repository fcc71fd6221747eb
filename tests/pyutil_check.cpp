// Drives the CPython/NumPy marshalling of the drop-in modules
// (pptk_amd/csrc/pyext/pyutil.hpp) under AddressSanitizer and
// UndefinedBehaviorSanitizer: an embedded interpreter builds the argument
// objects (arrays of every dtype and ndim 0..3, strided and reversed views,
// scalars, lists, boolean masks, out-of-range and negative indices, huge
// integers) and the checks call the helpers directly, including the result
// lists built as views of one output array (kNN rows and CSR lists), whose
// contents and base references are verified. Test infrastructure: built and
// run by tests/test_sanitizers_cpu.py (with PYTHONMALLOC=malloc so that every
// Python allocation is visible to the sanitizer).
#include "pyext/pyutil.hpp"

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

// The helpers read the library's last error when they raise; this driver
// links no library.
extern "C" const char* pptk_last_error(void) { return "status raised by the driver"; }

static int g_failures = 0;
#define CHECK(c, ...)                                   \
  do {                                                  \
    if (!(c)) {                                         \
      std::printf("FAIL %s:%d: ", __FILE__, __LINE__);  \
      std::printf(__VA_ARGS__);                         \
      std::printf("\n");                                \
      ++g_failures;                                     \
    }                                                   \
  } while (0)

static PyObject* g_ns = nullptr;

// New reference to the value of a Python expression (numpy is `np`).
static PyObject* ev(const char* expr) {
  PyObject* r = PyRun_String(expr, Py_eval_input, g_ns, g_ns);
  if (!r) {
    PyErr_Print();
    std::printf("FAIL: could not evaluate %s\n", expr);
    std::exit(1);
  }
  return r;
}

static void check_array2d() {
  struct Case {
    const char* expr;
    bool ok;
    long long m, n;
  };
  const Case cases[] = {
      {"np.arange(12, dtype=np.float32).reshape(4, 3)", true, 4, 3},
      {"np.arange(12, dtype=np.float64).reshape(3, 4)[:, ::2]", true, 3, 2},
      {"np.arange(24, dtype=np.int16).reshape(6, 4)[::-1]", true, 6, 4},
      {"np.arange(5, dtype=np.uint8)", true, 1, 5},
      {"np.float32(2.5)", true, 1, 1},
      {"np.array(7, dtype=np.int64)", true, 1, 1},
      {"3.0", true, 1, 1},
      {"5", true, 1, 1},
      {"True", true, 1, 1},
      {"np.zeros((0, 3), np.float32)", true, 0, 3},
      {"np.zeros((2, 2, 2))", false, 0, 0},
      {"[1.0, 2.0]", false, 0, 0},
      {"'text'", false, 0, 0},
      {"None", false, 0, 0},
  };
  for (const Case& c : cases) {
    PyObject* o = ev(c.expr);
    pyx::Array2D x;
    const bool ok = pyx::extract_array2d(o, x);
    CHECK(ok == c.ok, "extract_array2d(%s) = %d", c.expr, (int)ok);
    if (ok) {
      CHECK(x.m == c.m && x.n == c.n, "extract_array2d(%s) shape %lld x %lld", c.expr,
            (long long)x.m, (long long)x.n);
      CHECK(PyArray_IS_C_CONTIGUOUS((PyArrayObject*)x.arr.p), "contiguous copy of %s", c.expr);
      // read every element of the contiguous copy
      const char* p = (const char*)x.data();
      const size_t bytes = (size_t)x.m * x.n * PyArray_ITEMSIZE((PyArrayObject*)x.arr.p);
      volatile char sink = 0;
      for (size_t i = 0; i < bytes; ++i) sink ^= p[i];
      (void)sink;
    }
    PyErr_Clear();
    Py_DECREF(o);
  }
}

static void check_dtypes() {
  const int types[] = {NPY_BOOL, NPY_BYTE, NPY_UBYTE, NPY_SHORT, NPY_USHORT, NPY_INT,
                       NPY_UINT, NPY_LONG, NPY_ULONG, NPY_LONGLONG, NPY_ULONGLONG, NPY_HALF,
                       NPY_FLOAT, NPY_DOUBLE, NPY_LONGDOUBLE, NPY_CFLOAT, NPY_OBJECT, 9999};
  for (int t : types) {
    pptk_dtype d;
    const bool ok = pyx::map_dtype(t, &d);
    if (t == NPY_FLOAT) CHECK(ok && d == PPTK_F32, "float32");
    if (t == NPY_DOUBLE) CHECK(ok && d == PPTK_F64, "float64");
    if (t == NPY_OBJECT || t == 9999) CHECK(!ok, "type %d rejected", t);
    PyErr_Clear();
  }
}

static void check_indices() {
  struct Case {
    const char* expr;
    long long n;
    bool ok;
    std::vector<int32_t> want;
  };
  const std::vector<Case> cases = {
      {"[0, 3, -1]", 5, true, {0, 3, 4}},
      {"[]", 5, true, {}},
      {"np.array([4, -5, 2], dtype=np.int8)", 5, true, {4, 0, 2}},
      {"np.array([1, 2], dtype=np.uint64)", 5, true, {1, 2}},
      {"np.arange(10, dtype=np.int64)[::3]", 10, true, {0, 3, 6, 9}},
      {"np.array([True, False, True, False])", 4, true, {0, 2}},
      {"np.array([True, False])", 4, false, {}},
      {"np.int32(-2)", 4, true, {2}},
      {"3", 4, true, {3}},
      {"[0, 5]", 5, false, {}},
      {"[-6]", 5, false, {}},
      {"[2**70]", 5, false, {}},
      {"[1.5]", 5, false, {}},
      {"[None]", 5, false, {}},
      {"np.zeros((2, 2), dtype=np.int32)", 5, false, {}},
      {"np.array([0.5, 1.0])", 5, false, {}},
      {"2.0", 5, false, {}},
      {"True", 5, false, {}},
      {"np.float64(1.0)", 5, false, {}},
      {"'abc'", 5, false, {}},
      {"2**40", 5, false, {}},
  };
  for (const Case& c : cases) {
    PyObject* o = ev(c.expr);
    std::vector<int32_t> out;
    const bool ok = pyx::extract_indices(o, (int64_t)c.n, out);
    CHECK(ok == c.ok, "extract_indices(%s) = %d", c.expr, (int)ok);
    if (ok) CHECK(out == c.want, "extract_indices(%s) values", c.expr);
    PyErr_Clear();
    Py_DECREF(o);
  }
}

static void check_result_lists() {
  // kNN rows: (m, k) int32 array with per-row counts
  const int64_t m = 257, k = 5;
  npy_intp dims[2] = {(npy_intp)m, (npy_intp)k};
  PyObject* base = PyArray_SimpleNew(2, dims, NPY_INT32);
  int32_t* idx = (int32_t*)PyArray_DATA((PyArrayObject*)base);
  std::vector<int32_t> cnt((size_t)m);
  for (int64_t i = 0; i < m; ++i) {
    cnt[(size_t)i] = (int32_t)(i % (k + 1));
    for (int64_t s = 0; s < k; ++s) idx[i * k + s] = (int32_t)(i * 100 + s);
  }
  PyObject* rows = pyx::list_from_rows(base, idx, m, k, cnt.data());
  CHECK(rows && PyList_Size(rows) == m, "list_from_rows");
  Py_DECREF(base);  // the views keep the array alive
  for (int64_t i = 0; rows && i < m; ++i) {
    PyArrayObject* v = (PyArrayObject*)PyList_GetItem(rows, (Py_ssize_t)i);
    CHECK(PyArray_DIM(v, 0) == cnt[(size_t)i], "row %lld length", (long long)i);
    const int32_t* d = (const int32_t*)PyArray_DATA(v);
    for (npy_intp s = 0; s < PyArray_DIM(v, 0); ++s)
      CHECK(d[s] == (int32_t)(i * 100 + s), "row %lld value", (long long)i);
  }
  Py_XDECREF(rows);  // frees the views and then the array

  // CSR lists with empty rows
  const int64_t mm = 100;
  std::vector<int64_t> off((size_t)mm + 1, 0);
  for (int64_t i = 0; i < mm; ++i) off[(size_t)i + 1] = off[(size_t)i] + (i % 3 == 0 ? 0 : i % 7);
  npy_intp tot = (npy_intp)off[(size_t)mm];
  PyObject* flat = PyArray_SimpleNew(1, &tot, NPY_INT32);
  int32_t* f = (int32_t*)PyArray_DATA((PyArrayObject*)flat);
  for (npy_intp j = 0; j < tot; ++j) f[j] = (int32_t)j;
  PyObject* lists = pyx::list_from_csr(flat, f, mm, off.data());
  CHECK(lists && PyList_Size(lists) == mm, "list_from_csr");
  Py_DECREF(flat);
  for (int64_t i = 0; lists && i < mm; ++i) {
    PyArrayObject* v = (PyArrayObject*)PyList_GetItem(lists, (Py_ssize_t)i);
    CHECK(PyArray_DIM(v, 0) == off[(size_t)i + 1] - off[(size_t)i], "csr row %lld", (long long)i);
    const int32_t* d = (const int32_t*)PyArray_DATA(v);
    for (npy_intp s = 0; s < PyArray_DIM(v, 0); ++s)
      CHECK(d[s] == (int32_t)(off[(size_t)i] + s), "csr value");
  }
  Py_XDECREF(lists);
}

static void check_raise() {
  const pptk_status st[] = {PPTK_E_VALUE, PPTK_E_TYPE, PPTK_E_MEMORY, PPTK_E_NOTIMPL,
                            PPTK_E_RUNTIME, PPTK_E_DEVICE};
  PyObject* want[] = {PyExc_ValueError, PyExc_TypeError, PyExc_MemoryError,
                      PyExc_NotImplementedError, PyExc_RuntimeError, PyExc_RuntimeError};
  for (int i = 0; i < 6; ++i) {
    CHECK(pyx::raise_status(st[i]) == nullptr, "raise returns null");
    CHECK(PyErr_ExceptionMatches(want[i]), "exception type for status %d", (int)st[i]);
    PyErr_Clear();
  }
}

static int init_numpy() {
  import_array1(-1);
  return 0;
}

int main() {
  Py_Initialize();
  if (init_numpy() < 0) {
    PyErr_Print();
    return 2;
  }
  g_ns = PyDict_New();
  PyDict_SetItemString(g_ns, "__builtins__", PyEval_GetBuiltins());
  PyObject* np = PyImport_ImportModule("numpy");
  if (!np) {
    PyErr_Print();
    return 2;
  }
  PyDict_SetItemString(g_ns, "np", np);
  Py_DECREF(np);
  for (int round = 0; round < 3; ++round) {
    check_array2d();
    check_dtypes();
    check_indices();
    check_result_lists();
    check_raise();
  }
  Py_DECREF(g_ns);
  if (g_failures) {
    std::printf("pyutil_check: %d failures\n", g_failures);
    return 1;
  }
  std::printf("pyutil_check: ok\n");
  return 0;
}
